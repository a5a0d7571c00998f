"""Python handle on the oracle (test infrastructure only).

liboracle.so is the CPU restatement of the reference simulator (sdesim_oracle.cpp). Only
tests/, the smoke check and the benchmark's CPU-baseline leg use it, to compare against. The Sobol
direction-number table is loaded from the copy scipy ships (Joe & Kuo, new-joe-kuo-6.21201).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("oracle/liboracle.so is not built; run `python oracle/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        D, I32, I64, U64 = ctypes.c_double, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64
        P = ctypes.POINTER
        L.oracle_chacha_block.argtypes = [P(ctypes.c_uint32), I32, P(ctypes.c_uint32)]
        L.oracle_seed_from_u64.argtypes = [U64, P(ctypes.c_uint32)]
        L.oracle_pcg32.argtypes = [U64, U64, I32, P(ctypes.c_uint32)]
        L.oracle_pseudo_f64.argtypes = [U64, U64, I64, P(D)]
        L.oracle_sobol_init.argtypes = [I32, P(ctypes.c_int64), P(ctypes.c_int64)]
        L.oracle_sobol_dims.restype = I32
        L.oracle_sobol_directions.argtypes = [I32, P(ctypes.c_uint32)]
        L.oracle_sobol_points.argtypes = [I32, U64, I64, P(ctypes.c_uint32)]
        L.oracle_xor_masks.argtypes = [U64, I32, P(ctypes.c_uint32)]
        L.oracle_inv_normal.argtypes = [D]
        L.oracle_inv_normal.restype = D
        L.oracle_inv_poisson.argtypes = [D, D]
        L.oracle_inv_poisson.restype = U64
        L.oracle_eval_expr.argtypes = [ctypes.c_char_p, P(ctypes.c_char_p), P(D), I32, P(D), ctypes.c_char_p, I32]
        L.oracle_parse.argtypes = [P(ctypes.c_char_p), I32, P(I32), P(I32), ctypes.c_char_p, I32, P(I32),
                                   ctypes.c_char_p, I32]
        L.oracle_simulate.argtypes = [P(ctypes.c_char_p), I32, P(D), I32, P(ctypes.c_char_p), P(D), I32, I64, I32, I32,
                                      U64, I32, U64, I64, P(D), I32, I32, P(D), ctypes.c_char_p, I32]
        _lib = L
        _init_sobol()
    return _lib


def sobol_table():
    import scipy
    root = os.path.dirname(scipy.__file__)
    for rel in ("stats/_sobol_direction_numbers.npz", "stats/_qmc/_sobol_direction_numbers.npz"):
        path = os.path.join(root, rel)
        if os.path.exists(path):
            d = np.load(path)
            return d["poly"].astype(np.int64), d["vinit"].astype(np.int64)
    raise RuntimeError("scipy Sobol table not found")


def _init_sobol():
    poly, vinit = sobol_table()
    poly = np.ascontiguousarray(poly)
    vinit = np.ascontiguousarray(vinit)
    _lib.oracle_sobol_init(len(poly), poly.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                           vinit.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))


def _u32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def chacha_block(state16, rounds):
    inp = np.ascontiguousarray(state16, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint32)
    lib().oracle_chacha_block(_u32p(inp), rounds, _u32p(out))
    return out


def seed_from_u64(seed):
    key = np.zeros(8, dtype=np.uint32)
    lib().oracle_seed_from_u64(seed, _u32p(key))
    return key


def pcg32(state, inc, n):
    """n outputs of the seed expansion's PCG32 step (advance, then XSH-RR of the new state)."""
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_pcg32(state & (2 ** 64 - 1), inc & (2 ** 64 - 1), n, _u32p(out))
    return out


def pseudo_f64(seed, n, first=0):
    out = np.zeros(n, dtype=np.float64)
    lib().oracle_pseudo_f64(seed, first, n, _dp(out))
    return out


def sobol_points(dims, n, first=0):
    out = np.zeros((n, dims), dtype=np.uint32)
    lib().oracle_sobol_points(dims, first, n, _u32p(out))
    return out


def sobol_directions(d):
    out = np.zeros(32, dtype=np.uint32)
    lib().oracle_sobol_directions(d, _u32p(out))
    return out


def xor_masks(seed, dims):
    out = np.zeros(dims, dtype=np.uint32)
    lib().oracle_xor_masks(seed, dims, _u32p(out))
    return out


def inv_normal(p):
    return lib().oracle_inv_normal(float(p))


def inv_poisson(u, lam):
    return int(lib().oracle_inv_poisson(float(u), float(lam)))


def eval_expr(expr, namespace):
    names = list(namespace.keys())
    arr = (ctypes.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
    vals = np.ascontiguousarray([float(namespace[n]) for n in names] or [0.0])
    out = ctypes.c_double()
    err = ctypes.create_string_buffer(512)
    rc = lib().oracle_eval_expr(expr.encode(), arr, _dp(vals), len(names), ctypes.byref(out), err, 512)
    if rc == 1:
        raise ValueError(err.value.decode())
    if rc == 2:
        raise RuntimeError(err.value.decode())
    return out.value


def parse(equations):
    arr = (ctypes.c_char_p * max(1, len(equations)))(*[e.encode() for e in equations])
    P, m = ctypes.c_int(), ctypes.c_int()
    names = ctypes.create_string_buffer(1 << 16)
    alg = np.zeros(max(1, len(equations)), dtype=np.int32)
    err = ctypes.create_string_buffer(1024)
    rc = lib().oracle_parse(arr, len(equations), ctypes.byref(P), ctypes.byref(m), names, 1 << 16,
                            alg.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), err, 1024)
    if rc != 0:
        raise ValueError(err.value.decode())
    nm = names.value.decode().split("\n") if P.value else []
    return {"P": P.value, "m": m.value, "names": nm, "algebraic": [bool(a) for a in alg[:P.value]]}


def simulate(equations, times, scenarios, initial_values, scheme="euler", rng="pseudo", seed=0, scramble="shift",
             scramble_seed=0, sobol_first=None, injected=None, compat="reference", threads=0):
    """Returns values [S][T][P] (f64), the reference's raw filtrations stacked in scenario order."""
    info = parse(equations)
    P, m = info["P"], info["m"]
    T = len(times)
    tarr = np.ascontiguousarray(times, dtype=np.float64)
    names = list(initial_values.keys())
    narr = (ctypes.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
    varr = np.ascontiguousarray([float(initial_values[n]) for n in names] or [0.0])
    out = np.zeros((scenarios, T, P), dtype=np.float64)
    rng_kind = {"pseudo": 0, "sobol": 1, "injected": 2}[rng]
    scr = {"shift": 0, "xor": 1}[scramble]
    first = (5 if scr == 0 else 0) if sobol_first is None else sobol_first
    inj = None
    if injected is not None:
        # oracle layout: [S][N*m]
        inj = np.ascontiguousarray(injected, dtype=np.float64)
        assert inj.shape == (scenarios, (T - 1) * m)
    err = ctypes.create_string_buffer(1024)
    rc = lib().oracle_simulate(arr_eq(equations), len(equations), _dp(tarr), T, narr, _dp(varr), len(names),
                               scenarios, 0 if scheme == "euler" else 1, rng_kind, seed & (2 ** 64 - 1), scr,
                               scramble_seed & (2 ** 64 - 1), first, _dp(inj) if inj is not None else None,
                               {"reference": 0, "textbook": 1}[compat], threads, _dp(out), err, 1024)
    if rc == 1:
        raise ValueError(err.value.decode())
    if rc == 2:
        raise RuntimeError(err.value.decode())
    return out


def arr_eq(equations):
    return (ctypes.c_char_p * max(1, len(equations)))(*[e.encode() for e in equations])
