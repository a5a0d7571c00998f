"""Host-side product tests (CPU): the C ABI, the DSL parser, expression semantics, code
generation (hipRTC compiles for gfx950 without a device), Sobol tables and error behaviour."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import GBM, HESTON, JUMP, JUMP_X0, ROOT

import sde_sim_rs
from sde_sim_rs import PanicException
from sde_sim_rs import _native as nat


def test_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "sdesim.h")).read()
    decls = set(re.findall(r"\b(sdesim_[a-z_]+)\s*\(", hdr))
    assert len(decls) >= 20
    lib = ctypes.CDLL(nat.LIB_PATH)
    for name in sorted(decls):
        assert hasattr(lib, name), name
        assert name in nat.SIGNATURES, name
    assert set(nat.SIGNATURES) == decls


# ---------------------------------------------------------------------------------------------
# DSL (src/proc/util.rs): product parser against the oracle's
# ---------------------------------------------------------------------------------------------
UNIVERSES = [
    GBM, HESTON, JUMP,
    ["dX1 = ( sin(t) ) * dt + (0.01 * X1) * dW1 + (0.001 * X1) * dN1(0.5 * cos(t))", "X2 = max(X1 - 100.0, 0.0)"],
    ["dX = (1.0) * dt - (2.0) * dW2 + (3) * dW1", "dY = (X) * dW1"],
    ["dX = (1.0) * dt + (2.0) dW1"],                 # stops quietly at ')' not followed by '*'
    ["dX = 0.05 * X * dt"],                          # no parenthesised coefficient: no terms
    ["dA = (1) * dN1(A)", "dB = (1) * dN1( A )", "dC = (1) * dN1(A)"],  # dN keys include lambda text
    ["Y = 3", "X = Y * 2"],
    ["dX = ((1 + 2) * (3)) * dt"],
]


@pytest.mark.parametrize("eqs", UNIVERSES)
def test_parser_matches_oracle(oracle, eqs):
    u = sde_sim_rs.Universe(eqs)
    o = oracle.parse(eqs)
    assert u.num_processes == o["P"]
    assert u.num_increments == o["m"]
    assert u.names == o["names"]
    assert u.is_algebraic == o["algebraic"]


def test_parser_registry_order_and_names():
    u = sde_sim_rs.Universe(["dX = (1.0) * dt - (2.0) * dW2 + (3) * dW1", "dY = (X) * dW1"])
    assert u.names == ["X", "Y"] and u.num_increments == 2
    u = sde_sim_rs.Universe(["dA = (1) * dN1(A)", "dB = (1) * dN1( A )", "dC = (1) * dN1(A)"])
    assert u.num_increments == 2
    u = sde_sim_rs.Universe(["dX = (1) * dt", "dX = (2) * dt"])
    assert u.index("X") == 1  # the registry keeps the last index of a repeated name


BAD = [
    ("dX = (0.1 * X * dt", "Unbalanced parentheses in coefficient"),
    ("dX = (1) * dZ", "Unknown incrementor type: dZ"),
    ("X", "Missing '='"),
    ("a = b = c", "Missing '='"),
    ("dX = (1) * dN1(X", "Unbalanced parentheses in dN intensity"),
    ("dX = (1) * dN1", "dN missing opening bracket"),
    ("dX = (2 +) * dt", "Math error in coefficient"),
    ("dX=(1)*dt+(2)*dW1", "Unknown incrementor type: dt+(2)*dW1"),
    ("X = (1 + ", "Parse Error"),
    ("dX = (1) * dN1(1 +)", "Math error in jump lambda '1 +'"),
]


@pytest.mark.parametrize("eq,msg", BAD)
def test_parse_errors(oracle, eq, msg):
    with pytest.raises(ValueError) as e:
        sde_sim_rs.Universe([eq])
    assert str(e.value).startswith("Failed to parse equations: ")
    assert msg in str(e.value)
    with pytest.raises(ValueError) as e2:
        oracle.parse([eq])
    assert msg in str(e2.value)


# ---------------------------------------------------------------------------------------------
# Expression semantics (fasteval restatement): product host evaluator vs oracle evaluator
# ---------------------------------------------------------------------------------------------
def product_eval(expr, ns):
    names = list(ns)
    arr = (ctypes.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
    vals = np.ascontiguousarray([float(ns[n]) for n in names] or [0.0])
    out = ctypes.c_double()
    nat.check(nat.lib.sdesim_eval_expression(expr.encode(), arr, nat.dptr(vals), len(names), ctypes.byref(out)),
              value_prefix="")
    return out.value


EXPRS = [
    ("2 - 3 + 4", 3.0), ("2 + 3 - 4", 1.0), ("2 * 3 / 4", 1.5), ("2 / 4 * 3", 1.5), ("2 ^ 3 ^ 2", 512.0),
    ("-2 ^ 2", 4.0), ("10 % 4 * 2", 4.0), ("1 + 2 * 3", 7.0), ("(1 + 2) * 3", 9.0), ("1 < 2 == 1", 1.0),
    ("2k", 2000.0), ("3m", 0.003), ("log(100)", 2.0), ("log(2, 8)", 3.0), ("round(0.5, 1.26)", 1.5),
    ("min(3, 1, 2)", 1.0), ("max(3, x, 2)", 5.0), ("0 && 5", 0.0), ("2 && 5", 5.0), ("0 || 7", 7.0),
    ("3 || 7", 3.0), ("!0", 1.0), ("!3", 0.0), ("int(-2.7)", -2.0), ("sign(-0.5)", -1.0), ("abs(-3)", 3.0),
    ("e()", np.e), ("pi()", np.pi), ("x(1, 2)", 5.0), ("1.5e3", 1500.0), (".5 + 1.", 1.5), ("+-+1", -1.0),
    ("-7 % 3", -1.0), ("2 ^ -1", 0.5), ("floor(-1.5) + ceil(1.2)", 0.0), ("t * 2", 0.5), ("3 == 3 + 1e-16", 1.0),
]


@pytest.mark.parametrize("expr,val", EXPRS)
def test_expression_semantics(oracle, expr, val):
    ns = {"x": 5.0, "t": 0.25}
    p = product_eval(expr, ns)
    o = oracle.eval_expr(expr, ns)
    assert p == o or (np.isnan(p) and np.isnan(o))
    assert p == pytest.approx(val, rel=1e-12, abs=1e-15)


def test_expression_undefined_name_panics(oracle):
    with pytest.raises(PanicException):
        product_eval("y + 1", {"x": 1.0})
    with pytest.raises(RuntimeError):
        oracle.eval_expr("y + 1", {"x": 1.0})


def test_expression_random_agreement(oracle):
    """Random expressions over the operator set: the two independent parsers/evaluators agree
    bit for bit; on the subset without '-' and '/' (where fasteval's precedence coincides with
    Python's) they also agree with Python's evaluation."""
    rng = np.random.default_rng(11)
    ops = ["+", "-", "*", "/", "%", "^", "<", ">", "<=", ">=", "==", "!=", "&&", "||"]

    def gen(depth, allowed):
        if depth == 0 or rng.random() < 0.3:
            c = rng.integers(0, 4)
            if c == 0:
                return "a"
            if c == 1:
                return "b"
            return "%.3g" % rng.uniform(0.1, 3)
        if rng.random() < 0.15:
            return "(" + gen(depth - 1, allowed) + ")"
        if rng.random() < 0.1:
            return rng.choice(["min", "max"]) + "(" + gen(depth - 1, allowed) + ", " + gen(depth - 1, allowed) + ")"
        return gen(depth - 1, allowed) + " " + rng.choice(allowed) + " " + gen(depth - 1, allowed)

    ns = {"a": 1.7, "b": 0.6}
    for _ in range(400):
        e = gen(4, ops)
        try:
            o = oracle.eval_expr(e, ns)
        except (ValueError, RuntimeError):
            continue
        p = product_eval(e, ns)
        assert (p == o) or (np.isnan(p) and np.isnan(o)), e
    for _ in range(300):
        e = gen(4, ["+", "*"])
        p = product_eval(e, ns)
        py = eval(e, {"min": min, "max": max}, dict(ns))
        assert p == pytest.approx(py, rel=1e-12), e


# ---------------------------------------------------------------------------------------------
# Sobol tables the kernels read, against scipy's generator (golden fixtures)
# ---------------------------------------------------------------------------------------------
def product_sobol(dims, n, first=0):
    out = np.zeros((n, dims), dtype=np.uint32)
    nat.check(nat.lib.sdesim_sobol_points(dims, first, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
    return out


def test_device_sobol_tables_match_scipy():
    g = np.load(os.path.join(ROOT, "tests", "golden", "sobol_scipy.npz"))
    assert np.array_equal(product_sobol(2048, 300)[:, g["dims2048_cols"]], g["dims2048_first300"])
    assert np.array_equal(product_sobol(2048, 64, 1000)[:, g["ff1000_cols"]], g["dims2048_from1000"])
    assert np.array_equal(product_sobol(21201, 64)[:, g["dims21201_cols"]], g["dims21201_first64"])


def test_device_sobol_tables_far_indices(oracle):
    # indices that need 3, 4 and 5 table slices
    for first in (1 << 20, (1 << 27) + 12345, (1 << 31) + 777):
        assert np.array_equal(product_sobol(16, 32, first), oracle.sobol_points(16, 32, first=first))


# ---------------------------------------------------------------------------------------------
# Code generation: every variant compiles for gfx950 (hipRTC, no device needed)
# ---------------------------------------------------------------------------------------------
def generate(eqs, n_times, payoff=None, **opts):
    u = sde_sim_rs.Universe(eqs)
    o = nat.default_options()
    for k, v in opts.items():
        setattr(o, k, v)
    need = ctypes.c_int64()
    pb = payoff.encode() if payoff else None
    nat.check(nat.lib.sdesim_generate_source(u._h, n_times, ctypes.byref(o), pb, None, 0, ctypes.byref(need)))
    buf = ctypes.create_string_buffer(need.value)
    nat.check(nat.lib.sdesim_generate_source(u._h, n_times, ctypes.byref(o), pb, buf, need.value, ctypes.byref(need)))
    return buf.value.decode()


@pytest.mark.parametrize("eqs", [GBM, HESTON, JUMP], ids=["gbm", "heston", "jump"])
@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("rng", [(0, 0), (1, 0), (1, 1)], ids=["pseudo", "sobolshift", "sobolxor"])
def test_generated_kernels_compile(eqs, scheme, precision, rng):
    src = generate(eqs, 33, payoff=None, scheme=scheme, precision=precision, rng=rng[0], scramble=rng[1],
                   outputs=3 if precision else 1)
    size = ctypes.c_int64()
    rc = nat.lib.sdesim_compile_source(src.encode(), ctypes.byref(size))
    assert rc == 0, nat.last_error()[:2000]
    assert size.value > 1000


def test_codegen_reference_rk_carries_probe_coefficients():
    # RK with reference compat: the stage-1 coefficients of step n are the stage-2 values of step
    # n-1 (same expressions on the same probe cache at the same time), carried in registers.
    # FP64: three coefficients hold a square root (sqrt(v) S, -0.35 sqrt(v), 0.357 sqrt(v)); with
    # the textbook cache both stages evaluate them. FP32 factors sqrt(v) out of the variance's two
    # Wiener coefficients, so two expressions hold one.
    for precision, once in ((0, 3), (1, 2)):
        src = generate(HESTON, 33, scheme=1, precision=precision, rng=1, scramble=1)
        body = src[src.index("void step"):]
        assert "K[" in body and body.count("sde_sqrt(") == once
        src_tb = generate(HESTON, 33, scheme=1, precision=precision, rng=1, scramble=1, compat=1)
        assert src_tb[src_tb.index("void step"):].count("sde_sqrt(") == 2 * once


def test_codegen_fp32_factors_constants_out_of_coefficients():
    # FP32: c * r dt terms become r * (c * dt); the variance's Wiener terms -0.35 sqrt(v) dW1 and
    # 0.357 sqrt(v) dW2 become sqrt(v) * g with g = -0.35 dW1 + dW2', where dW2' = 0.357 sqrt(dt)
    # Psi(u_1) (dW2 has no other use and 0.357 > 0); the RK probe term is
    # sqrt(v) * ((-0.35 + 0.357) * sk sqrt(dt)). FP64 keeps every coefficient as written.
    def hx(v):  # C's %a form of a double
        mant, exp = float.hex(v).split("p")
        return mant.rstrip("0").rstrip(".") + "p" + exp

    for scheme in (0, 1):
        src = generate(HESTON, 33, scheme=scheme, precision=1, rng=1, scramble=1)
        step = src[src.index("void step"):src.index("payoff")]
        c1, c2 = -0.7 * 0.5, 0.7141428428542850 * 0.5
        assert "w1 = sde_dw(u[1], (((float)%s) * sdt));" % hx(c2) in step
        assert "g0 = ((float)%s) * w0 + w1;" % hx(c1) in step
        assert "* (((float)%s) * dt)" % hx(0.02) in step  # 0.02 * S
        assert "* (((float)%s) * dt)" % hx(2.0) in step  # 2 * (0.04 - max(v, 0))
        assert "w0 = sde_dw(u[0], sdt);" in step  # dW1 is also used by dS, unscaled
        if scheme == 1:
            assert "* (((float)%s) * sks)" % hx(c1 + c2) in step
        f64 = generate(HESTON, 33, scheme=scheme, precision=0, rng=1, scramble=1)
        assert "g0" not in f64 and "* dt)" not in f64 and "sde_dw(u[1], sdt)" in f64
    # A constant < 0 is not folded into the multiplier (the XOR form takes the sign from u).
    neg = ["dX = (0.1 * X) * dt + (0.2 * X) * dW1 + (-0.3 * X) * dW2"]
    src = generate(neg, 9, scheme=0, precision=1, rng=1, scramble=1)
    assert "w0 = sde_dw(u[0], (((float)%s) * sdt));" % hx(0.2) in src and "w1 = sde_dw(u[1], sdt);" in src
    assert "g0 = w0 + ((float)%s) * w1;" % hx(-0.3) in src


def test_codegen_fp32_reassociation_is_confined_to_the_step():
    # FP32 steps may be reassociated by the compiler, but the state update goes through a strict
    # addition (so results do not depend on how far the step loop is unrolled); FP64 steps keep
    # the reference's evaluation order.
    for scheme in (0, 1):
        f32 = generate(HESTON, 33, scheme=scheme, precision=1, rng=1, scramble=1)
        step = f32[f32.index("void step"):]
        assert "#pragma clang fp reassociate(on)" in step
        assert step.count("sde_add_strict(X[") == 2
        f64 = generate(HESTON, 33, scheme=scheme, precision=0, rng=1, scramble=1)
        assert "reassociate" not in f64 and "sde_add_strict" not in f64


def test_codegen_wave_aligned_table_form():
    # XOR mode with a first point index that is a multiple of 64 uses the wave-aligned table form
    # (last template argument); other first indices and the shift mode use per-lane gathers.
    def aligned(**kw):
        src = generate(HESTON, 33, scheme=1, precision=1, rng=1, **kw)
        call = re.search(r"sde_path_kernel<[^>]*>", src).group(0)
        return call.endswith("true>")
    assert aligned(scramble=1)
    assert aligned(scramble=1, sobol_first=128)
    assert not aligned(scramble=1, sobol_first=37)
    assert not aligned(scramble=0, sobol_first=0)


def test_codegen_payoff_and_names():
    src = generate(HESTON, 33, payoff="max(S - 100, 0)", scheme=1, precision=1, rng=1, scramble=1, outputs=2)
    assert "sdesim_rk_f32_sobolxor_term" in src and "X[0]" in src[src.index("payoff("):]


# ---------------------------------------------------------------------------------------------
# Error behaviour of the Python API (src/py_binding.rs:19-47), raised before any device work
# ---------------------------------------------------------------------------------------------
def test_scenarios_must_be_positive():
    with pytest.raises(ValueError, match="scenarios must be a positive integer"):
        sde_sim_rs.simulate(GBM, [0.0, 1.0], 0, {"X": 1.0}, "pseudo", "euler")
    with pytest.raises(TypeError):
        sde_sim_rs.simulate(GBM, [0.0, 1.0], 1.5, {"X": 1.0}, "pseudo", "euler")
    with pytest.raises(OverflowError):
        sde_sim_rs.simulate(GBM, [0.0, 1.0], 2 ** 31, {"X": 1.0}, "pseudo", "euler")
    with pytest.raises(TypeError):
        sde_sim_rs.simulate("dX = (1) * dt", [0.0, 1.0], 1, {"X": 1.0}, "pseudo", "euler")


def test_parse_error_through_simulate():
    with pytest.raises(ValueError, match="^Failed to parse equations: Unknown incrementor type: dQ"):
        sde_sim_rs.simulate(["dX = (1) * dQ"], [0.0, 1.0], 1, {}, "pseudo", "euler")


def test_panics_where_the_reference_panics():
    with pytest.raises(PanicException, match="not implemented"):
        sde_sim_rs.simulate(GBM, [0.0, 1.0], 1, {"X": 1.0}, "pseudo", "milstein")
    with pytest.raises(PanicException, match="Undefined|undefined"):
        sde_sim_rs.simulate(["dX = (q) * dt"], [0.0, 1.0], 1, {}, "pseudo", "euler")
    with pytest.raises(PanicException, match="RNG Index 0 out of bounds"):
        sde_sim_rs.simulate(["dX = (1) * dt"], [0.0, 1.0], 1, {}, "pseudo", "runge-kutta")
    with pytest.raises(PanicException):
        sde_sim_rs.simulate(GBM, [], 1, {"X": 1.0}, "pseudo", "euler")
    with pytest.raises(PanicException, match="Sobol dimension"):
        sde_sim_rs.simulate(HESTON, np.linspace(0, 1, 12000), 1, {}, "sobol", "euler")


def test_duplicate_time_points_rejected():
    with pytest.raises(ValueError, match="duplicate"):
        sde_sim_rs.simulate(GBM, [0.0, 0.5, 0.5, 1.0], 1, {"X": 1.0}, "pseudo", "euler")


def test_empty_universe_returns_empty_table():
    t = sde_sim_rs.simulate([], [0.0, 1.0], 3, {}, "pseudo", "euler")
    tab = t if hasattr(t, "column") else t.to_arrow()
    assert tab.num_rows == 0 and tab.column_names == ["scenario", "time", "process_name", "value"]


def test_fp32_q_from_signed_word():
    """The f32 device forms q = min(u, 1 - u) of u = (y + 0.5) 2^-32 from yc = ~y without folding
    the word: q = |fma((float)(int)yc, 2^-32, 2^-33)|, and a lane is in the |z| > 5 tail when
    q < 1233 2^-32. Checked here against the definition qi = min(y, ~y), q = (qi + 0.5) 2^-32, with
    numpy's round-to-nearest f32 conversion and a single rounding of the exact f64 sum (the fma):
    the tail test must agree with qi < 1233 exactly, q must be exact for qi < 2^23, and within one
    f32 rounding of the conversion otherwise."""
    rng = np.random.default_rng(5)
    edge = np.arange(0, 4096, dtype=np.uint64)
    yc = np.concatenate([edge, 0xFFFFFFFF - edge, 0x7FFFFFFF - edge, 0x80000000 + edge,
                         (1 << 23) + np.arange(-64, 64, dtype=np.int64).astype(np.uint64),
                         0xFFFFFFFF - (1 << 23) + np.arange(-64, 64, dtype=np.int64).astype(np.uint64),
                         rng.integers(0, 1 << 32, 200000, dtype=np.uint64)]).astype(np.uint32)
    f = yc.view(np.int32).astype(np.float32)
    q = np.abs((f.astype(np.float64) * 2.0**-32 + 2.0**-33).astype(np.float32))
    y = ~yc
    qi = np.minimum(y, ~y).astype(np.float64)
    exact = (qi + 0.5) * 2.0**-32
    in_tail = q < np.float32(1233.0 * 2.0**-32)
    np.testing.assert_array_equal(in_tail, qi < 1233)
    small = qi < (1 << 23)
    np.testing.assert_array_equal(q[small].astype(np.float64), exact[small])
    assert np.all(np.abs(q.astype(np.float64) - exact) <= 2.0**-23 * exact)


def test_fp32_q_from_chacha_words():
    """The f32 device forms q = min(u, 1 - u) of rand's f64 uniform u = (w >> 11) 2^-53 of the
    64-bit word w = hi:lo straight from the words: with m the sign extension of bit 31 of hi,
    q' = (f32)(hi ^ m) 2^-32 + (f32)(lo ^ m) 2^-64. Checked against the exact q: within 2^-22
    relative wherever q' takes the f32 path (q' >= 2.87e-7), and m set exactly when u >= 1/2 (the
    sign of Psi)."""
    rng = np.random.default_rng(9)
    n = 400000
    hi = rng.integers(0, 1 << 32, n, dtype=np.uint64)
    lo = rng.integers(0, 1 << 32, n, dtype=np.uint64)
    # Near 0, 1/2 and 1, and around the tail threshold, on both halves.
    t = int(2.87e-7 * 2.0**32)
    hi[:4096] = np.arange(4096) % 64
    hi[4096:8192] = (1 << 32) - 1 - np.arange(4096) % 64
    hi[8192:12288] = (1 << 31) + np.arange(-2048, 2048)
    hi[12288:14336] = t + np.arange(-1024, 1024)
    hi[14336:16384] = (1 << 32) - 1 - t + np.arange(-1024, 1024)
    w = (hi << np.uint64(32)) | lo
    u = (w >> np.uint64(11)).astype(np.float64) * 2.0**-53
    q = np.minimum(u, 1.0 - u)
    m = np.where(hi >= (1 << 31), np.uint64(0xFFFFFFFF), np.uint64(0))
    fh = (hi ^ m).astype(np.uint32).astype(np.float32).astype(np.float64)
    fl = ((lo ^ m).astype(np.uint32).astype(np.float32).astype(np.float64)) * 2.0**-64
    qp = (fh * 2.0**-32 + fl).astype(np.float32).astype(np.float64)
    np.testing.assert_array_equal(m != 0, u >= 0.5)
    fast = qp >= np.float32(2.87e-7)
    assert np.all(np.abs(qp[fast] - q[fast]) <= 2.0**-22 * q[fast])
    # Every draw the f32 path would mis-place around the threshold lies within that tolerance of it.
    assert np.all(q[~fast] < 2.87e-7 * (1 + 2.0**-22))


@pytest.mark.parametrize("chunk_rows", [1 << 24, 7, 12, 1])
def test_long_table_layout_and_chunks(chunk_rows):
    """The long table's columns, dtypes and row order (scenario -> time -> process), the same
    whatever the chunking of the columns (whole scenarios per chunk)."""
    from sde_sim_rs.api import long_table
    names, times, S = ["X0", "X1", "X2"], [0.0, 0.25, 1.0], 4
    values = np.arange(S * len(times) * len(names), dtype=np.float64).reshape(S, len(times), len(names)) / 7.0
    tab = long_table(names, times, S, values, chunk_rows=chunk_rows)
    tab = tab if hasattr(tab, "column") else tab.to_arrow()
    assert tab.column_names == ["scenario", "time", "process_name", "value"]
    assert [str(f.type) for f in tab.schema] == ["int32", "double", "string", "double"]
    assert tab.num_rows == S * len(times) * len(names)
    sc, tm, pn, vv = (tab.column(c).to_pylist() for c in tab.column_names)
    k = 0
    for s in range(S):
        for ti, t in enumerate(times):
            for pi, n in enumerate(names):
                assert (sc[k], tm[k], pn[k], vv[k]) == (s, t, n, values[s, ti, pi])
                k += 1


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and says why, before touching
    a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_abi_argument_validation():
    """Invalid arguments are refused with SDESIM_EVALUE and a message, before any device call
    (so these run without a GPU): plan shapes and launch options, basket and forward-rate specs."""
    u = sde_sim_rs.Universe(GBM)
    times = np.ascontiguousarray([0.0, 0.5, 1.0])
    x0 = np.ascontiguousarray([1.0])
    plan = ctypes.c_void_p()

    def plan_rc(S=8, **kw):
        o = nat.default_options()
        for k, v in kw.items():
            setattr(o, k, v)
        return nat.lib.sdesim_plan_create(u._h, nat.dptr(times), 3, nat.dptr(x0), S, ctypes.byref(o), None,
                                          ctypes.byref(plan))

    cases = [(dict(S=0), "scenarios must be a positive integer"),
             (dict(block_size=96), "block_size must be a multiple of 64"),
             (dict(block_size=2048), "block_size must be a multiple of 64"),
             (dict(paths_per_lane=9), "paths_per_lane must be at most 8")]
    for kw, msg in cases:
        assert plan_rc(**kw) == nat.EVALUE, kw
        assert msg in nat.last_error(), (kw, nat.last_error())

    def basket_rc(**kw):
        s = nat.BasketSpec()
        s.assets, s.steps, s.paths, s.horizon, s.strike, s.width, s.precision = 8, 10, 1024, 1.0, 100.0, 0.5, nat.F32
        for k, v in kw.items():
            setattr(s, k, v)
        h = ctypes.c_void_p()
        return nat.lib.sdesim_basket_create(ctypes.byref(s), ctypes.byref(h))

    for kw, msg in [(dict(assets=9), "assets must be between 1 and 8"), (dict(steps=0), "steps must be >= 1"),
                    (dict(paths=0), "paths must be >= 1"), (dict(horizon=0.0), "horizon must be > 0"),
                    (dict(width=-1.0), "width must be >= 0"), (dict(storage=3), "storage must be")]:
        assert basket_rc(**kw) == nat.EVALUE, kw
        assert msg in nat.last_error(), (kw, nat.last_error())

    def lmm_rc(**kw):
        s = nat.LmmSpec()
        s.forwards, s.factors, s.steps, s.steps_per_period, s.expiry_index = 40, 3, 160, 40, 4
        s.precision, s.paths, s.tenor, s.strike = nat.F32, 1024, 0.25, 0.03
        for k, v in kw.items():
            setattr(s, k, v)
        h = ctypes.c_void_p()
        return nat.lib.sdesim_lmm_create(ctypes.byref(s), ctypes.byref(h))

    for kw, msg in [(dict(forwards=20), "compiled for 40 forwards"), (dict(expiry_index=40), "expiry_index out of range"),
                    (dict(steps=150), "steps must equal steps_per_period * expiry_index"),
                    (dict(paths=(1 << 22) + 1), "paths must be in [1, 2^22]"), (dict(tenor=0.0), "tenor must be > 0"),
                    (dict(storage=-1), "storage must be"),
                    (dict(steps=7100, steps_per_period=1775, expiry_index=4), "Sobol dimension (steps x factors)")]:
        assert lmm_rc(**kw) == nat.EVALUE, kw
        assert msg in nat.last_error(), (kw, nat.last_error())

    # Batches: instance parameters and the work-item count (paths / 64 summed over the batch must
    # fit 32 bits).
    def batch_rc(insts):
        arr = (nat.Instance * len(insts))(*insts)
        h = ctypes.c_void_p()
        return nat.lib.sdesim_batch_create(arr, len(insts), 1, 0, ctypes.byref(h))

    ok = dict(kind=0, steps=10, paths=1000, horizon=1.0)
    for bad, msg in [(dict(kind=3), "unknown model"), (dict(steps=0), "steps must be >= 1"),
                     (dict(paths=0), "paths must be >= 1"), (dict(horizon=0.0), "horizon must be > 0")]:
        assert batch_rc([nat.Instance(**dict(ok, **bad))]) == nat.EVALUE, bad
        assert msg in nat.last_error(), (bad, nat.last_error())
    huge = [nat.Instance(**dict(ok, paths=1 << 36)) for _ in range(3)]  # 3 * 2^30 items at 64 paths each
    assert batch_rc(huge) == nat.EVALUE and "too many work items" in nat.last_error()

    # The one-call long simulation checks its pointers before planning.
    vals = np.zeros(3 * 8)
    names = (ctypes.c_char_p * 1)(b"X")
    inits = (ctypes.c_double * 1)(1.0)
    sim = nat.lib.sdesim_simulate_long
    assert sim(u._h, nat.dptr(times), 3, names, inits, 1, 8, None, None, None) == nat.EVALUE
    assert "null output buffer" in nat.last_error()
    assert sim(u._h, nat.dptr(times), 3, None, None, 1, 8, None, nat.dptr(vals), None) == nat.EVALUE
    assert "bad initial values" in nat.last_error()
    assert sim(u._h, nat.dptr(times), 3, (ctypes.c_char_p * 1)(None), inits, 1, 8, None, nat.dptr(vals),
               None) == nat.EVALUE
    assert "null initial value name" in nat.last_error()

    # The long-layout transpose: shapes are checked before any device call (null buffers here).
    to_long = nat.lib.sdesim_filtration_to_long
    for args, msg in [((1, 0, 4, 2), "bad shape"), ((7, 8, 4, 2), "precision must be f32 or f64"),
                      ((1, 8, 1 << 16, 1 << 16), "exceeds 2^31 - 1"),
                      ((1, 1 << 40, 1025, 2), "too large for one long-layout transpose"),
                      ((0, 1 << 39, 1025, 2), "too large for one long-layout transpose"),
                      ((1, 8, 4, 2), "null argument")]:
        precision, S, T, P = args
        assert to_long(None, precision, S, T, P, None, None) == nat.EVALUE, args
        assert msg in nat.last_error(), (args, nat.last_error())


def test_plan_refuses_initial_values_of_the_wrong_length():
    """A sequence of initial values must hold one value per process (registry order): the library
    copies exactly that many, so a shorter or misshaped array is refused before any device call."""
    u = sde_sim_rs.Universe(HESTON)
    times = np.linspace(0.0, 1.0, 5)
    for bad in ([100.0], [100.0, 0.04, 1.0], [[100.0, 0.04]], np.zeros((2, 1))):
        with pytest.raises(ValueError, match="one per process"):
            sde_sim_rs.Plan(u, times, 8, bad, "sobol", "runge-kutta")


def test_bench_quotes_counters_only_for_the_kernel_they_measured():
    """bench.py takes traffic and VALU figures from the committed summaries only when the summary
    records the running kernel's code hash; otherwise it reports none, with a reason."""
    import importlib
    import sys
    sys.path.insert(0, ROOT)
    bench = importlib.import_module("bench")
    for name, keys in ((bench.PROFILE_FILT, ["traffic_bytes_per_launch"]),
                       (bench.PROFILE_TERM, ["valu_instructions_per_path_step", "trans_instructions_per_path_step"])):
        recorded = bench._profile_lines(name).get("kernel_code_hash")
        assert recorded and re.fullmatch(r"[0-9a-f]{16}", recorded), name
        vals, why = bench._profiled(name, recorded, keys)
        assert why is None and all(vals[k] > 0 for k in keys)
        vals, why = bench._profiled(name, "0" * 16, keys)
        assert vals is None and recorded in why and "0" * 16 in why
    assert bench._profiled("no_such_summary.md", "0" * 16, ["x"])[0] is None


def test_log_lines_only_when_asked():
    """SDESIM_LOG=1 writes one host-side line per event to stderr (here: a hipRTC compile); without
    it the library prints nothing."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import ctypes; from sde_sim_rs import _native as nat; "
            "n = ctypes.c_int64(); rc = nat.lib.sdesim_compile_source("
            "b'extern \"C\" __global__ void k(float* x) { x[0] = 1.0f; }', ctypes.byref(n)); "
            "print(rc, n.value > 0)" % os.path.join(ROOT, "sde-sim-rs_amd"))
    for value, expect in (("1", True), ("0", False), (None, False)):
        env = dict(os.environ)
        env.pop("SDESIM_LOG", None)
        if value is not None:
            env["SDESIM_LOG"] = value
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
        assert r.stdout.strip() == "0 True", (r.stdout, r.stderr[-2000:])
        assert ("[sdesim] source compiled with hipRTC" in r.stderr) == expect, r.stderr[-2000:]
