"""ctypes binding of libsdesim.so (the C ABI declared in include/sdesim.h).

The shared library is built in-tree by ``sde-sim-rs_amd/build.py``. Importing this module fails
loudly when it is missing: there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsdesim.so")

OK, EVALUE, ERUNTIME, EPANIC = 0, 1, 2, 3
EULER, RUNGE_KUTTA = 0, 1
PSEUDO, SOBOL, INJECTED = 0, 1, 2
SCRAMBLE_SHIFT, SCRAMBLE_XOR = 0, 1
F64, F32 = 0, 1
COMPAT_REFERENCE, COMPAT_TEXTBOOK = 0, 1
OUT_FILTRATION, OUT_TERMINAL = 1, 2
STORAGE_AUTO, STORAGE_TAPE, STORAGE_CHECKPOINTS = 0, 1, 2
SCRATCH_BYTES = 4 << 20  # SDESIM_SCRATCH_BYTES


class Options(ctypes.Structure):
    _fields_ = [
        ("scheme", ctypes.c_int32),
        ("rng", ctypes.c_int32),
        ("scramble", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("compat", ctypes.c_int32),
        ("outputs", ctypes.c_int32),
        ("paths_per_lane", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("scramble_seed", ctypes.c_uint64),
        ("sobol_first", ctypes.c_int64),
    ]


class Instance(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("paths", ctypes.c_int64),
        ("x0", ctypes.c_double),
        ("horizon", ctypes.c_double),
        ("p0", ctypes.c_double),
        ("p1", ctypes.c_double),
        ("p2", ctypes.c_double),
        ("seed", ctypes.c_uint64),
    ]


class BasketSpec(ctypes.Structure):
    _fields_ = [
        ("assets", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("paths", ctypes.c_int64),
        ("horizon", ctypes.c_double),
        ("strike", ctypes.c_double),
        ("width", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("precision", ctypes.c_int32),
        ("storage", ctypes.c_int32),
    ]


class LmmSpec(ctypes.Structure):
    _fields_ = [
        ("forwards", ctypes.c_int32),
        ("factors", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("steps_per_period", ctypes.c_int32),
        ("expiry_index", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("paths", ctypes.c_int64),
        ("tenor", ctypes.c_double),
        ("strike", ctypes.c_double),
        ("scramble_seed", ctypes.c_uint64),
        ("storage", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
    ]


# name -> (restype, argtypes); every symbol include/sdesim.h declares.
SIGNATURES = {
    "sdesim_default_options": (None, [ctypes.POINTER(Options)]),
    "sdesim_parse_equations": (ctypes.c_int, [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int32,
                                              ctypes.POINTER(ctypes.c_void_p)]),
    "sdesim_free_universe": (None, [ctypes.c_void_p]),
    "sdesim_num_processes": (ctypes.c_int32, [ctypes.c_void_p]),
    "sdesim_num_increments": (ctypes.c_int32, [ctypes.c_void_p]),
    "sdesim_process_name": (ctypes.c_char_p, [ctypes.c_void_p, ctypes.c_int32]),
    "sdesim_process_is_algebraic": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "sdesim_process_index": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_char_p]),
    "sdesim_plan_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_double), ctypes.c_int64, ctypes.POINTER(Options),
                                          ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    "sdesim_plan_free": (None, [ctypes.c_void_p]),
    "sdesim_plan_filtration_bytes": (ctypes.c_int64, [ctypes.c_void_p]),
    "sdesim_plan_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "sdesim_plan_moments": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]),
    "sdesim_plan_code_hash": (ctypes.c_uint64, [ctypes.c_void_p]),
    "sdesim_plan_run_scratch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p]),
    "sdesim_filtration_to_long": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                                 ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]),
    "sdesim_simulate_long": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int32,
                                            ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double),
                                            ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(Options),
                                            ctypes.c_void_p, ctypes.c_void_p]),
    "sdesim_plan_source": (ctypes.c_char_p, [ctypes.c_void_p]),
    "sdesim_generate_source": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Options),
                                              ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64,
                                              ctypes.POINTER(ctypes.c_int64)]),
    "sdesim_compile_source": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
    "sdesim_eval_expression": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p),
                                              ctypes.POINTER(ctypes.c_double), ctypes.c_int32,
                                              ctypes.POINTER(ctypes.c_double)]),
    "sdesim_sobol_points": (ctypes.c_int, [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int64,
                                           ctypes.POINTER(ctypes.c_uint32)]),
    "sdesim_batch_create": (ctypes.c_int, [ctypes.POINTER(Instance), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "sdesim_batch_free": (None, [ctypes.c_void_p]),
    "sdesim_batch_set_config": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "sdesim_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                                             ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)]),
    "sdesim_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "sdesim_batch_capture": (ctypes.c_int, [ctypes.c_void_p]),
    "sdesim_batch_replay": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "sdesim_batch_moments": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]),
    "sdesim_batch_terminal": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "sdesim_batch_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_int64)]),
    "sdesim_basket_create": (ctypes.c_int, [ctypes.POINTER(BasketSpec), ctypes.POINTER(ctypes.c_void_p)]),
    "sdesim_kernel_cache_reset": (None, []),
    "sdesim_basket_free": (None, [ctypes.c_void_p]),
    "sdesim_basket_storage": (ctypes.c_int32, [ctypes.c_void_p]),
    "sdesim_basket_eval": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_double)] * 5 +
                           [ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]),
    "sdesim_lmm_create": (ctypes.c_int, [ctypes.POINTER(LmmSpec), ctypes.POINTER(ctypes.c_void_p)]),
    "sdesim_lmm_free": (None, [ctypes.c_void_p]),
    "sdesim_lmm_storage": (ctypes.c_int32, [ctypes.c_void_p]),
    "sdesim_lmm_eval": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_double)] * 3 +
                        [ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]),
    "sdesim_last_error": (ctypes.c_char_p, []),
    "sdesim_version": (ctypes.c_int32, []),
}


def _load():
    # torch and libsdesim.so must share one HIP runtime (both bind the SONAME libamdhip64.so.7);
    # loading torch first makes the choice deterministic: torch's bundled runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError("libsdesim.so is not built (%s); run `python sde-sim-rs_amd/build.py`" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class PanicException(BaseException):
    """Raised where the reference panics (PyO3 surfaces a Rust panic as
    pyo3_runtime.PanicException, a BaseException subclass)."""


def last_error():
    msg = lib.sdesim_last_error()
    return msg.decode("utf-8", "replace") if msg else ""


def check(rc, value_prefix="", runtime_prefix=""):
    if rc == OK:
        return
    msg = last_error()
    if rc == EVALUE:
        raise ValueError(value_prefix + msg)
    if rc == EPANIC:
        raise PanicException(msg)
    raise RuntimeError(runtime_prefix + msg)


def default_options():
    o = Options()
    lib.sdesim_default_options(ctypes.byref(o))
    return o


def dptr(arr):
    return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
