"""Host side of the simulator: argument handling, plans and result assembly.

Mirrors the reference's PyO3 entry point (src/py_binding.rs:8-56) and sim::simulate
(src/sim/mod.rs:20-92). Parsing, kernel generation and the simulation itself run in
libsdesim.so; this module converts arguments, allocates result buffers and builds the long
DataFrame layout of src/filtration.rs:81-116.
"""
import ctypes
import os
import numbers
from collections.abc import Mapping

import numpy as np

from . import _native as nat

_PRECISIONS = {"f64": nat.F64, "float64": nat.F64, "f32": nat.F32, "float32": nat.F32}
_SCRAMBLES = {"shift": nat.SCRAMBLE_SHIFT, "xor": nat.SCRAMBLE_XOR}
_COMPAT = {"reference": nat.COMPAT_REFERENCE, "textbook": nat.COMPAT_TEXTBOOK}


def _random_seed():
    # The reference draws a fresh u64 from the OS-seeded thread RNG on every call
    # (src/sim/mod.rs:28-29).
    return int.from_bytes(os.urandom(8), "little")


def _as_str_list(equations):
    if isinstance(equations, (str, bytes)):
        raise TypeError("Can't extract `str` to `Vec`")
    out = []
    for e in equations:
        if not isinstance(e, str):
            raise TypeError("'%s' object cannot be converted to 'PyString'" % type(e).__name__)
        out.append(e)
    return out


def _as_i32(v, name):
    if isinstance(v, bool) or not isinstance(v, numbers.Integral):
        raise TypeError("'%s' object cannot be interpreted as an integer" % type(v).__name__)
    v = int(v)
    if not -(2 ** 31) <= v < 2 ** 31:
        raise OverflowError("%s out of range for i32" % name)
    return v


class Universe:
    """A parsed set of process equations (proc::util::parse_equations, src/proc/util.rs:52-66)."""

    def __init__(self, processes_equations):
        eqs = _as_str_list(processes_equations)
        arr = (ctypes.c_char_p * max(1, len(eqs)))(*[e.encode("utf-8") for e in eqs])
        h = ctypes.c_void_p()
        rc = nat.lib.sdesim_parse_equations(arr, len(eqs), ctypes.byref(h))
        nat.check(rc, value_prefix="Failed to parse equations: ")
        self._h = h
        self.equations = eqs
        self.num_processes = nat.lib.sdesim_num_processes(h)
        self.num_increments = nat.lib.sdesim_num_increments(h)
        self.names = [nat.lib.sdesim_process_name(h, p).decode("utf-8") for p in range(self.num_processes)]
        self.is_algebraic = [bool(nat.lib.sdesim_process_is_algebraic(h, p)) for p in range(self.num_processes)]

    def index(self, name):
        """Registry index of a process name (the last equation with that name), or -1."""
        return nat.lib.sdesim_process_index(self._h, name.encode("utf-8"))

    def initial_row(self, initial_values):
        """Initial values by name -> one value per process, as ScenarioFiltration::new does
        (src/filtration.rs:42-50): unknown names are ignored, missing processes start at 0."""
        x0 = np.zeros(max(1, self.num_processes), dtype=np.float64)
        for name, val in dict(initial_values).items():
            if not isinstance(name, str):
                raise TypeError("'%s' object cannot be converted to 'PyString'" % type(name).__name__)
            idx = self.index(name)
            if idx >= 0:
                x0[idx] = float(val)
        return x0

    def __del__(self):
        # At interpreter exit the module globals may already be cleared (the process releases
        # the device memory itself then).
        h = getattr(self, "_h", None)
        lib = getattr(nat, "lib", None)
        if h is not None and h.value and lib is not None:
            lib.sdesim_free_universe(h)
            self._h = None


def parse_equations(processes_equations):
    return Universe(processes_equations)


def _options(rng_method, scheme, seed, precision, scramble, scramble_seed, compat, outputs,
             paths_per_lane, block_size, sobol_first, n_times):
    o = nat.default_options()
    if scheme == "euler":
        o.scheme = nat.EULER
    elif scheme == "runge-kutta":
        o.scheme = nat.RUNGE_KUTTA
    elif n_times >= 2:
        # src/sim/mod.rs:82: any other scheme reaches unimplemented!() at the first step.
        raise nat.PanicException("not implemented")
    if rng_method == "sobol":
        o.rng = nat.SOBOL
    elif rng_method == "injected":
        o.rng = nat.INJECTED
    else:
        o.rng = nat.PSEUDO  # src/sim/mod.rs:54-66: anything but "sobol" is pseudo-random
    try:
        o.precision = _PRECISIONS[precision]
        o.scramble = _SCRAMBLES[scramble]
        o.compat = _COMPAT[compat]
    except KeyError as e:
        raise ValueError("unknown option value %s" % e) from None
    o.outputs = outputs
    o.paths_per_lane = int(paths_per_lane)
    o.block_size = int(block_size)
    o.seed = (_random_seed() if seed is None else int(seed)) & (2 ** 64 - 1)
    o.scramble_seed = int(scramble_seed) & (2 ** 64 - 1)
    o.sobol_first = -1 if sobol_first is None else int(sobol_first)
    return o


def _stream_handle(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


class Plan:
    """Parse once, compile once, run many times on the device.

    ``outputs`` is a subset of {"filtration", "terminal"}. The filtration is the step-major
    device tensor [T][P][S] (float32 or float64); terminal moments are the mean and sample
    variance of each process at the last time point and of ``payoff``, an expression over the
    process names and ``t`` evaluated on the terminal row.
    """

    def __init__(self, universe, time_steps, scenarios, initial_values, rng_method="pseudo", scheme="euler", *,
                 seed=None, precision="f64", scramble="shift", scramble_seed=0, compat="reference",
                 outputs=("filtration",), payoff=None, paths_per_lane=0, block_size=0, sobol_first=None):
        if not isinstance(universe, Universe):
            universe = Universe(universe)
        self.universe = universe
        self.times = np.ascontiguousarray(np.asarray([float(t) for t in time_steps], dtype=np.float64))
        self.scenarios = int(scenarios)
        mask = (nat.OUT_FILTRATION if "filtration" in outputs else 0) | (nat.OUT_TERMINAL if "terminal" in outputs else 0)
        self.options = _options(rng_method, scheme, seed, precision, scramble, scramble_seed, compat, mask,
                                paths_per_lane, block_size, sobol_first, len(self.times))
        self.precision = precision
        self.payoff = payoff
        if isinstance(initial_values, Mapping):
            self.x0 = universe.initial_row(initial_values)
        else:
            # One value per process, in registry order: the library copies exactly that many.
            x0 = np.ascontiguousarray(np.asarray(initial_values, dtype=np.float64))
            if x0.ndim != 1 or x0.size != universe.num_processes:
                raise ValueError("initial_values must be a mapping of process names or a sequence of %d values "
                                 "(one per process), got shape %s" % (universe.num_processes, x0.shape))
            self.x0 = x0 if x0.size else np.zeros(1, dtype=np.float64)
        h = ctypes.c_void_p()
        rc = nat.lib.sdesim_plan_create(universe._h, nat.dptr(self.times), len(self.times), nat.dptr(self.x0),
                                        self.scenarios, ctypes.byref(self.options),
                                        payoff.encode("utf-8") if payoff else None, ctypes.byref(h))
        nat.check(rc, runtime_prefix="Simulation failed: ")
        self._h = h

    @property
    def shape(self):
        return (len(self.times), self.universe.num_processes, self.scenarios)

    @property
    def dtype(self):
        import torch
        return torch.float32 if self.options.precision == nat.F32 else torch.float64

    @property
    def source(self):
        s = nat.lib.sdesim_plan_source(self._h)
        return s.decode("utf-8") if s else ""

    @property
    def code_hash(self):
        """Identity of the compiled kernel (source, device headers, options, hipRTC version), as
        16 hex digits; '' for a plan with no kernel."""
        h = nat.lib.sdesim_plan_code_hash(self._h)
        return "%016x" % h if h else ""

    def new_filtration(self, device="cuda"):
        import torch
        return torch.empty(self.shape, dtype=self.dtype, device=device)

    def _check_filtration(self, filtration):
        if (tuple(filtration.shape) != self.shape or filtration.dtype != self.dtype or not filtration.is_contiguous()
                or filtration.device.type != "cuda"):
            raise ValueError("filtration must be a contiguous %s device tensor of shape %s" % (self.dtype, self.shape))

    def run(self, filtration=None, injected=None, stream=None):
        """Launch on `stream` (default: torch's current stream). Asynchronous."""
        fptr = None
        if filtration is not None:
            self._check_filtration(filtration)
            fptr = ctypes.c_void_p(filtration.data_ptr())
        iptr = None
        if injected is not None:
            import torch
            n = (len(self.times) - 1) * self.universe.num_increments
            if (tuple(injected.shape) != (n, self.scenarios) or not injected.is_contiguous()
                    or injected.dtype != torch.float64 or injected.device.type != "cuda"):
                raise ValueError("injected uniforms must be a contiguous float64 device tensor [N*m][S] = %s"
                                 % ((n, self.scenarios),))
            iptr = ctypes.c_void_p(injected.data_ptr())
        rc = nat.lib.sdesim_plan_run(self._h, fptr, iptr, _stream_handle(stream))
        nat.check(rc, runtime_prefix="Simulation failed: ")

    def moments(self, stream=None):
        """{name: (mean, variance)} at the last time point (+ "payoff"). Synchronises."""
        kq = self.universe.num_processes + (1 if self.payoff else 0)
        out = np.zeros(2 * kq, dtype=np.float64)
        rc = nat.lib.sdesim_plan_moments(self._h, nat.dptr(out), _stream_handle(stream))
        nat.check(rc, runtime_prefix="Simulation failed: ")
        names = list(self.universe.names) + (["payoff"] if self.payoff else [])
        return {n: (float(out[2 * q]), float(out[2 * q + 1])) for q, n in enumerate(names)}

    def to_long(self, filtration, stream=None):
        """Device [T][P][S] -> device [S][T][P] float64 (the long layout's value column)."""
        import torch
        self._check_filtration(filtration)
        T, P, S = self.shape
        out = torch.empty((S, T, P), dtype=torch.float64, device=filtration.device)
        rc = nat.lib.sdesim_filtration_to_long(ctypes.c_void_p(filtration.data_ptr()), self.options.precision, S, T, P,
                                               ctypes.c_void_p(out.data_ptr()), _stream_handle(stream))
        nat.check(rc, runtime_prefix="Simulation failed: ")
        return out

    def __del__(self):
        # At interpreter exit the module globals may already be cleared (the process releases
        # the device memory itself then).
        h = getattr(self, "_h", None)
        lib = getattr(nat, "lib", None)
        if h is not None and h.value and lib is not None:
            lib.sdesim_plan_free(h)
            self._h = None


def long_table(names, times, scenarios, values, chunk_rows=1 << 24):
    """The reference's long layout (src/filtration.rs:81-116, concatenated in scenario order by
    src/sim/mod.rs:88-91): columns scenario:i32, time:f64, process_name:str, value:f64, rows
    ordered scenario -> time -> process. Returns a polars.DataFrame when polars is importable,
    otherwise the equivalent pyarrow.Table.

    The columns are built in chunks of whole scenarios (about `chunk_rows` rows each), so no
    column is materialised at once beside the values and no string chunk reaches Arrow's 2^31-byte
    limit (a 2^20-path, 1025-step, 2-process run has 2.15e9 rows). The time and process_name
    chunks repeat from one chunk to the next, so each is built once and shared (a shorter last
    chunk is a zero-copy prefix); `values` ([S][T][P] f64) is wrapped without a copy."""
    import pyarrow as pa
    names = list(names)
    T, P, S = len(times), len(names), int(scenarios)
    per = T * P
    vals = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    if vals.size != S * per:
        raise ValueError("values must hold scenarios x len(times) x len(names) entries")
    longest = max([len(n.encode("utf-8")) for n in names] + [1])
    rows = max(1, min(int(chunk_rows), ((1 << 31) - 1) // longest))
    spc = max(1, rows // max(1, per))  # scenarios per chunk
    spc = min(spc, max(1, S))
    tpat = np.repeat(np.asarray(times, dtype=np.float64), P)
    dictionary = pa.array(names, type=pa.string())
    time_chunk = pa.array(np.tile(tpat, spc))
    idx = pa.array(np.tile(np.arange(P, dtype=np.int32), spc * T))
    name_chunk = pa.DictionaryArray.from_arrays(idx, dictionary).dictionary_decode()
    cols = {"scenario": [], "time": [], "process_name": [], "value": []}
    for s0 in range(0, S, spc):
        n = min(spc, S - s0)
        cols["scenario"].append(pa.array(np.repeat(np.arange(s0, s0 + n, dtype=np.int32), per)))
        cols["time"].append(time_chunk if n == spc else time_chunk.slice(0, n * per))
        cols["process_name"].append(name_chunk if n == spc else name_chunk.slice(0, n * per))
        cols["value"].append(pa.array(vals[s0 * per:(s0 + n) * per]))
    types = {"scenario": pa.int32(), "time": pa.float64(), "process_name": pa.string(), "value": pa.float64()}
    table = pa.table({k: pa.chunked_array(v, type=types[k]) for k, v in cols.items()})
    try:
        import polars as pl
    except ImportError:
        return table
    return pl.from_arrow(table)


def simulate(processes_equations, time_steps, scenarios, initial_values, rng_method="pseudo", scheme="euler", *,
             seed=None, precision="f64", scramble="shift", scramble_seed=0, compat="reference", output="dataframe",
             payoff=None, paths_per_lane=0, block_size=0, sobol_first=None, stream=None):
    """Simulate SDE paths on the GPU (src/py_binding.rs:8-56 with src/sim/mod.rs:20-92).

    Positional arguments and errors are the reference's: ``scenarios <= 0`` raises ValueError,
    a malformed equation raises ValueError("Failed to parse equations: ..."), a failed run raises
    RuntimeError("Simulation failed: ..."), and conditions under which the reference panics raise
    PanicException. Path ``s`` draws from ``seed + s``; ``seed=None`` draws a fresh seed as the
    reference does on every call.

    output: "dataframe" (the reference's long table), "filtration" (device tensor [T][P][S]),
    or "terminal" ({name: (mean, variance)} at the last time point, plus "payoff").
    """
    eqs = _as_str_list(processes_equations)
    times = [float(t) for t in time_steps]
    scenarios = _as_i32(scenarios, "scenarios")
    if not isinstance(initial_values, Mapping):
        raise TypeError("'%s' object cannot be converted to 'PyDict'" % type(initial_values).__name__)
    init = {}
    for k, v in initial_values.items():
        if not isinstance(k, str):
            raise TypeError("'%s' object cannot be converted to 'PyString'" % type(k).__name__)
        init[k] = float(v)
    if not isinstance(rng_method, str) or not isinstance(scheme, str):
        raise TypeError("rng_method and scheme must be str")
    if scenarios <= 0:
        raise ValueError("scenarios must be a positive integer")
    universe = Universe(eqs)
    if output == "dataframe":
        opts = _options(rng_method, scheme, seed, precision, scramble, scramble_seed, compat, nat.OUT_FILTRATION,
                        paths_per_lane, block_size, sobol_first, len(times))
        tarr = np.ascontiguousarray(times, dtype=np.float64)
        names = list(init.keys())
        carr = (ctypes.c_char_p * max(1, len(names)))(*[n.encode("utf-8") for n in names])
        varr = np.ascontiguousarray([init[n] for n in names] or [0.0], dtype=np.float64)
        values = np.empty((scenarios, len(times), universe.num_processes), dtype=np.float64)
        rc = nat.lib.sdesim_simulate_long(universe._h, nat.dptr(tarr), len(times), carr, nat.dptr(varr), len(names),
                                          scenarios, ctypes.byref(opts), values.ctypes.data_as(ctypes.c_void_p),
                                          None if stream is None else _stream_handle(stream))
        nat.check(rc, runtime_prefix="Simulation failed: ")
        return long_table(universe.names, times, scenarios, values)
    outs = {"filtration": ("filtration",), "terminal": ("terminal",)}.get(output)
    if outs is None:
        raise ValueError("output must be 'dataframe', 'filtration' or 'terminal'")
    plan = Plan(universe, times, scenarios, init, rng_method, scheme, seed=seed, precision=precision,
                scramble=scramble, scramble_seed=scramble_seed, compat=compat, outputs=outs, payoff=payoff,
                paths_per_lane=paths_per_lane, block_size=block_size, sobol_first=sobol_first)
    if output == "filtration":
        filt = plan.new_filtration()
        plan.run(filt, stream=stream)
        return filt
    plan.run(None, stream=stream)
    return plan.moments(stream=stream)
