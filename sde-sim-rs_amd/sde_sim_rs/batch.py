"""Batched instances: many independent 1-D SDEs of uneven size in one persistent launch.

No reference counterpart: the reference runs one universe per simulate() call
(src/sim/mod.rs:20-92). Each instance here equals simulate() of its built-in model's DSL form
(models.instance_equations) on the grid t_n = n * horizon / steps with its own seed.
"""
import ctypes

import numpy as np

from . import _native as nat
from .api import _stream_handle

_FIELDS = ("kind", "steps", "paths", "x0", "horizon", "p0", "p1", "p2", "seed")


class Batch:
    """instances: sequence of dicts (keys kind, steps, paths, x0, horizon, p0, p1, p2, seed)."""

    def __init__(self, instances, precision="f32", keep_terminal=False):
        arr = (nat.Instance * len(instances))()
        for i, inst in enumerate(instances):
            for f in _FIELDS:
                setattr(arr[i], f, inst[f] if f != "seed" else int(inst[f]) & (2 ** 64 - 1))
        self.instances = list(instances)
        self.n = len(instances)
        self.precision = precision
        prec = nat.F32 if precision in ("f32", "float32") else nat.F64
        h = ctypes.c_void_p()
        nat.check(nat.lib.sdesim_batch_create(arr, self.n, prec, 1 if keep_terminal else 0, ctypes.byref(h)))
        self._h = h
        self.total_paths = int(sum(int(x["paths"]) for x in instances))
        self.path_steps = int(sum(int(x["paths"]) * int(x["steps"]) for x in instances))

    def set_config(self, paths_per_lane, block_size):
        nat.check(nat.lib.sdesim_batch_set_config(self._h, paths_per_lane, block_size))

    def autotune(self, stream=None):
        p, b, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        nat.check(nat.lib.sdesim_batch_autotune(self._h, _stream_handle(stream), ctypes.byref(p), ctypes.byref(b),
                                                ctypes.byref(ms)))
        return p.value, b.value, ms.value

    def info(self):
        p, b, g, ni, tp = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        nat.check(nat.lib.sdesim_batch_info(self._h, ctypes.byref(p), ctypes.byref(b), ctypes.byref(g),
                                            ctypes.byref(ni), ctypes.byref(tp)))
        return {"paths_per_lane": p.value, "block_size": b.value, "grid": g.value, "work_items": ni.value,
                "total_paths": tp.value}

    def run(self, stream=None):
        nat.check(nat.lib.sdesim_batch_run(self._h, _stream_handle(stream)))

    def capture(self):
        nat.check(nat.lib.sdesim_batch_capture(self._h))

    def replay(self, stream=None):
        nat.check(nat.lib.sdesim_batch_replay(self._h, _stream_handle(stream)))

    def moments(self, stream=None):
        out = np.zeros((self.n, 2), dtype=np.float64)
        nat.check(nat.lib.sdesim_batch_moments(self._h, nat.dptr(out), _stream_handle(stream)))
        return out

    def terminal(self, stream=None):
        out = np.zeros(self.total_paths, dtype=np.float32 if self.precision in ("f32", "float32") else np.float64)
        nat.check(nat.lib.sdesim_batch_terminal(self._h, out.ctypes.data_as(ctypes.c_void_p), _stream_handle(stream)))
        return out

    def __del__(self):
        # At interpreter exit the module globals may already be cleared (the process releases
        # the device memory itself then).
        h = getattr(self, "_h", None)
        lib = getattr(nat, "lib", None)
        if h is not None and h.value and lib is not None:
            lib.sdesim_batch_free(h)
            self._h = None
