"""sde_sim_rs -- MI355X implementation of the sde-sim-rs Python API.

``simulate`` keeps the reference's call shape (python/sde_sim_rs/sde_sim_rs.pyi:6-13,
src/py_binding.rs:8-56): the six positional arguments, their meaning, the long
``(scenario, time, process_name, value)`` result and the reference's error behaviour. Keyword-only
extensions select the seed, precision, Sobol scrambling, cache semantics and output form.
``Universe`` and ``Plan`` expose the parse-once / plan-once / run-many form used by benchmarks and
by callers that keep results on the device.
"""
from ._native import PanicException
from . import models
from .api import Plan, Universe, parse_equations, simulate
from .batch import Batch

__all__ = ["simulate", "parse_equations", "Universe", "Plan", "Batch", "models", "PanicException"]
