"""sdesim_plan_run_scratch (the allocation-free terminal run) through the C ABI against
sdesim_plan_run + sdesim_plan_moments: identical moments on two calls in a row, the second on a
smaller plan, and a plan too large for the scratch array refused before any launch."""
import ctypes

import numpy as np
import pytest

from conftest import HESTON

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _scratch_run(plan, stream):
    from sde_sim_rs import _native as nat
    kq = plan.universe.num_processes + (1 if plan.payoff else 0)
    out = torch.full((2 * kq,), float("nan"), dtype=torch.float64, device="cuda")
    rc = nat.lib.sdesim_plan_run_scratch(plan._h, None, None, ctypes.c_void_p(out.data_ptr()),
                                         ctypes.c_void_p(stream.cuda_stream))
    nat.check(rc)
    stream.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_scratch_run_matches_plan_run(precision):
    import sde_sim_rs
    stream = torch.cuda.current_stream()
    init = {"S": 100.0, "v": 0.04}
    for S, T in ((1 << 18, 129), (5000, 33)):  # the second call is smaller
        print("case: scratch run %s S=%d T=%d" % (precision, S, T), flush=True)
        times = np.linspace(0.0, 1.0, T)
        plan = sde_sim_rs.Plan(HESTON, times, S, init, "sobol", "runge-kutta", precision=precision, scramble="xor",
                               scramble_seed=7, seed=0, outputs=("terminal",), payoff="max(S - 100, 0)")
        got = _scratch_run(plan, stream)
        plan.run(None, stream=stream)
        m = plan.moments(stream=stream)
        want = np.array([v for name in ("S", "v", "payoff") for v in m[name]])
        assert np.array_equal(got, want), (got, want)


def test_scratch_run_refuses_plans_beyond_the_array():
    import sde_sim_rs
    from sde_sim_rs import _native as nat
    # At 64 paths per block, one block more than the array holds (2 x 3 quantities x 8 B each).
    S = (nat.SCRATCH_BYTES // (2 * 3 * 8)) * 64 + 64
    print("case: scratch run refusal S=%d" % S, flush=True)
    plan = sde_sim_rs.Plan(HESTON, [0.0, 1.0], S, {"S": 100.0, "v": 0.04}, "pseudo", "euler", seed=0,
                           outputs=("terminal",), payoff="max(S - 100, 0)", block_size=64)
    out = torch.zeros(6, dtype=torch.float64, device="cuda")
    rc = nat.lib.sdesim_plan_run_scratch(plan._h, None, None, ctypes.c_void_p(out.data_ptr()), None)
    assert rc == nat.EVALUE and "exceed the scratch array" in nat.last_error()
    assert torch.all(out == 0)
