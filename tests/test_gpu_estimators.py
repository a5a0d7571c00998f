"""Estimators on the GPU against closed forms and across generators (3 standard errors)."""
import numpy as np
import pytest

from conftest import HESTON

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _terminal(eqs, times, S, init, rng, scheme, **kw):
    import sde_sim_rs
    plan = sde_sim_rs.Plan(eqs, times, S, init, rng, scheme, outputs=("terminal",), **kw)
    plan.run(None)
    torch.cuda.synchronize()
    return plan.moments()


def test_c1_terminal_mean_within_3se_of_closed_form():
    """Config 1 (GBM, Euler, 252 steps, 10,000 paths, pseudo seed 42, FP64): E[X_T] = X0 e^{mu T}.
    The Euler mean X0 (1 + mu dt)^252 differs from it by 5e-4, far below the standard error."""
    times = np.linspace(0.0, 1.0, 253)
    S = 10000
    print("case: c1 terminal mean vs closed form S=%d" % S, flush=True)
    m = _terminal(["dX = (0.05 * X) * dt + (0.2 * X) * dW1"], times, S, {"X": 100.0}, "pseudo", "euler", seed=42)
    mean, var = m["X"]
    se = np.sqrt(var / S)
    exact = 100.0 * np.exp(0.05)
    print("mean %.5f exact %.5f se %.4f" % (mean, exact, se))
    assert abs(mean - exact) <= 3.0 * se, (mean, exact, se)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_heston_mc_and_rqmc_means_agree(precision):
    """Heston Runge-Kutta: the pseudo-random (MC) and Sobol-XOR (RQMC) estimates of E[S_T] and of
    the call payoff agree within 3 combined standard errors."""
    times = np.linspace(0.0, 1.0, 65)
    S = 1 << 16
    init = {"S": 100.0, "v": 0.04}
    print("case: heston mc vs rqmc %s S=%d" % (precision, S), flush=True)
    mc = _terminal(HESTON, times, S, init, "pseudo", "runge-kutta", seed=17, precision=precision,
                   payoff="max(S - 100, 0)")
    qmc = _terminal(HESTON, times, S, init, "sobol", "runge-kutta", scramble="xor", scramble_seed=7,
                    precision=precision, payoff="max(S - 100, 0)")
    for name in ("S", "payoff"):
        se = np.sqrt(mc[name][1] / S + qmc[name][1] / S)
        print("%s mc %.5f rqmc %.5f combined se %.4f" % (name, mc[name][0], qmc[name][0], se))
        assert abs(mc[name][0] - qmc[name][0]) <= 3.0 * se, (name, mc[name][0], qmc[name][0], se)
