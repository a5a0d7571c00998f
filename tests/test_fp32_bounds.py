"""The FP32 rounding bound of the Heston step (tests/fp32_bounds.py) on the CPU: its FP64 path is
the oracle's, an f32 evaluation of the step stays inside it, and planted mistakes leave it; with
Sobol-XOR uniforms and with the pseudo-random draws."""
import numpy as np
import pytest

from conftest import HESTON

import fp32_bounds as B

S, N = 256, 1024
FIRST = (1 << 20) - 8192  # the last 8192 points of the 2^20-path workload start here
SEED = 42


@pytest.fixture(scope="module", params=["xor", "words"])
def case(oracle, request):
    """The uniforms and the oracle's paths: Sobol-XOR points (the last 8192 of the 2^20 points
    start at FIRST) or the pseudo-random draws of ChaCha8(seed + s)."""
    times = np.linspace(0.0, 1.0, N + 1)
    if request.param == "xor":
        y = oracle.sobol_points(2 * N, S, FIRST) ^ oracle.xor_masks(7, 2 * N)[None, :]
        ref = oracle.simulate(HESTON, times, S, {"S": 100.0, "v": 0.04}, scheme="runge-kutta", rng="sobol",
                              scramble="xor", scramble_seed=7, sobol_first=FIRST)
    else:
        y = np.stack([oracle.pseudo_f64(SEED + s, 2 * N) for s in range(S)])
        ref = oracle.simulate(HESTON, times, S, {"S": 100.0, "v": 0.04}, scheme="runge-kutta", rng="pseudo",
                              seed=SEED)
    return times, y, ref, request.param


def test_bound_path_is_the_oracle_path(case):
    times, y, ref, draws = case
    X, _ = B.heston_rk_bound(y, times, draws=draws)
    assert np.all(np.abs(X - ref) <= 1e-12 * np.abs(ref) + 1e-12)


def test_step_bound_contains_an_f32_evaluation(case):
    """The f32 evaluation rounds every operation (more roundings than the fused kernel) and its
    q, log2, sqrt and reciprocal are those of numpy; every one of its steps stays inside the
    one-step bound, and the bound is tight (a few roundings of the values)."""
    times, y, ref, draws = case
    em = B.emulate_f32(y, times, draws=draws)
    P, R = B.heston_rk_bound(y, times, states=em, draws=draws)
    d = np.abs(em[:, 1:] - P[:, 1:])
    assert np.all(d <= R[:, 1:])
    assert np.median(R[:, 1:, 0] / np.abs(em[:, 1:, 0])) < 1e-6
    assert np.median(R[:, 1:, 1]) < 1e-7


def test_propagated_bound_contains_an_f32_path_over_the_first_steps(case):
    times, y, ref, draws = case
    em = B.emulate_f32(y, times, draws=draws)
    X, E = B.heston_rk_bound(y, times, draws=draws)
    n = 65
    assert np.all(np.abs(em[:, :n] - ref[:, :n]) <= E[:, :n] + 1e-12 * np.abs(ref[:, :n]) + 1e-13)
    assert np.max(E[:, 32, 0] / ref[:, 32, 0]) < 1e-5


@pytest.mark.parametrize("kw", [dict(c2=0.7141428428542850 * 0.5 * (1 + 1e-4)), dict(theta=0.04 * (1 + 1e-4)),
                                dict(kappa=2.0 * (1 - 1e-4)), dict(r=0.02 * (1 + 1e-3))])
def test_step_bound_detects_a_wrong_constant(case, kw):
    """An f32 evaluation with one model constant off by 1e-4 (1e-3 for the small drift r) leaves
    the one-step bound computed with the right constants."""
    times, y, ref, draws = case
    bad = B.emulate_f32(y, times, draws=draws, **kw)
    P, R = B.heston_rk_bound(y, times, states=bad, draws=draws)
    assert np.any(np.abs(bad[:, 1:] - P[:, 1:]) > R[:, 1:])


def test_propagated_bound_detects_the_textbook_cache(case):
    """The scheme with stage 1 on the settled row (instead of the previous probe) leaves the bound
    of the reference's scheme within the first 64 steps."""
    times, y, ref, draws = case
    Xt, _ = B.heston_rk_bound(y, times, perturb="textbook", draws=draws)
    _, E = B.heston_rk_bound(y, times, draws=draws)
    n = 65
    assert np.any(np.abs(Xt[:, :n] - ref[:, :n]) > E[:, :n] + 1e-12 * np.abs(ref[:, :n]) + 1e-13)
