"""Pins for the oracle (CPU): published vectors, library outputs, closed forms and invariants.

Each check is chosen so that a plausible mistake in the restatement (a wrong round count, word
order, polynomial bit, constant, index or cache rule) fails it.
"""
import os

import numpy as np
import pytest

from conftest import GBM, JUMP, JUMP_X0

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------------------------------------
# ChaCha core (src/rng/pseudo.rs via rand_chacha): published block vectors
# ---------------------------------------------------------------------------------------------
def _state(key_words, counter_lo, counter_hi, nonce0, nonce1):
    return np.array([0x61707865, 0x3320646E, 0x79622D32, 0x6B206574] + list(key_words) +
                    [counter_lo, counter_hi, nonce0, nonce1], dtype=np.uint32)


def test_chacha20_rfc8439_block_vector(oracle):
    # RFC 8439 section 2.3.2: key 00..1f, nonce 000000090000004a00000000, block counter 1.
    key = [int.from_bytes(bytes(range(4 * i, 4 * i + 4)), "little") for i in range(8)]
    st = _state(key, 0x00000001, 0x09000000, 0x4A000000, 0x00000000)
    out = oracle.chacha_block(st, 20)
    expect = [0xE4E7F110, 0x15593BD1, 0x1FDD0F50, 0xC47120A3, 0xC7F4D1C7, 0x0368C033, 0x9AAA2204, 0x4E6CD4C3,
              0x466482D2, 0x09AA9F07, 0x05D7C214, 0xA2028BD9, 0xD19C12B5, 0xB94E16DE, 0xE883D0CB, 0x4E3C50A2]
    assert [int(x) for x in out] == expect


def test_chacha20_zero_key_keystream(oracle):
    # The all-zero key / nonce keystream (first block, counter 0): 76 b8 e0 ad a0 f1 3d 90 ...
    ks = bytes.fromhex("76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
                       "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
    out = oracle.chacha_block(_state([0] * 8, 0, 0, 0, 0), 20)
    assert out.astype("<u4").tobytes() == ks


def test_chacha8_zero_key_keystream(oracle):
    # 8 rounds, 256-bit all-zero key and IV (the ChaCha test-vector set's TC1 at 8 rounds).
    ks = bytes.fromhex("3e00ef2f895f40d67f5bb8e81f09a5a12c840ec3ce9a7f3b181be188ef711a1e"
                       "984ce172b9216f419f445367456d5619314a42a3da86b001387bfdb80e0cfe42")
    out = oracle.chacha_block(_state([0] * 8, 0, 0, 0, 0), 8)
    assert out.astype("<u4").tobytes() == ks


def test_pcg32_step_matches_the_pcg_reference_vectors(oracle):
    """The seed expansion's PCG32 step (LCG multiplier, XSH-RR output of the advanced state,
    rotation direction) against the PCG reference's published pcg32 demo output for
    pcg32_srandom_r(42, 54). srandom sets inc = 2 * 54 + 1, advances from state 0, adds 42 and
    advances again; each pcg32_random_r then outputs the state before advancing, which is the
    advance-then-output step (rand_core's order) started one state earlier, at 0 * MUL + inc + 42.
    The increment seed_from_u64 uses (11634580027462260723) is rand_core's constant and stays
    unpinned."""
    inc = (54 << 1) | 1
    start = (0 * 6364136223846793005 + inc + 42) % 2 ** 64
    want = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    assert [int(x) for x in oracle.pcg32(start, inc, 6)] == want


def test_pseudo_stream_structure(oracle):
    """seed_from_u64 + f64 conversion are restated (parity unpinned); what the maths fixes:
    the f64 draws are multiples of 2^-53 in [0, 1), consecutive blocks use consecutive counters,
    different seeds give different keys, and the stream is uniform (KS test)."""
    from scipy import stats
    u = oracle.pseudo_f64(42, 20000)
    assert np.all((u >= 0) & (u < 1))
    assert np.all(np.floor(u * 2.0 ** 53) == u * 2.0 ** 53)
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    # Draw j uses words 2j, 2j+1 of the keystream: draws 8..15 come from block counter 1.
    key = oracle.seed_from_u64(42)
    blk = oracle.chacha_block(_state(key, 1, 0, 0, 0), 8).astype(np.uint64)
    w = (blk[1::2] << np.uint64(32)) | blk[0::2]
    assert np.array_equal((w >> np.uint64(11)).astype(np.float64) * 2.0 ** -53, u[8:16])
    assert not np.array_equal(oracle.seed_from_u64(1), oracle.seed_from_u64(2))
    assert np.array_equal(oracle.pseudo_f64(42, 5, first=3), u[3:8])


# ---------------------------------------------------------------------------------------------
# Sobol (src/rng/sobol.rs via the sobol crate's Joe-Kuo D6 table): scipy's library generator
# ---------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def sobol_golden():
    return np.load(os.path.join(HERE, "golden", "sobol_scipy.npz"))


def test_sobol_oracle_matches_scipy(oracle, sobol_golden):
    g = sobol_golden
    pts = oracle.sobol_points(2048, 300)
    assert np.array_equal(pts[:, g["dims2048_cols"]], g["dims2048_first300"])
    pts = oracle.sobol_points(2048, 64, first=1000)
    assert np.array_equal(pts[:, g["ff1000_cols"]], g["dims2048_from1000"])
    pts = oracle.sobol_points(21201, 64)
    assert np.array_equal(pts[:, g["dims21201_cols"]], g["dims21201_first64"])


def test_sobol_invariants(oracle):
    # Dimension 0 is the van der Corput sequence in Gray-code order; every dimension's first
    # 2^k points are a permutation of the 2^k strata (a (0, k, 1)-net).
    n = 1 << 10
    pts = oracle.sobol_points(64, n).astype(np.uint64)
    gray = np.arange(n, dtype=np.uint64) ^ (np.arange(n, dtype=np.uint64) >> np.uint64(1))
    rev = np.array([int("{:032b}".format(int(g))[::-1], 2) for g in gray], dtype=np.uint64)
    assert np.array_equal(pts[:, 0], rev)
    for d in range(64):
        assert len(np.unique(pts[:, d] >> np.uint64(22))) == n


def test_xor_masks_are_chacha_words(oracle):
    m = oracle.xor_masks(7, 20)
    key = oracle.seed_from_u64(7)
    assert np.array_equal(m[:16], oracle.chacha_block(_state(key, 0, 0, 0, 0), 8))


# ---------------------------------------------------------------------------------------------
# Inverse CDFs (src/proc/increment.rs:160-200)
# ---------------------------------------------------------------------------------------------
def test_inverse_normal_against_ndtri(oracle):
    from scipy.special import ndtri
    p = np.concatenate([np.logspace(-12, -1, 400), np.linspace(0.1, 0.9, 801), 1 - np.logspace(-12, -1, 400)])
    psi = np.array([oracle.inv_normal(x) for x in p])
    err = np.abs(psi - ndtri(p))
    assert err.max() <= 4.5e-4  # A&S 26.2.23 error bound
    assert err.max() > 1e-5     # and not an exact inverse: it is the approximation
    # antisymmetry up to the rounding of 1 - (1 - p)
    q = np.linspace(0.01, 0.49, 49)
    a = np.array([oracle.inv_normal(x) for x in q])
    b = np.array([oracle.inv_normal(1 - x) for x in q])
    assert np.allclose(a, -b, rtol=1e-12, atol=1e-12)
    assert 1e-7 < abs(oracle.inv_normal(0.5)) < 5e-6


def test_inverse_poisson_against_scipy(oracle):
    from scipy.stats import poisson
    rng = np.random.default_rng(0)
    for lam in (0.05, 0.7, 3.0, 25.0):
        u = rng.random(300)
        cdf = poisson.cdf(np.arange(400), lam)
        keep = np.min(np.abs(u[:, None] - cdf[None, :]), axis=1) > 1e-9  # away from ties
        got = np.array([oracle.inv_poisson(x, lam) for x in u[keep]])
        assert np.array_equal(got, poisson.ppf(u[keep], lam).astype(int))
    assert oracle.inv_poisson(0.9, 0.0) == 0 and oracle.inv_poisson(0.9, -1.0) == 0
    assert oracle.inv_poisson(1.0, 1.0) == 200 or oracle.inv_poisson(1.0, 1.0) < 200


# ---------------------------------------------------------------------------------------------
# Stepping schemes: closed forms with injected uniforms
# ---------------------------------------------------------------------------------------------
def test_euler_gbm_closed_form(oracle):
    mu, sig, x0 = 0.05, 0.2, 100.0
    times = np.array([0.0, 0.1, 0.25, 0.7, 1.0])
    S = 7
    u = np.random.default_rng(3).random((S, 4))
    got = oracle.simulate(GBM, times, S, {"X": x0}, rng="injected", injected=u)
    x = np.full(S, x0)
    for n in range(4):
        dt = times[n + 1] - times[n]
        z = np.array([oracle.inv_normal(v) for v in u[:, n]])
        x = x + (mu * x) * dt + (sig * x) * (np.sqrt(dt) * z)
        assert np.allclose(got[:, n + 1, 0], x, rtol=1e-14, atol=0)


def test_runge_kutta_reference_cache_quirk_gbm(oracle):
    """RK as the reference runs it (src/sim/runge_kutta.rs:5-107 with the cache rule of
    src/func.rs:37-39): coefficients of step n's first stage see the previous step's probe row.
    For GBM the recursion is P_0 = X_0;  k1 = mu P_n dt + s P_n dW;
    P_{n+1} = X_n + k1 + s P_n sk sqrt(dt);  k2 = mu P_{n+1} dt + s P_{n+1} dW;
    X_{n+1} = X_n + (k1 + k2) / 2, with sk = +1 if u > 0.5 else -1."""
    mu, sig, x0 = 0.05, 0.2, 100.0
    times = np.linspace(0, 1, 6)
    S = 5
    u = np.random.default_rng(4).random((S, 5))
    got = oracle.simulate(GBM, times, S, {"X": x0}, scheme="runge-kutta", rng="injected", injected=u)
    tb = oracle.simulate(GBM, times, S, {"X": x0}, scheme="runge-kutta", rng="injected", injected=u,
                         compat="textbook")
    X = np.full(S, x0)
    Pn = X.copy()
    Xt = X.copy()
    for n in range(5):
        dt = times[n + 1] - times[n]
        dW = np.sqrt(dt) * np.array([oracle.inv_normal(v) for v in u[:, n]])
        sk = np.where(u[:, n] > 0.5, 1.0, -1.0)
        k1 = mu * Pn * dt + sig * Pn * dW
        P1 = X + k1 + sig * Pn * sk * np.sqrt(dt)
        k2 = mu * P1 * dt + sig * P1 * dW
        X = X + 0.5 * (k1 + k2)
        Pn = P1
        assert np.allclose(got[:, n + 1, 0], X, rtol=1e-13, atol=0)
        # textbook: stage 1 on the settled row
        k1t = mu * Xt * dt + sig * Xt * dW
        P1t = Xt + k1t + sig * Xt * sk * np.sqrt(dt)
        k2t = mu * P1t * dt + sig * P1t * dW
        Xt = Xt + 0.5 * (k1t + k2t)
        assert np.allclose(tb[:, n + 1, 0], Xt, rtol=1e-13, atol=0)


def test_euler_algebraic_stale_cache(oracle):
    """Euler with an algebraic process (src/sim/euler.rs:30-36, src/filtration.rs:70-79): the
    algebraic pass refreshes the cache from row n+1 before any algebraic value of that row is
    set, and the next step's Levy coefficients reuse that cache, so they read 0 for Y."""
    eqs = ["dX = (0.1 * Y) * dt + (1.0) * dW1", "Y = X + 1"]
    times = np.linspace(0, 1, 5)
    S = 3
    u = np.random.default_rng(5).random((S, 4))
    got = oracle.simulate(eqs, times, S, {"X": 2.0, "Y": 10.0}, rng="injected", injected=u)
    X = np.full(S, 2.0)
    Yc = np.full(S, 10.0)  # cache value of Y at step 0 is the initial value
    for n in range(4):
        dt = times[n + 1] - times[n]
        z = np.array([oracle.inv_normal(v) for v in u[:, n]])
        X = X + (0.1 * Yc) * dt + 1.0 * (np.sqrt(dt) * z)
        Y = X + 1
        assert np.allclose(got[:, n + 1, 0], X, rtol=1e-14)
        assert np.allclose(got[:, n + 1, 1], Y, rtol=1e-14)
        Yc = np.zeros(S)
    assert np.all(got[:, 0, 1] == 10.0)


def test_euler_gbm_mean_matches_closed_form(oracle):
    times = np.linspace(0, 1, 253)
    for rng, kw in (("pseudo", {}), ("sobol", {"scramble": "shift"}), ("sobol", {"scramble": "xor", "scramble_seed": 9})):
        v = oracle.simulate(GBM, times, 4000, {"X": 100.0}, rng=rng, seed=42, **kw)[:, -1, 0]
        se = v.std(ddof=1) / np.sqrt(v.size)
        assert abs(v.mean() - 100 * np.exp(0.05)) < 4 * se, (rng, kw, v.mean())


def test_runge_kutta_reference_bias_regression(oracle):
    # The reference's RK is biased for GBM (sk = sign of the driving uniform, stale probe cache):
    # E[X_T] ~ 109.02 for mu=.05, sigma=.2, X0=100, 252 steps, against the exact 105.13.
    v = oracle.simulate(GBM, np.linspace(0, 1, 253), 20000, {"X": 100.0}, scheme="runge-kutta", seed=1)[:, -1, 0]
    se = v.std(ddof=1) / np.sqrt(v.size)
    assert abs(v.mean() - 109.02) < 4 * se + 0.05
    assert v.mean() - 100 * np.exp(0.05) > 10 * se


def test_poisson_jump_counts(oracle):
    # dN with constant intensity: increments are Poisson(lambda dt); with lambda = 0, none.
    times = np.linspace(0, 1, 101)
    v = oracle.simulate(["dN = (1.0) * dN1(3.0)"], times, 2000, {}, seed=8)[:, -1, 0]
    assert abs(v.mean() - 3.0) < 4 * np.sqrt(3.0 / 2000)
    v0 = oracle.simulate(["dN = (1.0) * dN1(0.0)"], times, 50, {}, seed=8)
    assert np.all(v0 == 0.0)


def test_example_universe_runs_and_is_deterministic(oracle):
    times = np.arange(0, 1, 0.01)
    a = oracle.simulate(JUMP, times, 64, JUMP_X0, scheme="runge-kutta", seed=5)
    b = oracle.simulate(JUMP, times, 64, JUMP_X0, scheme="runge-kutta", seed=5, threads=1)
    assert np.array_equal(a, b)
    assert np.all(a[:, :, 2] >= 0)  # X2 = max(X1 - 100, 0)


def test_long_layout_order(oracle):
    from sde_sim_rs.api import long_table
    vals = np.arange(2 * 3 * 2, dtype=np.float64).reshape(2, 3, 2)
    t = long_table(["A", "B"], [0.0, 0.5, 1.0], 2, vals)
    tab = t if hasattr(t, "column") else t.to_arrow()
    assert tab.column("scenario").to_pylist() == [0] * 6 + [1] * 6
    assert tab.column("time").to_pylist() == [0.0, 0.0, 0.5, 0.5, 1.0, 1.0] * 2
    assert tab.column("process_name").to_pylist() == ["A", "B"] * 6
    assert tab.column("value").to_pylist() == list(range(12))
    import pyarrow as pa
    assert tab.schema.field("scenario").type == pa.int32()
    assert tab.schema.field("value").type == pa.float64()
