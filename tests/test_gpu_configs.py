"""The benchmarked workloads themselves against the oracle.

* Heston RK with Sobol-XOR points in FP32 (the kernel the benchmark times, 1024 steps): every
  value of 8192 of its 2^20 paths against the oracle's FP64 step, and the full 2^20-path run's
  terminal moments against the terminal-only FP32 kernel and the FP64 kernel.
* The 4096 batched instances: 32 of them path by path against the oracle in FP64, all of them
  in FP32 against FP64 and the discrete closed forms.
* The basket and forward-rate sensitivities at their full sizes against central differences.
"""
import numpy as np
import pytest

from conftest import HESTON

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

INIT = {"S": 100.0, "v": 0.04}
FULL_S, FULL_N = 1 << 20, 1024
BENCH_OPTS = dict(seed=0, scramble="xor", scramble_seed=7, paths_per_lane=0, block_size=0)


def _close(gpu, ref, rtol, atol):
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    bad = ~(np.abs(gpu - ref) <= rtol * np.abs(ref) + atol)
    if bad.any():
        i = tuple(np.argwhere(bad)[0])
        raise AssertionError("mismatch at %s: gpu=%r ref=%r (%d bad of %d)" % (i, gpu[i], ref[i], int(bad.sum()), bad.size))


def test_c2_fp32_timed_kernel_against_oracle(oracle):
    """The timed kernel (FP32 state, wave-aligned Sobol-XOR tables, 2^20-path table slices) on the
    last 8192 of the 2^20 points, 1024 steps, against the oracle's FP64 Heston RK.

    A bound on the whole-path difference does not stay useful over 1024 steps: sqrt(max(v, 0))
    is not Lipschitz at 0 and the variance's diffusion derivative xi / (2 sqrt(v)) compounds, so
    the propagated rounding bound passes 1e-3 relative within ~100 steps. Every value is
    therefore checked one step at a time: row n + 1 of the kernel must equal the oracle's FP64
    step applied to the kernel's row n (stage 1 on the probe row rebuilt from the kernel's rows)
    within what one FP32 step can round (tests/fp32_bounds.py, derived from the step's
    operations, the inverse normal's and the kink). Over the first 64 steps, where the
    propagated bound is still informative, the kernel's paths must also stay within it of the
    oracle's paths. The terminal payoff mean of the 8192 paths must be within 1e-6 of the
    oracle's."""
    import sde_sim_rs
    import fp32_bounds as B
    S = 8192
    first = FULL_S - S
    times = np.linspace(0.0, 1.0, FULL_N + 1)
    print("case: c2 fp32 timed kernel S=%d first=%d N=%d" % (S, first, FULL_N), flush=True)
    full = sde_sim_rs.Plan(HESTON, times, FULL_S, INIT, "sobol", "runge-kutta", precision="f32", **BENCH_OPTS)
    sub = sde_sim_rs.Plan(HESTON, times, S, INIT, "sobol", "runge-kutta", precision="f32", sobol_first=first,
                          **BENCH_OPTS)
    assert sub.source == full.source and "sdesim_rk_f32_sobolxor_filt" in sub.source
    f = sub.new_filtration()
    sub.run(f)
    torch.cuda.synchronize()
    gpu = f.cpu().numpy().astype(np.float64).transpose(2, 0, 1)  # [S][T][P]
    del full
    ref = oracle.simulate(HESTON, times, S, INIT, scheme="runge-kutta", rng="sobol", scramble="xor",
                          scramble_seed=7, sobol_first=first)
    y = oracle.sobol_points(2 * FULL_N, S, first) ^ oracle.xor_masks(7, 2 * FULL_N)[None, :]

    X, E = B.heston_rk_bound(y, times)
    _close(X, ref, 1e-12, 1e-12)  # the bound follows the oracle's paths
    n = 65
    err = np.abs(gpu[:, :n] - ref[:, :n])
    tol = E[:, :n] + 1e-12 * np.abs(ref[:, :n]) + 1e-13
    assert np.all(err <= tol), ("first 64 steps", float(np.max(err / tol)))

    P, R = B.heston_rk_bound(y, times, states=gpu)
    d = np.abs(gpu[:, 1:] - P[:, 1:])
    r = R[:, 1:] + 1e-12 * np.abs(P[:, 1:]) + 1e-13
    worst = np.unravel_index(np.argmax(d / r), d.shape)
    print("one-step check: worst |gpu - step(gpu)| / bound = %.3g at path %d step %d process %d"
          % ((d / r)[worst], worst[0], worst[1] + 1, worst[2]), flush=True)
    assert np.all(d <= r)

    rel = np.abs(gpu - ref) / np.abs(ref)
    print("whole paths: max relative difference S %.3g, v (|v| > 1e-3) %.3g"
          % (rel[..., 0].max(), rel[..., 1][np.abs(ref[..., 1]) > 1e-3].max()), flush=True)
    pay_g = np.maximum(gpu[:, -1, 0] - 100.0, 0.0).mean()
    pay_r = np.maximum(ref[:, -1, 0] - 100.0, 0.0).mean()
    assert abs(pay_g - pay_r) <= 1e-6 * abs(pay_r), (pay_g, pay_r)


@pytest.mark.parametrize("rng", ["pseudo", "shift"])
def test_c2_model_fp32_words_form_against_oracle(oracle, rng):
    """The same Heston RK model in FP32 with the reference's own generators: pseudo-random ChaCha8
    draws (path s seeded seed + s; the reference's default), and Sobol points with the reference's
    per-path f64 random shift (src/rng/sobol.rs:62-79). Both FP32 kernels form the inverse normal
    from the uniform held in two words (the draw, or the shifted point as an exact 53-bit
    fraction). 4096 paths x 1024 steps: every value against the oracle's FP64 step applied to the
    kernel's own row within one FP32 step's rounding (the words form's budget in
    tests/fp32_bounds.py), the first 64 steps against the propagated bound, and the terminal
    payoff mean within 1e-6 of the oracle's."""
    import sde_sim_rs
    import fp32_bounds as B
    S, seed = 4096, 42
    times = np.linspace(0.0, 1.0, FULL_N + 1)
    dims = 2 * FULL_N
    print("case: c2 model fp32 %s S=%d N=%d" % (rng, S, FULL_N), flush=True)
    gen = "pseudo" if rng == "pseudo" else "sobol"
    plan = sde_sim_rs.Plan(HESTON, times, S, INIT, gen, "runge-kutta", precision="f32", seed=seed, scramble="shift")
    assert ("sdesim_rk_f32_%s_filt" % ("pseudo" if rng == "pseudo" else "sobolshift")) in plan.source
    f = plan.new_filtration()
    plan.run(f)
    torch.cuda.synchronize()
    gpu = f.cpu().numpy().astype(np.float64).transpose(2, 0, 1)
    ref = oracle.simulate(HESTON, times, S, INIT, scheme="runge-kutta", rng=gen, seed=seed, scramble="shift")
    draws = np.stack([oracle.pseudo_f64(seed + s, dims) for s in range(S)])
    if rng == "pseudo":
        y = draws
    else:
        # The reference's composition: frac(point 2^-32 + shift), point index 5 + s.
        yx = oracle.sobol_points(dims, S, 5).astype(np.float64) * 2.0 ** -32 + draws
        y = yx - np.trunc(yx)
    X, E = B.heston_rk_bound(y, times, draws="words")
    _close(X, ref, 1e-12, 1e-12)
    n = 65
    err = np.abs(gpu[:, :n] - ref[:, :n])
    tol = E[:, :n] + 1e-12 * np.abs(ref[:, :n]) + 1e-13
    assert np.all(err <= tol), ("first 64 steps", float(np.max(err / tol)))
    P, R = B.heston_rk_bound(y, times, states=gpu, draws="words")
    d = np.abs(gpu[:, 1:] - P[:, 1:])
    r = R[:, 1:] + 1e-12 * np.abs(P[:, 1:]) + 1e-13
    print("one-step check: worst |gpu - step(gpu)| / bound = %.3g" % float(np.max(d / r)), flush=True)
    assert np.all(d <= r)
    pay_g = np.maximum(gpu[:, -1, 0] - 100.0, 0.0).mean()
    pay_r = np.maximum(ref[:, -1, 0] - 100.0, 0.0).mean()
    assert abs(pay_g - pay_r) <= 1e-6 * abs(pay_r), (pay_g, pay_r)


def test_c2_full_size_terminal_moments():
    """2^20 paths x 1024 steps, the timed configuration: the FP32 filtration's terminal row gives
    the FP32 terminal-only kernel's moments (the same paths, summed in another order), and both
    are within 1e-6 of the FP64 kernel's (S_T mean and the call payoff max(S_T - 100, 0))."""
    import sde_sim_rs
    times = np.linspace(0.0, 1.0, FULL_N + 1)
    print("case: c2 full size terminal moments S=%d N=%d" % (FULL_S, FULL_N), flush=True)
    pf = sde_sim_rs.Plan(HESTON, times, FULL_S, INIT, "sobol", "runge-kutta", precision="f32", **BENCH_OPTS)
    f = pf.new_filtration()
    pf.run(f)
    last = f[-1].double()
    payoff = torch.clamp(last[0] - 100.0, min=0.0)
    filt = {"S": (float(last[0].mean()), float(last[0].var())), "v": (float(last[1].mean()), float(last[1].var())),
            "payoff": (float(payoff.mean()), float(payoff.var()))}
    del f, last, payoff
    torch.cuda.empty_cache()
    term = {}
    for prec in ("f32", "f64"):
        p = sde_sim_rs.Plan(HESTON, times, FULL_S, INIT, "sobol", "runge-kutta", precision=prec,
                            outputs=("terminal",), payoff="max(S - 100, 0)", **BENCH_OPTS)
        p.run(None)
        term[prec] = p.moments()
    print("filtration row %r\nterminal f32 %r\nterminal f64 %r" % (filt, term["f32"], term["f64"]), flush=True)
    for name in ("S", "v", "payoff"):
        assert abs(filt[name][0] - term["f32"][name][0]) <= 1e-12 * abs(term["f32"][name][0]), name
        assert abs(filt[name][1] - term["f32"][name][1]) <= 1e-9 * term["f32"][name][1], name
    for name in ("S", "payoff"):
        for m in (filt[name][0], term["f32"][name][0]):
            assert abs(m - term["f64"][name][0]) <= 1e-6 * abs(term["f64"][name][0]), (name, m, term["f64"][name][0])


def _c3_sample(insts):
    """32 instances of the batched workload: the 8 longest of each model kind and 8 others drawn
    at random (seed 0)."""
    from sde_sim_rs import models
    pick = []
    for kind in (models.GBM, models.OU, models.CIR):
        idx = [i for i, x in enumerate(insts) if x["kind"] == kind]
        idx.sort(key=lambda i: -insts[i]["steps"])
        pick += idx[:8]
    rest = [i for i in range(len(insts)) if i not in pick]
    pick += list(np.random.default_rng(0).choice(rest, 8, replace=False))
    return sorted(int(i) for i in pick)


def test_c3_batch_against_oracle(oracle):
    """The 4096-instance workload (seed 11) in FP64 with every terminal value kept: for 32
    instances, the first (up to) 1024 paths equal the oracle's simulation of the instance's DSL
    form on its grid t_n = n T / N with its seed (src/sim/euler.rs:5-37), within 1e-12 relative
    + 1e-12. The batch steps with dt = T / N; the grid's differences t_{n+1} - t_n that the
    oracle uses differ from it by an ulp, which moves the paths by ~1e-14 (CIR through the
    max(X, 0)^0.5 kink included)."""
    import sde_sim_rs
    from sde_sim_rs import models
    insts = models.batched_instances(n=4096, seed=11)
    print("case: c3 batch f64 4096 instances vs oracle on 32", flush=True)
    b = sde_sim_rs.Batch(insts, precision="f64", keep_terminal=True)
    b.run()
    term = b.terminal()
    offs = np.concatenate([[0], np.cumsum([x["paths"] for x in insts])])
    for i in _c3_sample(insts):
        x = insts[i]
        n = min(x["paths"], 1024)
        times = np.arange(x["steps"] + 1) * (x["horizon"] / x["steps"])
        eqs = models.instance_equations(x["kind"], x["p0"], x["p1"], x["p2"])
        ref = oracle.simulate(eqs, times, n, {"X": x["x0"]}, seed=x["seed"])[:, -1, 0]
        got = term[offs[i]:offs[i] + n]
        try:
            _close(got, ref, 1e-12, 1e-12)
        except AssertionError as e:
            raise AssertionError("instance %d %r: %s" % (i, x, e)) from None


def test_c3_full_batch_fp32_fp64_and_closed_forms():
    """All 4096 instances in FP32 and FP64 with every terminal value kept.

    GBM and OU: the FP32 mean within 1e-5 of the FP64 mean, relative to the instance's scale
    |mean| + sd, and the FP64 mean within 5 standard errors of the Euler chain's exact mean.
    CIR: sqrt(max(X, 0)) is not Lipschitz at 0, so near 0 the FP32 and FP64 copies of a path
    decorrelate (an instance with theta ~ 1e-4 differs on ~93 % of its paths) and the comparison
    is statistical: the mean of the per-path differences x32 - x64 within 5 of its standard errors
    (a paired test; 5 sigma over ~1400 instances) plus the FP32 allowance of the other models,
    1e-5 (|mean| + sd) (where the copies stay together, the paired test resolves the f32 rounding
    of the parameters themselves: theta ~ 0.09 moves the mean by ~3e-8 relative)."""
    import sde_sim_rs
    from sde_sim_rs import models
    insts = models.batched_instances(n=4096, seed=11)
    print("case: c3 batch 4096 instances f32/f64 terminal values", flush=True)
    m, term = {}, {}
    for prec in ("f64", "f32"):
        b = sde_sim_rs.Batch(insts, precision=prec, keep_terminal=True)
        b.run()
        m[prec] = b.moments()
        term[prec] = b.terminal().astype(np.float64)
        del b
    offs = np.concatenate([[0], np.cumsum([x["paths"] for x in insts])]).astype(np.int64)
    d = term["f32"] - term["f64"]
    paths = np.diff(offs).astype(np.float64)
    dmean = np.add.reduceat(d, offs[:-1]) / paths
    dvar = (np.add.reduceat(d * d, offs[:-1]) - paths * dmean ** 2) / np.maximum(paths - 1.0, 1.0)
    se_d = np.sqrt(np.maximum(dvar, 0.0) / paths)
    worst = worst_cir = 0.0
    for i, x in enumerate(insts):
        mean64, var64 = m["f64"][i]
        mean32 = m["f32"][i][0]
        sd = np.sqrt(var64)
        if x["kind"] == models.CIR:
            z = abs(dmean[i]) / (se_d[i] + 1e-300)
            worst_cir = max(worst_cir, z if abs(dmean[i]) > 1e-12 else 0.0)
            assert abs(dmean[i]) <= 5.0 * se_d[i] + 1e-5 * (abs(mean64) + sd), (i, x, dmean[i], se_d[i])
        else:
            e = abs(mean32 - mean64) / (abs(mean64) + sd)
            worst = max(worst, e)
            assert e <= 1e-5, (i, x, mean32, mean64)
            exact = models.discrete_euler_mean(x)
            se = sd / np.sqrt(x["paths"])
            assert abs(mean64 - exact) <= 5.0 * se + 1e-12 * abs(exact), (i, x, mean64, exact, se)
    print("GBM/OU: worst |mean32 - mean64| / (|mean| + sd) = %.3g; CIR: worst paired |mean(d)| / se(d) = %.3g"
          % (worst, worst_cir))


def test_c4_full_size_greeks_against_central_differences():
    """The benchmarked basket (8 assets, 2^22 paths, 365 steps, seed 3, softplus call K=100 w=0.5):
    the 52 FP32 adjoint Greeks against FP64 central differences with common random numbers
    (h = 1e-3 relative; the correlation entries bumped symmetrically), within 1e-4 relative (the
    scale is max(|FD|, 1e-3 of the largest Greek of its kind))."""
    from sde_sim_rs.adjoint import Basket, basket_config
    c = basket_config(seed=3)
    S, N = 1 << 22, 365
    print("case: c4 full size basket f32 adjoint vs f64 FD S=%d N=%d" % (S, N), flush=True)
    bk = Basket(8, N, S, horizon=1.0, strike=100.0, width=0.5, seed=3, precision="f32")
    adj = bk.evaluate(c["x0"], c["mu"], c["sigma"], c["rho"], adjoint=True)
    del bk
    fwd = Basket(8, N, S, horizon=1.0, strike=100.0, width=0.5, seed=3, precision="f64", storage="checkpoints")
    price = lambda **kw: fwd.evaluate(**{**c, **kw}, adjoint=False)["price"]  # noqa: E731
    worst = 0.0
    for name in ("x0", "mu", "sigma"):
        for i in range(8):
            v = c[name].copy()
            h = 1e-3 * abs(v[i])
            vp, vm = v.copy(), v.copy()
            vp[i] += h
            vm[i] -= h
            fd = (price(**{name: vp}) - price(**{name: vm})) / (2 * h)
            err = abs(adj[name][i] - fd) / max(abs(fd), 1e-3 * np.max(np.abs(adj[name])))
            worst = max(worst, err)
            assert err <= 1e-4, (name, i, adj[name][i], fd)
        print("  %s done (worst so far %.3g)" % (name, worst), flush=True)
    rho = c["rho"]
    for i in range(8):
        for j in range(i):
            h = 1e-3 * max(abs(rho[i, j]), 0.1)
            rp, rm = rho.copy(), rho.copy()
            rp[i, j] = rp[j, i] = rho[i, j] + h
            rm[i, j] = rm[j, i] = rho[i, j] - h
            fd = (price(rho=rp) - price(rho=rm)) / (2 * h)
            an = adj["rho"][i, j] + adj["rho"][j, i]
            err = abs(an - fd) / max(abs(fd), 1e-3 * np.max(np.abs(adj["rho"])))
            worst = max(worst, err)
            assert err <= 1e-4, ("rho", i, j, an, fd)
    print("max relative error of 52 adjoint Greeks vs FD at 2^22 x 365: %.3g" % worst)


def test_c5_full_size_sensitivities_against_central_differences():
    """The benchmarked forward-rate model (2^20 paths, 160 steps, 1y x 9y receiver swaption K=0.03,
    Sobol-XOR seed 5): the 40 FP32 curve sensitivities against FP64 central differences with
    common random numbers (h = 1e-3 relative), within 1e-4 relative (scale max(|FD|, 1e-3 of the
    largest))."""
    from sde_sim_rs.lmm import ForwardRateModel, forward_rate_config
    c = forward_rate_config(seed=5)
    S, N = 1 << 20, 160
    print("case: c5 full size forward-rate f32 adjoint vs f64 FD S=%d N=%d" % (S, N), flush=True)
    adj = ForwardRateModel(steps=N, paths=S, expiry_index=4, strike=0.03, scramble_seed=5,
                           precision="f32").evaluate(c["L0"], c["vol"], c["loadings"])
    fwd = ForwardRateModel(steps=N, paths=S, expiry_index=4, strike=0.03, scramble_seed=5, precision="f64",
                           storage="checkpoints")
    scale = np.max(np.abs(adj["dL0"]))
    worst = 0.0
    for i in range(40):
        h = 1e-3 * c["L0"][i]
        lp, lm = c["L0"].copy(), c["L0"].copy()
        lp[i] += h
        lm[i] -= h
        fd = (fwd.evaluate(lp, c["vol"], c["loadings"], adjoint=False)["price"]
              - fwd.evaluate(lm, c["vol"], c["loadings"], adjoint=False)["price"]) / (2 * h)
        err = abs(adj["dL0"][i] - fd) / max(abs(fd), 1e-3 * scale)
        worst = max(worst, err)
        assert err <= 1e-4, (i, adj["dL0"][i], fd)
    print("max relative error of 40 curve sensitivities vs FD at 2^20 x 160: %.3g" % worst)


def test_benchmarked_workloads_rerun_bitwise():
    """Run again, every benchmarked workload gives bitwise the same result: the timed filtration
    (2^20 x 1024, FP32), the 4096-instance batch (its waves claim work items from an atomic
    counter, so which wave runs which item changes between runs, while the partial sums are
    placed by item), and the basket and forward-rate adjoints at their full sizes."""
    import sde_sim_rs
    from sde_sim_rs import models
    from sde_sim_rs.adjoint import Basket, basket_config
    from sde_sim_rs.lmm import ForwardRateModel, forward_rate_config
    times = np.linspace(0.0, 1.0, FULL_N + 1)
    print("case: rerun bitwise: filtration, batch, basket, forward-rate", flush=True)
    p = sde_sim_rs.Plan(HESTON, times, FULL_S, INIT, "sobol", "runge-kutta", precision="f32", **BENCH_OPTS)
    f1, f2 = p.new_filtration(), p.new_filtration()
    p.run(f1)
    p.run(f2)
    torch.cuda.synchronize()
    assert torch.equal(f1, f2)
    del f1, f2
    torch.cuda.empty_cache()
    b = sde_sim_rs.Batch(models.batched_instances(n=4096, seed=11), precision="f32")
    b.run()
    m1 = b.moments()
    b.run()
    assert np.array_equal(m1, b.moments())
    del b
    c = basket_config(seed=3)
    bk = Basket(8, 365, 1 << 22, horizon=1.0, strike=100.0, width=0.5, seed=3, precision="f32")
    r1 = bk.evaluate(c["x0"], c["mu"], c["sigma"], c["rho"], adjoint=True)
    r2 = bk.evaluate(c["x0"], c["mu"], c["sigma"], c["rho"], adjoint=True)
    for k in ("price", "x0", "mu", "sigma", "rho"):
        assert np.array_equal(r1[k], r2[k]), k
    del bk
    c = forward_rate_config(seed=5)
    fm = ForwardRateModel(steps=160, paths=1 << 20, expiry_index=4, strike=0.03, scramble_seed=5, precision="f32")
    r1 = fm.evaluate(c["L0"], c["vol"], c["loadings"])
    r2 = fm.evaluate(c["L0"], c["vol"], c["loadings"])
    assert r1["price"] == r2["price"] and np.array_equal(r1["dL0"], r2["dL0"])
