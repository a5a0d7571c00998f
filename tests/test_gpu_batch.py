"""Batched instances on the GPU: built-in models against their DSL forms, estimators, graph
replay and the autotuner."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _small_instances():
    out = []
    rng = np.random.default_rng(2)
    for i, (kind, S, N) in enumerate([(0, 300, 37), (1, 129, 64), (2, 1000, 8), (0, 64, 5), (2, 700, 50),
                                      (1, 1, 16)]):
        p = (0.05, 0.3, 0.0) if kind == 0 else (2.0, 0.07, 0.25)
        x0 = 0.05 if kind == 2 else 100.0
        # dyadic step: t_n = n * 2^-6 is exact, so every t[n+1] - t[n] equals horizon / steps
        out.append(dict(kind=kind, steps=N, paths=S, x0=x0, horizon=N * 2.0 ** -6, p0=p[0], p1=p[1], p2=p[2],
                        seed=int(rng.integers(0, 2 ** 62))))
    return out


def test_builtin_equals_dsl_form_bitwise():
    import sde_sim_rs
    from sde_sim_rs import models
    insts = _small_instances()
    print("case: batch f64 builtins vs DSL forms (%d instances)" % len(insts), flush=True)
    b = sde_sim_rs.Batch(insts, precision="f64", keep_terminal=True)
    b.run()
    term = b.terminal()
    off = 0
    for inst in insts:
        eqs = models.instance_equations(inst["kind"], inst["p0"], inst["p1"], inst["p2"])
        times = np.arange(inst["steps"] + 1) * 2.0 ** -6
        plan = sde_sim_rs.Plan(eqs, times, inst["paths"], {"X": inst["x0"]}, "pseudo", "euler", seed=inst["seed"])
        f = plan.new_filtration()
        plan.run(f)
        torch.cuda.synchronize()
        last = f[-1, 0].cpu().numpy()
        got = term[off:off + inst["paths"]]
        off += inst["paths"]
        assert np.array_equal(got, last), (inst, np.max(np.abs(got - last)))


def test_batch_moments_fp32_vs_fp64_and_closed_form():
    import sde_sim_rs
    from sde_sim_rs import models
    insts = models.batched_instances(n=192, seed=11)
    print("case: batch 192 instances f32/f64 moments", flush=True)
    m, term = {}, {}
    for prec in ("f64", "f32"):
        b = sde_sim_rs.Batch(insts, precision=prec, keep_terminal=True)
        b.run()
        m[prec] = b.moments()
        term[prec] = b.terminal().astype(np.float64)
    off = 0
    for i, inst in enumerate(insts):
        mean64, var64 = m["f64"][i]
        mean32 = m["f32"][i][0]
        se = np.sqrt(var64 / inst["paths"])
        d = term["f32"][off:off + inst["paths"]] - term["f64"][off:off + inst["paths"]]
        off += inst["paths"]
        if inst["kind"] == models.CIR:
            # sigma * sqrt(max(X, 0)) is not Lipschitz at 0: once a path comes near zero its f32
            # and f64 copies decorrelate, so the comparison is statistical: the mean of the
            # per-path differences within 5 of its standard errors, plus the FP32 allowance of the
            # other models (the f32 rounding of the parameters moves the mean itself).
            se_d = d.std(ddof=1) / np.sqrt(d.size) if d.size > 1 else 0.0
            assert abs(d.mean()) <= 5.0 * se_d + 1e-5 * (abs(mean64) + np.sqrt(var64)), (i, inst, d.mean(), se_d)
        else:
            assert abs(mean32 - mean64) <= 1e-5 * abs(mean64), (i, inst, mean32, mean64)
        if inst["kind"] in (models.GBM, models.OU):
            exact = models.discrete_euler_mean(inst)
            assert abs(mean64 - exact) <= 5.0 * se + 1e-12 * abs(exact), (i, inst, mean64, exact, se)


def test_graph_replay_matches_eager():
    import sde_sim_rs
    from sde_sim_rs import models
    insts = models.batched_instances(n=64, seed=5)
    print("case: batch graph replay", flush=True)
    b = sde_sim_rs.Batch(insts, precision="f32")
    b.set_config(2, 128)
    b.run()
    eager = b.moments()
    b.capture()
    stream = torch.cuda.current_stream()
    for _ in range(3):
        b.replay(stream)
        got = b.moments(stream)
        assert np.array_equal(got, eager)
        x = torch.randn(1 << 20, device="cuda")  # unrelated eager work between replays
        torch.cuda.synchronize()
        b.run(stream)
        assert np.array_equal(b.moments(stream), eager)
        del x


def test_autotune_selects_valid_config_and_keeps_results(tmp_path, monkeypatch):
    import sde_sim_rs
    from sde_sim_rs import models
    monkeypatch.setenv("SDESIM_CACHE_DIR", str(tmp_path))
    insts = models.batched_instances(n=48, seed=9)
    print("case: batch autotune", flush=True)
    b = sde_sim_rs.Batch(insts, precision="f32")
    b.run()
    ref = b.moments()
    p, blk, ms = b.autotune()
    assert p in (1, 2, 4) and blk in (64, 128, 256) and ms > 0
    b.run()
    got = b.moments()
    # The same paths, summed in the groups of the chosen (paths per lane, block) configuration.
    # The moments come from compensated f64 sums of d = x - x0, so regrouping moves the mean by
    # a few ulps of |mean| + |mean - x0| and the variance, (sum d^2 - sum d * mean_d) / (S - 1),
    # by a few ulps of var + (mean - x0)^2 (an OU instance far from x0 with a small spread
    # cancels by up to ~1e6 there, well below the Monte Carlo error).
    x0 = np.array([i["x0"] for i in insts])
    shift = np.abs(ref[:, 0] - x0)
    assert np.all(np.abs(got[:, 0] - ref[:, 0]) <= 1e-13 * (np.abs(ref[:, 0]) + shift))
    assert np.all(np.abs(got[:, 1] - ref[:, 1]) <= 1e-13 * (ref[:, 1] + shift ** 2))
    assert (tmp_path / "autotune.txt").exists()
    b2 = sde_sim_rs.Batch(insts, precision="f32")
    p2, blk2, _ = b2.autotune()  # remembered
    assert (p2, blk2) == (p, blk)
