"""Device parity: the fused HIP kernels against the oracle on identical inputs.

FP64 bar: |gpu - oracle| <= 1e-12 * |oracle| + 1e-12 per value (identical uniforms; the only
differences are last-ulp libm results and FMA contraction on the device). The pseudo-random
stream is the same ChaCha8 word stream on both sides, so "identical uniforms" holds for the
pseudo, Sobol and injected modes alike.
"""
import os

import numpy as np
import pytest

from conftest import GBM, HESTON, JUMP, JUMP_X0, ROOT

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _close(gpu, ref, rtol=1e-12, atol=1e-12):
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert gpu.shape == ref.shape
    bad = ~(np.abs(gpu - ref) <= rtol * np.abs(ref) + atol)
    bad &= ~(np.isnan(gpu) & np.isnan(ref))
    if bad.any():
        i = np.argwhere(bad)[0]
        raise AssertionError("mismatch at %s: gpu=%r ref=%r (%d bad of %d)"
                             % (tuple(i), gpu[tuple(i)], ref[tuple(i)], int(bad.sum()), bad.size))


def _filtration(eqs, times, S, init, *args, **kw):
    import sde_sim_rs
    plan = sde_sim_rs.Plan(eqs, times, S, init, *args, **kw)
    f = plan.new_filtration()
    plan.run(f)
    torch.cuda.synchronize()
    return plan, f.cpu().numpy()  # [T][P][S]


def test_c1_gbm_euler_pseudo_full_filtration(oracle):
    """Config 1: GBM, Euler, 252 steps, 10,000 paths, pseudo seed 42, FP64."""
    times = np.linspace(0.0, 1.0, 253)
    print("case: c1 gbm euler pseudo S=10000 T=253", flush=True)
    plan, f = _filtration(GBM, times, 10000, {"X": 100.0}, "pseudo", "euler", seed=42)
    ref = oracle.simulate(GBM, times, 10000, {"X": 100.0}, seed=42)  # [S][T][P]
    _close(f.transpose(2, 0, 1), ref)
    xT = f[-1, 0, :]
    se = xT.std(ddof=1) / np.sqrt(xT.size)
    assert abs(xT.mean() - 100.0 * np.exp(0.05)) < 3.5 * se


@pytest.mark.parametrize("scheme", ["euler", "runge-kutta"])
@pytest.mark.parametrize("compat", ["reference", "textbook"])
def test_heston_injected_uniforms(oracle, scheme, compat):
    times = np.linspace(0.0, 1.0, 65)
    S = 777
    rng = np.random.default_rng(1)
    u = rng.random((S, 64 * 2))
    print("case: heston injected %s %s S=%d" % (scheme, compat, S), flush=True)
    import sde_sim_rs
    plan = sde_sim_rs.Plan(HESTON, times, S, {"S": 100.0, "v": 0.04}, "injected", scheme, compat=compat)
    f = plan.new_filtration()
    inj = torch.from_numpy(np.ascontiguousarray(u.T)).cuda()
    plan.run(f, injected=inj)
    torch.cuda.synchronize()
    ref = oracle.simulate(HESTON, times, S, {"S": 100.0, "v": 0.04}, scheme=scheme, rng="injected", injected=u,
                          compat=compat)
    _close(f.cpu().numpy().transpose(2, 0, 1), ref)


@pytest.mark.parametrize("scheme", ["euler", "runge-kutta"])
def test_jump_algebraic_example_pseudo(oracle, scheme):
    """examples/example.py universe: dN jumps shared across processes plus an algebraic process."""
    times = np.arange(0.0, 2.0, 0.01)
    S = 3000
    print("case: jump+algebraic %s S=%d" % (scheme, S), flush=True)
    _, f = _filtration(JUMP, times, S, JUMP_X0, "pseudo", scheme, seed=12345)
    ref = oracle.simulate(JUMP, times, S, JUMP_X0, scheme=scheme, seed=12345)
    _close(f.transpose(2, 0, 1), ref, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("scramble", ["shift", "xor"])
@pytest.mark.parametrize("scheme", ["euler", "runge-kutta"])
def test_heston_sobol(oracle, scramble, scheme):
    times = np.linspace(0.0, 1.0, 33)
    S = 1000
    print("case: heston sobol %s %s S=%d" % (scramble, scheme, S), flush=True)
    _, f = _filtration(HESTON, times, S, {"S": 100.0, "v": 0.04}, "sobol", scheme, seed=99, scramble=scramble,
                       scramble_seed=7)
    ref = oracle.simulate(HESTON, times, S, {"S": 100.0, "v": 0.04}, scheme=scheme, rng="sobol", seed=99,
                          scramble=scramble, scramble_seed=7)
    _close(f.transpose(2, 0, 1), ref)


@pytest.mark.parametrize("scramble,first,ppl,block", [("xor", 37, 1, 256), ("xor", 192, 2, 128),
                                                      ("xor", 0, 2, 64), ("shift", 0, 1, 64)])
def test_sobol_first_index_and_launch_shape(oracle, scramble, first, ppl, block):
    """A first point index that is a multiple of 64 lets a wave read its upper table slices with
    scalar loads; any other index gathers every slice per lane. Both, over partial last waves
    and several launch shapes, give the oracle's points."""
    times = np.linspace(0.0, 1.0, 17)
    S = 1000
    print("case: heston sobol %s first=%d ppl=%d block=%d S=%d" % (scramble, first, ppl, block, S), flush=True)
    _, f = _filtration(HESTON, times, S, {"S": 100.0, "v": 0.04}, "sobol", "runge-kutta", seed=99,
                       scramble=scramble, scramble_seed=7, sobol_first=first, paths_per_lane=ppl, block_size=block)
    ref = oracle.simulate(HESTON, times, S, {"S": 100.0, "v": 0.04}, scheme="runge-kutta", rng="sobol", seed=99,
                          scramble=scramble, scramble_seed=7, sobol_first=first)
    _close(f.transpose(2, 0, 1), ref)


@pytest.mark.parametrize("m", [1, 3, 5])
@pytest.mark.parametrize("scheme", ["euler", "runge-kutta"])
def test_xor_increment_counts(oracle, m, scheme):
    """XOR mode with 1, 3 and 5 increments per step: the wave-aligned table form loads a lane's m
    slice-0 words of a step in one to two vector loads (one, three, four + one words)."""
    eqs = ["dX = (0.05 * X) * dt + " + " + ".join("(%r * X) * dW%d" % (0.1 + 0.02 * k, k + 1) for k in range(m))]
    times = np.linspace(0.0, 1.0, 9)
    S = 1000
    print("case: xor m=%d %s S=%d" % (m, scheme, S), flush=True)
    _, f = _filtration(eqs, times, S, {"X": 100.0}, "sobol", scheme, seed=5, scramble="xor", scramble_seed=11)
    ref = oracle.simulate(eqs, times, S, {"X": 100.0}, scheme=scheme, rng="sobol", seed=5, scramble="xor",
                          scramble_seed=11)
    _close(f.transpose(2, 0, 1), ref)


def test_xor_four_table_slices(oracle):
    """Beyond 2^22 paths the Gray codes need 23 bits: four table slices, three of them read with
    scalar loads in the wave-aligned form. 2^22 + 100 paths also leave a partial last wave."""
    eqs = ["dX = (0.05 * X) * dt + (0.2 * X) * dW1"]
    times = np.linspace(0.0, 1.0, 3)
    S = (1 << 22) + 100
    print("case: xor four slices S=%d" % S, flush=True)
    _, f = _filtration(eqs, times, S, {"X": 100.0}, "sobol", "euler", seed=1, scramble="xor", scramble_seed=3)
    ref = oracle.simulate(eqs, times, S, {"X": 100.0}, scheme="euler", rng="sobol", seed=1, scramble="xor",
                          scramble_seed=3)
    _close(f.transpose(2, 0, 1), ref)


def test_fp32_xor_tail_increments(oracle):
    """FP32 XOR increments with qi = min(y, ~y) < 1233 (|z| > 5) read the host-tabulated tail. Over
    2^21 points of dimension 0 (van der Corput) some scrambled points land there; after one GBM
    step every path's FP32 value matches FP64, the tail paths included."""
    import sde_sim_rs
    S = 1 << 21
    eqs = ["dX = (0.05 * X) * dt + (0.2 * X) * dW1"]
    times = np.array([0.0, 1.0])
    y = oracle.sobol_points(1, S)[:, 0] ^ oracle.xor_masks(7, 1)[0]
    qi = np.minimum(y, ~y)
    tail = np.nonzero(qi < 1233)[0]
    print("case: fp32 xor tail S=%d tail paths %d" % (S, tail.size), flush=True)
    assert tail.size > 0
    out = {}
    for prec in ("f32", "f64"):
        plan = sde_sim_rs.Plan(eqs, times, S, {"X": 100.0}, "sobol", "euler", precision=prec, scramble="xor",
                               scramble_seed=7, seed=0)
        f = plan.new_filtration()
        plan.run(f)
        torch.cuda.synchronize()
        out[prec] = f[1, 0].double().cpu().numpy()
    z = (out["f64"][tail] - 105.0) / 20.0
    assert np.all(np.abs(z) > 5.0), z
    np.testing.assert_allclose(out["f32"][tail], out["f64"][tail], rtol=0, atol=2e-4)
    np.testing.assert_allclose(out["f32"], out["f64"], rtol=0, atol=2e-4)


@pytest.mark.parametrize("rng", ["pseudo", "sobol"])
@pytest.mark.parametrize("scheme", ["euler", "runge-kutta"])
def test_fp32_pseudo_matches_oracle(oracle, scheme, rng):
    """FP32 pseudo-random and Sobol-shift paths (the uniform held as a 53-bit fraction in two
    words: the ChaCha draw, or the Sobol point plus the draw's shift modulo 1; the inverse normal
    and the RK sign formed from the words) follow the oracle's FP64 paths to FP32 accuracy: GBM,
    252 steps, 5000 paths (a partial last block), every value within 5e-5 relative."""
    times = np.linspace(0.0, 1.0, 253)
    S = 5000
    print("case: fp32 %s %s gbm S=%d N=252" % (rng, scheme, S), flush=True)
    _, f = _filtration(GBM, times, S, {"X": 100.0}, rng, scheme, seed=42, precision="f32")
    ref = oracle.simulate(GBM, times, S, {"X": 100.0}, scheme=scheme, rng=rng, seed=42)
    _close(f.transpose(2, 0, 1), ref, rtol=5e-5, atol=0.0)


def test_terminal_moments_match_filtration():
    import sde_sim_rs
    times = np.linspace(0.0, 1.0, 101)
    S = 50000
    kw = dict(seed=3, scramble="xor", scramble_seed=7)
    print("case: terminal vs filtration heston rk S=%d" % S, flush=True)
    p1 = sde_sim_rs.Plan(HESTON, times, S, {"S": 100.0, "v": 0.04}, "sobol", "runge-kutta", outputs=("filtration",), **kw)
    f = p1.new_filtration()
    p1.run(f)
    p2 = sde_sim_rs.Plan(HESTON, times, S, {"S": 100.0, "v": 0.04}, "sobol", "runge-kutta", outputs=("terminal",),
                         payoff="max(S - 100, 0)", **kw)
    p2.run(None)
    m = p2.moments()
    torch.cuda.synchronize()
    last = f[-1].cpu().numpy()
    payoff = np.maximum(last[0] - 100.0, 0.0)
    for name, vals in (("S", last[0]), ("v", last[1]), ("payoff", payoff)):
        assert abs(m[name][0] - vals.mean()) <= 1e-12 * abs(vals.mean()) + 1e-14
        assert abs(m[name][1] - vals.var(ddof=1)) <= 1e-9 * vals.var(ddof=1)


def test_fp32_estimators_close_to_fp64():
    import sde_sim_rs
    times = np.linspace(0.0, 1.0, 1025)
    S = 1 << 16
    out = {}
    print("case: fp32 vs fp64 heston rk sobol-xor S=%d N=1024" % S, flush=True)
    for prec in ("f64", "f32"):
        p = sde_sim_rs.Plan(HESTON, times, S, {"S": 100.0, "v": 0.04}, "sobol", "runge-kutta", precision=prec,
                            scramble="xor", scramble_seed=7, outputs=("terminal",), payoff="max(S - 100, 0)")
        p.run(None)
        out[prec] = p.moments()
    for name in ("S", "payoff"):
        m64, m32 = out["f64"][name][0], out["f32"][name][0]
        assert abs(m32 - m64) <= 1e-6 * abs(m64), (name, m32, m64)


def test_dataframe_layout_matches_oracle(oracle):
    import sde_sim_rs
    times = [0.0, 0.25, 0.5, 1.0]
    print("case: dataframe layout", flush=True)
    df = sde_sim_rs.simulate(JUMP, times, 5, JUMP_X0, "pseudo", "euler", seed=77)
    names = df.column_names if hasattr(df, "column_names") else df.columns
    assert list(names) == ["scenario", "time", "process_name", "value"]
    tab = df if hasattr(df, "column") else df.to_arrow()
    assert tab.num_rows == 5 * 4 * 3
    ref = oracle.simulate(JUMP, times, 5, JUMP_X0, seed=77)
    _close(np.asarray(tab.column("value").to_numpy()).reshape(5, 4, 3), ref)
    assert tab.column("scenario").to_pylist()[:13] == [0] * 12 + [1]
    assert tab.column("process_name").to_pylist()[:4] == ["X0", "X1", "X2", "X0"]
    assert tab.column("time").to_pylist()[:4] == [0.0, 0.0, 0.0, 0.25]


def test_single_time_point_and_single_scenario(oracle):
    """Edge shapes: one time point (no steps: the filtration is the initial row, the terminal
    moments are those of a constant) and one scenario (a single live lane in the grid)."""
    import sde_sim_rs
    print("case: one time point; one scenario", flush=True)
    f = sde_sim_rs.simulate(HESTON, [0.0], 3, {"S": 100.0, "v": 0.04}, "sobol", "runge-kutta", seed=1,
                            output="filtration")
    torch.cuda.synchronize()
    got = f.cpu().numpy()
    assert got.shape == (1, 2, 3)
    np.testing.assert_array_equal(got[0, 0], 100.0)
    np.testing.assert_array_equal(got[0, 1], 0.04)
    m = sde_sim_rs.simulate(GBM, [0.0], 7, {"X": 100.0}, "pseudo", "euler", seed=1, output="terminal")
    assert m["X"][0] == 100.0 and m["X"][1] == 0.0
    times = np.linspace(0.0, 1.0, 65)
    _, f1 = _filtration(HESTON, times, 1, {"S": 100.0, "v": 0.04}, "pseudo", "runge-kutta", seed=5)
    ref = oracle.simulate(HESTON, times, 1, {"S": 100.0, "v": 0.04}, scheme="runge-kutta", rng="pseudo", seed=5)
    _close(f1.transpose(2, 0, 1), ref)


@pytest.mark.parametrize("precision,rtol", [("f64", 1e-12), ("f32", 2e-5)])
def test_nonuniform_grid_xor(oracle, precision, rtol):
    """A non-uniform time grid (per-step dt and sqrt(dt) arrays instead of the uniform-grid
    constants), Sobol-XOR RK: FP64 at the parity bar, FP32 to FP32 accuracy (GBM, 3000 paths)."""
    rng = np.random.default_rng(4)
    times = np.concatenate([[0.0], np.sort(rng.uniform(0.0, 1.0, 39))])
    S = 3000
    print("case: nonuniform grid xor rk %s S=%d T=%d" % (precision, S, len(times)), flush=True)
    _, f = _filtration(GBM, times, S, {"X": 100.0}, "sobol", "runge-kutta", scramble="xor", scramble_seed=9,
                       precision=precision)
    ref = oracle.simulate(GBM, times, S, {"X": 100.0}, scheme="runge-kutta", rng="sobol", scramble="xor",
                          scramble_seed=9)
    _close(f.transpose(2, 0, 1), ref, rtol=rtol, atol=0.0 if precision == "f32" else 1e-12)


@pytest.mark.parametrize("model,T,precision,S", [("heston", 300, "f32", 300), ("heston", 300, "f64", 300),
                                                 ("heston", 300, "f32", 301), ("gbm", 301, "f32", 4096),
                                                 ("heston", 70, "f32", 4096)])
def test_long_layout_transpose(model, T, precision, S):
    """The long layout's value column ([S][T][P] f64) is the exact transpose of the device
    filtration ([T][P][S]), widened to f64: whole 256-row tiles and partial tiles in both
    directions, a path count that is not a multiple of the vector width, and an odd row count
    (element-wise accesses)."""
    eqs, init = (HESTON, {"S": 100.0, "v": 0.04}) if model == "heston" else (GBM, {"X": 100.0})
    times = np.linspace(0.0, 1.0, T)
    print("case: long layout %s %s S=%d T=%d" % (model, precision, S, T), flush=True)
    plan, _ = _filtration(eqs, times, S, init, "pseudo", "euler", seed=3, precision=precision)
    dev = plan.new_filtration()
    plan.run(dev)
    out = plan.to_long(dev)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = dev.cpu().numpy().astype(np.float64).transpose(2, 0, 1)
    assert got.dtype == np.float64 and got.shape == want.shape
    np.testing.assert_array_equal(got, want)


def test_concurrent_plans_from_threads():
    """Plans created and run from several host threads at once (the GIL is released in every
    native call; the compiled-module cache is shared and locked) give the results of the same
    plans run one after another. Each thread uses its own stream."""
    import threading
    import sde_sim_rs
    times = np.linspace(0.0, 1.0, 33)
    cases = [(GBM, {"X": 100.0}, "pseudo", "euler", s) for s in (1, 2)] + \
            [(HESTON, {"S": 100.0, "v": 0.04}, "sobol", "runge-kutta", s) for s in (3, 4)]
    print("case: %d concurrent plans" % len(cases), flush=True)

    def run(case, stream=None):
        eqs, init, rng, scheme, seed = case
        plan = sde_sim_rs.Plan(eqs, times, 2000, init, rng, scheme, seed=seed, scramble="xor", scramble_seed=seed)
        f = plan.new_filtration()
        if stream is None:
            plan.run(f)
        else:
            with torch.cuda.stream(stream):
                plan.run(f, stream=stream)
            stream.synchronize()
        torch.cuda.synchronize()
        return f.cpu().numpy()

    want = [run(c) for c in cases]
    got = [None] * len(cases)
    errors = []

    def worker(i):
        try:
            got[i] = run(cases[i], torch.cuda.Stream())
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(cases))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)


def test_plan_run_captured_in_a_graph():
    """Plan.run issues launches only (no allocation, memset or synchronisation), so a caller can
    capture it into a graph: the config-1 shape (10,000 paths x 252 steps, launch-bound) captured
    with torch.cuda.graph, replayed three times with eager work and synchronisation in between,
    gives the eager filtration and moments on every replay. The captured launch holds pointers to
    the plan's device buffers; a plan whose run was captured keeps them until the process ends
    (test_captured_plan_keeps_its_buffers)."""
    import sde_sim_rs
    times = np.linspace(0.0, 1.0, 253)
    print("case: graph capture of Plan.run (filtration + terminal)", flush=True)
    plan = sde_sim_rs.Plan(GBM, times, 10000, {"X": 100.0}, "pseudo", "euler", seed=42,
                           outputs=("filtration", "terminal"))
    f = plan.new_filtration()
    plan.run(f)
    torch.cuda.synchronize()
    want_f = f.cpu().numpy()
    want_m = plan.moments()
    g = torch.cuda.CUDAGraph()
    f.zero_()
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        plan.run(f)
    for rep in range(3):
        f.fill_(-1.0)  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(f.cpu().numpy(), want_f, err_msg="replay %d" % rep)
        assert plan.moments() == want_m, rep


def _random_coef(rng, names, depth):
    """A random continuous expression in the DSL's expression language over the process names and
    t: sums, differences, products, guarded quotients and powers, min/max, abs and bounded
    transcendental functions (no comparisons or rounding, whose discontinuities would turn
    last-ulp differences between the two libms into different paths)."""
    if depth == 0 or rng.random() < 0.25:
        c = rng.integers(0, 4)
        if c == 0:
            return str(rng.choice(names))
        if c == 1:
            return "t"
        return "%.4g" % rng.uniform(0.1, 2.0)
    a = _random_coef(rng, names, depth - 1)
    b = _random_coef(rng, names, depth - 1)
    k = rng.integers(0, 12)
    return ["(%s + %s)" % (a, b), "(%s - %s)" % (a, b), "(%s * %s)" % (a, b), "(%s / (1 + abs(%s)))" % (a, b),
            "(abs(%s) ^ 0.5)" % a, "min(%s, %s)" % (a, b), "max(%s, %s, 0.3)" % (a, b), "sin(%s)" % a,
            "cos(%s)" % a, "tanh(%s)" % a, "log(1 + abs(%s))" % a, "atan(%s)" % a][k]


def test_random_universes_match_oracle(oracle):
    """Random equation sets through the code generator and the device kernel against the oracle's
    interpreter: 1-2 processes, 1-2 Wiener increments per process, random bounded coefficient
    expressions, both schemes, pseudo / Sobol-shift / Sobol-XOR, FP64, 17 time points, 300 paths.
    The bar allows for last-ulp differences between the device's and the host's libm."""
    rng = np.random.default_rng(21)
    times = np.linspace(0.0, 1.0, 17)
    for case in range(16):
        P = int(rng.integers(1, 3))
        names = ["X", "Y"][:P]
        eqs = []
        for n in names:
            terms = ["(0.1 * sin(%s)) * dt" % _random_coef(rng, names, 3)]
            for k in range(int(rng.integers(1, 3))):
                terms.append("(0.2 * cos(%s) * %s) * dW%d" % (_random_coef(rng, names, 3), n, k + 1))
            eqs.append("d%s = %s" % (n, " + ".join(terms)))
        scheme = str(rng.choice(["euler", "runge-kutta"]))
        rng_method, scramble = [("pseudo", "shift"), ("sobol", "shift"), ("sobol", "xor")][int(rng.integers(0, 3))]
        init = {n: float(rng.uniform(0.5, 1.5)) for n in names}
        seed = int(rng.integers(0, 1000))
        print("case %d: %s %s/%s %s" % (case, scheme, rng_method, scramble, eqs), flush=True)
        _, f = _filtration(eqs, times, 300, init, rng_method, scheme, seed=seed, scramble=scramble,
                           scramble_seed=seed + 1)
        ref = oracle.simulate(eqs, times, 300, init, scheme=scheme, rng=rng_method, seed=seed, scramble=scramble,
                              scramble_seed=seed + 1)
        _close(f.transpose(2, 0, 1), ref, rtol=1e-10, atol=1e-12)


def test_kernel_code_cache_on_disk(tmp_path, monkeypatch):
    """A compiled kernel is written to the cache directory and loaded from it once the in-process
    table is forgotten (as by a new process); a damaged file is rejected by its checksum before
    the loader sees it, compiled again and rewritten. The results are the same each time."""
    import os
    import sde_sim_rs
    from sde_sim_rs import _native as nat
    monkeypatch.setenv("SDESIM_CACHE_DIR", str(tmp_path))
    eqs = ["dX = (0.0371 * X) * dt + (0.19 * X) * dW1"]  # a source no other test compiles
    times = np.linspace(0.0, 1.0, 9)
    print("case: kernel code cache in %s" % tmp_path, flush=True)

    def run():
        _, f = _filtration(eqs, times, 1000, {"X": 100.0}, "pseudo", "euler", seed=7)
        return f

    nat.lib.sdesim_kernel_cache_reset()
    f1 = run()
    files = sorted(p for p in os.listdir(tmp_path) if p.endswith(".co"))
    assert len(files) == 1, os.listdir(tmp_path)
    path = os.path.join(tmp_path, files[0])
    blob = open(path, "rb").read()
    assert len(blob) > 24
    nat.lib.sdesim_kernel_cache_reset()
    np.testing.assert_array_equal(run(), f1)
    assert open(path, "rb").read() == blob  # loaded, not rewritten
    damaged = blob[:24] + bytes(len(blob) - 24)
    open(path, "wb").write(damaged)
    nat.lib.sdesim_kernel_cache_reset()
    np.testing.assert_array_equal(run(), f1)
    again = open(path, "rb").read()  # recompiled and rewritten
    assert again != damaged and again[:8] == blob[:8] and len(again) > 24
    assert sorted(p for p in os.listdir(tmp_path)) == files  # no temporary files left behind


def _random_universe(rng, jumps=False):
    P = int(rng.integers(1, 3))
    names = ["X", "Y"][:P]
    eqs = []
    for n in names:
        terms = ["(0.1 * sin(%s)) * dt" % _random_coef(rng, names, 3)]
        for k in range(int(rng.integers(1, 3))):
            terms.append("(0.2 * cos(%s) * %s) * dW%d" % (_random_coef(rng, names, 3), n, k + 1))
        if jumps:
            terms.append("(0.05 * cos(%s)) * dN1(1 + abs(sin(%s)))" % (_random_coef(rng, names, 2),
                                                                     _random_coef(rng, names, 2)))
        eqs.append("d%s = %s" % (n, " + ".join(terms)))
    return names, eqs


def test_random_universes_fp32_and_jumps(oracle):
    """More random equation sets: (a) FP32 kernels (reassociated within a step) against the oracle's
    FP64 paths, every value within 1e-4 relative over 16 steps; (b) FP64 with a Poisson jump term
    dN1(lambda(X)) in every equation (the jump count from the sequential inverse CDF of the same
    uniform) against the oracle within 1e-10."""
    rng = np.random.default_rng(33)
    times = np.linspace(0.0, 1.0, 17)
    for case in range(12):
        jumps = case >= 6
        names, eqs = _random_universe(rng, jumps=jumps)
        scheme = str(rng.choice(["euler", "runge-kutta"]))
        rng_method, scramble = [("pseudo", "shift"), ("sobol", "shift"), ("sobol", "xor")][int(rng.integers(0, 3))]
        init = {n: float(rng.uniform(0.5, 1.5)) for n in names}
        seed = int(rng.integers(0, 1000))
        precision = "f64" if jumps else "f32"
        print("case %d: %s %s %s/%s %s" % (case, precision, scheme, rng_method, scramble, eqs), flush=True)
        _, f = _filtration(eqs, times, 300, init, rng_method, scheme, seed=seed, scramble=scramble,
                           scramble_seed=seed + 1, precision=precision)
        ref = oracle.simulate(eqs, times, 300, init, scheme=scheme, rng=rng_method, seed=seed, scramble=scramble,
                              scramble_seed=seed + 1)
        if jumps:
            _close(f.transpose(2, 0, 1), ref, rtol=1e-10, atol=1e-12)
        else:
            _close(f.transpose(2, 0, 1), ref, rtol=1e-4, atol=1e-6)


def test_long_time_grid(oracle):
    """A long grid: 10,000 steps of GBM Runge-Kutta with Sobol-XOR uniforms (10,000 dimensions of
    the direction-number table), 1000 paths, FP64 at the parity bar."""
    times = np.linspace(0.0, 10.0, 10001)
    S = 1000
    print("case: long grid gbm rk sobol-xor S=%d N=10000" % S, flush=True)
    _, f = _filtration(GBM, times, S, {"X": 100.0}, "sobol", "runge-kutta", scramble="xor", scramble_seed=4)
    ref = oracle.simulate(GBM, times, S, {"X": 100.0}, scheme="runge-kutta", rng="sobol", scramble="xor",
                          scramble_seed=4)
    _close(f.transpose(2, 0, 1), ref)


FACTORED = ["dX = (0.1 * X) * dt + (0.3 * X) * dW1 + (-0.2 * X) * dW2",
            "dY = (0.5 * (1 - Y)) * dt + (0.25 * X * 0.5) * dW2 + (0.4 * X) * dW3 + (2 * X) * dW3"]


@pytest.mark.parametrize("rng,scramble", [("pseudo", "shift"), ("sobol", "shift"), ("sobol", "xor")])
@pytest.mark.parametrize("scheme,compat", [("euler", "reference"), ("runge-kutta", "reference"),
                                           ("euler", "textbook"), ("runge-kutta", "textbook")])
def test_fp32_factored_coefficients(oracle, scheme, compat, rng, scramble):
    """FP32 kernels move constant factors from coefficients to increments. In this universe:
    * X's Wiener terms form the group X * (0.3 dW1 - 0.2 dW2);
    * Y's form X * (0.125 dW2 + 0.4 dW3 + 2 dW3);
    * dW1 is used only with 0.3 > 0, so it is formed as 0.3 sqrt(dt) Psi(u) (and its uniform still
      gives the RK sign);
    * dW2 (two constants, one < 0) and dW3 (two constants) are not.
    Both schemes, both cache modes, all three uniform generators. The paths follow the oracle's FP64 paths to FP32 accuracy: 3000 paths, 64 steps, every value
    within 5e-5 relative + 1e-5. The paths reach |value| ~ 11, where one FP32 ulp is ~1e-6, and Y
    passes through 0 by cancellation of O(1) increments, so an absolute term of ~10 ulps at the
    paths' scale is FP32 rounding; a wrong constant or scaling errs by 1e-2 or more."""
    times = np.linspace(0.0, 1.0, 65)
    S = 3000
    init = {"X": 1.0, "Y": 0.5}
    print("case: fp32 factored %s %s %s/%s S=%d" % (scheme, compat, rng, scramble, S), flush=True)
    _, f = _filtration(FACTORED, times, S, init, rng, scheme, seed=17, scramble=scramble, scramble_seed=5,
                       precision="f32", compat=compat)
    ref = oracle.simulate(FACTORED, times, S, init, scheme=scheme, rng=rng, seed=17, scramble=scramble,
                          scramble_seed=5, compat=compat)
    _close(f.transpose(2, 0, 1), ref, rtol=5e-5, atol=1e-5)


def test_plan_refuses_host_and_misshaped_tensors():
    """Tensors handed to a plan are checked before any launch: a host tensor, a wrong shape or
    dtype, or a non-contiguous view raises ValueError (for run, the injected uniforms and
    to_long), and the plan still runs afterwards."""
    import sde_sim_rs
    times = np.linspace(0.0, 1.0, 9)
    print("case: plan tensor checks", flush=True)
    plan = sde_sim_rs.Plan(GBM, times, 100, {"X": 100.0}, "injected", "euler")
    good = plan.new_filtration()
    inj = torch.rand((8, 100), dtype=torch.float64, device="cuda")
    bad = [torch.empty(plan.shape, dtype=torch.float64),                               # host memory
           torch.empty((9, 1, 99), dtype=torch.float64, device="cuda"),                # shape
           torch.empty(plan.shape, dtype=torch.float32, device="cuda"),                # dtype
           torch.empty((100, 1, 9), dtype=torch.float64, device="cuda").permute(2, 1, 0)]  # layout
    for f in bad:
        with pytest.raises(ValueError):
            plan.run(f, injected=inj)
        with pytest.raises(ValueError):
            plan.to_long(f)
    for u in (inj.cpu(), inj.float(), inj[:, :99], inj.t().contiguous().t()):
        with pytest.raises(ValueError):
            plan.run(good, injected=u)
    plan.run(good, injected=inj)
    torch.cuda.synchronize()
    assert torch.isfinite(good).all()


@pytest.mark.parametrize("precision,rtol", [("f64", 1e-12), ("f32", 2e-6)])
def test_universes_without_stochastic_terms(oracle, precision, rtol):
    """Universes with no stochastic increment run without a generator: a deterministic ODE, an
    algebraic-only universe (its value at t0 is the initial row's 0, as the reference leaves it),
    an ODE with an algebraic process, under Euler against the oracle; an empty equation list
    gives an empty table; Runge-Kutta without a stochastic increment panics as in the reference
    (src/sim/runge_kutta.rs:18)."""
    import sde_sim_rs
    times = np.linspace(0.0, 1.0, 11)
    S = 300
    cases = [(["dX = (0.1 * X) * dt + (0.5 * (1 - X)) * dt"], {"X": 2.0}),
             (["Y = 2 * t + 1"], {}),
             (["dX = (0.1 * X) * dt", "Y = X * t"], {"X": 1.0})]
    for eqs, init in cases:
        print("case: no stochastic terms %s %s" % (precision, eqs), flush=True)
        _, f = _filtration(eqs, times, S, init, "pseudo", "euler", seed=1, precision=precision)
        ref = oracle.simulate(eqs, times, S, init)
        _close(f.transpose(2, 0, 1), ref, rtol=rtol, atol=1e-12)
    table = sde_sim_rs.simulate([], list(times), 4, {}, "pseudo", "euler", precision=precision)
    assert table.num_rows == 0 if hasattr(table, "num_rows") else len(table) == 0
    with pytest.raises(sde_sim_rs.PanicException):
        sde_sim_rs.simulate(cases[0][0], list(times), 4, {"X": 2.0}, "pseudo", "runge-kutta", precision=precision)


def _smooth_coef(rng, names, depth):
    """A random smooth expression over the process names and t (sums, products, sin, cos, tanh,
    atan): no kinks, so FP32 rounding stays at FP32 size through the paths."""
    if depth == 0 or rng.random() < 0.3:
        c = rng.integers(0, 3)
        return str(rng.choice(names)) if c == 0 else ("t" if c == 1 else "%.4g" % rng.uniform(0.1, 1.5))
    a, b = _smooth_coef(rng, names, depth - 1), _smooth_coef(rng, names, depth - 1)
    return ["(%s + %s)" % (a, b), "(%s * %s)" % (a, b), "sin(%s)" % a, "cos(%s)" % a, "tanh(%s)" % a,
            "atan(%s)" % a][int(rng.integers(0, 6))]


def test_random_factored_universes_fp32(oracle):
    """Random equation sets built to exercise the FP32 constant factoring: each process has a dt
    term c0 * r0 and 2-3 Wiener terms c_k * r over one shared residual r, with random constants
    of either sign and random increments from dW1-dW3 (shared across processes, sometimes
    repeated within one). Random scheme, cache mode and generator; FP32 paths against the
    oracle's FP64 paths, every value within 1e-4 relative + 1e-5 times the largest |value|."""
    rng = np.random.default_rng(44)
    times = np.linspace(0.0, 1.0, 17)
    for case in range(10):
        P = int(rng.integers(1, 3))
        names = ["X", "Y"][:P]
        eqs = []
        for n in names:
            terms = ["(%.3g * %s) * dt" % (rng.uniform(-0.5, 0.5), _smooth_coef(rng, names, 2))]
            r = _smooth_coef(rng, names, 2)
            for _ in range(int(rng.integers(2, 4))):
                terms.append("(%.3g * %s) * dW%d" % (rng.choice([-1, 1]) * rng.uniform(0.05, 0.4), r,
                                                    int(rng.integers(1, 4))))
            eqs.append("d%s = %s" % (n, " + ".join(terms)))
        scheme = str(rng.choice(["euler", "runge-kutta"]))
        compat = str(rng.choice(["reference", "textbook"]))
        rng_method, scramble = [("pseudo", "shift"), ("sobol", "shift"), ("sobol", "xor")][int(rng.integers(0, 3))]
        init = {n: float(rng.uniform(0.5, 1.5)) for n in names}
        seed = int(rng.integers(0, 1000))
        print("case %d: %s %s %s/%s %s" % (case, scheme, compat, rng_method, scramble, eqs), flush=True)
        _, f = _filtration(eqs, times, 300, init, rng_method, scheme, seed=seed, scramble=scramble,
                           scramble_seed=seed + 1, precision="f32", compat=compat)
        ref = oracle.simulate(eqs, times, 300, init, scheme=scheme, rng=rng_method, seed=seed, scramble=scramble,
                              scramble_seed=seed + 1, compat=compat)
        _close(f.transpose(2, 0, 1), ref, rtol=1e-4, atol=1e-5 * max(1.0, float(np.abs(ref).max())))


def test_captured_plan_keeps_its_buffers():
    """A plan whose run was captured into a graph keeps its device buffers when it is freed (the
    graph may still launch with them); a plan that was only run eagerly frees them. Checked through
    the library's log in a separate process; the graph is not replayed after the free."""
    import subprocess
    import sys
    code = """
import sys, gc
sys.path[:0] = [%r, %r]
import numpy as np, torch, sde_sim_rs
from conftest import GBM
times = np.linspace(0.0, 1.0, 17)
eager = sde_sim_rs.Plan(GBM, times, 1000, {"X": 100.0}, "pseudo", "euler", seed=1)
f = eager.new_filtration(); eager.run(f); torch.cuda.synchronize()
del eager; gc.collect()
print("eager freed", flush=True)
plan = sde_sim_rs.Plan(GBM, times, 1000, {"X": 100.0}, "pseudo", "euler", seed=2)
f2 = plan.new_filtration()
plan.run(f2); torch.cuda.synchronize()
g = torch.cuda.CUDAGraph()
with torch.cuda.graph(g):
    plan.run(f2)
torch.cuda.synchronize()
del plan; gc.collect()
print("captured freed", flush=True)
""" % (os.path.join(ROOT, "sde-sim-rs_amd"), os.path.join(ROOT, "tests"))
    print("case: captured plan keeps its buffers", flush=True)
    env = dict(os.environ, SDESIM_LOG="1")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "eager freed" in r.stdout and "captured freed" in r.stdout
    kept = [ln for ln in r.stderr.splitlines() if "freed after a graph capture" in ln]
    assert len(kept) == 1, r.stderr[-3000:]
