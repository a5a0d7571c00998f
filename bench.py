"""Benchmark: Heston RQMC path simulation on one MI355X.

Workload (the reference's headline configuration): Heston (S, v) with S0=100, v0=0.04,
kappa=2, theta=0.04, xi=0.5, rho=-0.7, r=0.02, T=1, written in the reference's equation DSL;
"Runge-Kutta first order" scheme with the reference's cache semantics; 1024 steps; 2^20 paths;
Sobol points of dimension 2048 with a digital XOR scramble from seed 7.

One step = one launch of the fused kernel over all 2^20 paths x 1024 steps, inputs resident in
HBM. Headline: FP32 state, full filtration written to HBM (2^20 x 1025 x 2 x 4 B = 8.6 GB).
Reported beside it, in the same run: the plain FP64 kernel (full filtration, 17.2 GB), the
terminal-only modes (European call payoff max(S - 100, 0), FP64 estimator sums) in FP32 and FP64
with the FP32 estimator error against FP64, and the oracle (the C++ restatement of the
reference's CPU simulator) timed on the host's cores on a bounded sample.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "sde-sim-rs_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HESTON = [
    "dS = (0.02 * S) * dt + (max(v, 0)^0.5 * S) * dW1",
    "dv = (2.0 * (0.04 - max(v, 0))) * dt + (-0.7 * 0.5 * max(v, 0)^0.5) * dW1"
    " + (0.7141428428542850 * 0.5 * max(v, 0)^0.5) * dW2",
]
INIT = {"S": 100.0, "v": 0.04}
PAYOFF = "max(S - 100, 0)"
HBM_PEAK_GBS = 8000.0
PROFILE_FILT = "sdesim_rk_f32_sobolxor_filt_heston_2p20x1024.md"
PROFILE_TERM = "sdesim_rk_f32_sobolxor_term_heston_2p20x1024.md"


def _time_plan(plan, out, steps, warmup):
    """Returns (wall ms per step, mean kernel ms from per-launch HIP events)."""
    import torch

    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        plan.run(out, stream=stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        plan.run(out, stream=stream)
        b.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    kern = sum(a.elapsed_time(b) for a, b in ev) / steps
    return wall, kern


def _time_eval(fn, args):
    """Mean wall ms of a synchronous evaluation after max(1, warmup) untimed ones."""
    for _ in range(max(1, args.warmup)):
        fn()
    reps = max(1, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    return (time.perf_counter() - t0) * 1e3 / reps


def _time_fill(buf, steps, warmup):
    """Mean ms of a plain fill of the same buffer (torch's elementwise fill kernel): the write
    rate this many bytes reach on this GPU without any computation in the loop."""
    import torch

    stream = torch.cuda.current_stream()
    for _ in range(max(1, warmup)):
        buf.fill_(1.0)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    for _ in range(steps):
        buf.fill_(1.0)
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def _profile_lines(name):
    """`key: value` lines of a committed rocprofv3 summary under profiles/ (empty if absent)."""
    out = {}
    try:
        for ln in open(os.path.join(ROOT, "profiles", name)):
            if ":" in ln and not ln.startswith(("|", "#", " ", "`")):
                k, v = ln.split(":", 1)
                out[k.strip()] = v.strip()
    except OSError:
        pass
    return out


def _profiled(name, code_hash, keys):
    """The values of `keys` from a committed counter summary, if it records the counters of this
    very kernel (its `kernel_code_hash` line equals the plan's code hash): ({key: float}, None),
    or (None, reason)."""
    d = _profile_lines(name)
    if not d:
        return None, "no counter summary profiles/%s" % name
    if d.get("kernel_code_hash") != code_hash:
        return None, ("profiles/%s measured kernel %s; this run's kernel is %s"
                      % (name, d.get("kernel_code_hash", "(unrecorded)"), code_hash))
    try:
        return {k: float(d[k]) for k in keys}, None
    except (KeyError, ValueError):
        return None, "profiles/%s lacks %s" % (name, ", ".join(keys))


def host_cpu():
    """(model name, online CPUs, CPUs this process may run on)."""
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count() or 1, len(os.sched_getaffinity(0))


def cpu_baseline(paths, steps_n, threads, scramble="xor", runs=3):
    """The oracle (the reference simulator's CPU restatement) on `paths` paths of the headline
    workload, `threads` threads; the median of `runs` timed runs after a small warm-up.
    scramble="shift" runs the reference's own randomisation (a per-path f64 random shift of the
    Sobol points, src/rng/sobol.rs:62-79) instead of the digital XOR shift."""
    import numpy as np

    from oracle import oracle as O

    times = np.linspace(0.0, 1.0, steps_n + 1)
    O.simulate(HESTON, times, 64, INIT, scheme="runge-kutta", rng="sobol", scramble=scramble, scramble_seed=7,
               threads=threads)
    dts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        O.simulate(HESTON, times, paths, INIT, scheme="runge-kutta", rng="sobol", scramble=scramble, scramble_seed=7,
                   threads=threads)
        dts.append(time.perf_counter() - t0)
    dt = sorted(dts)[runs // 2]
    return paths * steps_n / dt, dts


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--paths", type=int, default=1 << 20)
    ap.add_argument("--nsteps", type=int, default=1024)
    ap.add_argument("--ppl", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--cpu-paths", type=int, default=1 << 16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib", action="store_true", help="run a 1 GiB device copy first (HBM counter calibration)")
    ap.add_argument("--only", default="", help="comma list of legs to run: f32,long,f64,term,batch,adjoint,lmm")
    ap.add_argument("--adjoint-storage", default="auto", choices=("auto", "tape", "checkpoints"),
                    help="state storage of the basket and forward-rate adjoints")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: paths are independent and the multi-GPU form "
              "(shard scenarios, one all-reduce of the estimator partials) is not built.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import sde_sim_rs

    S, N = args.paths, args.nsteps
    times = np.linspace(0.0, 1.0, N + 1)
    legs = set(args.only.split(",")) if args.only else {"f32", "long", "f64", "term", "batch", "adjoint", "lmm"}
    common = dict(seed=0, scramble="xor", scramble_seed=7, paths_per_lane=args.ppl, block_size=args.block)
    uni = sde_sim_rs.Universe(HESTON)
    res = {}

    if args.calib:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if "f32" in legs:
        p32 = sde_sim_rs.Plan(uni, times, S, INIT, "sobol", "runge-kutta", precision="f32", **common)
        res["f32_hash"] = p32.code_hash
        f32 = p32.new_filtration()
        res["f32"] = _time_plan(p32, f32, args.steps, args.warmup)
        res["fill_ms"] = _time_fill(f32, args.steps, args.warmup)
        if "long" in legs:
            # The long layout's value column [S][T][P] f64 from the same filtration (one
            # LDS-tiled transpose launch): reads 4 B and writes 8 B per value.
            stream = torch.cuda.current_stream()
            out = p32.to_long(f32, stream=stream)
            reps = max(1, args.steps // 4)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            a.record(stream)
            for _ in range(reps):
                out = p32.to_long(f32, stream=stream)
            b.record(stream)
            torch.cuda.synchronize()
            ms = a.elapsed_time(b) / reps
            nbytes = f32.numel() * (4 + 8)
            res["long"] = {"workload": "filtration [1025][2][2^20] f32 -> long value column [2^20][1025][2] f64",
                           "ms": ms, "bytes": nbytes, "gbs": nbytes / (ms * 1e-3) / 1e9}
            del out
        del f32
    if "f64" in legs:
        p64 = sde_sim_rs.Plan(uni, times, S, INIT, "sobol", "runge-kutta", precision="f64", **common)
        f64 = p64.new_filtration()
        res["f64"] = _time_plan(p64, f64, args.steps, args.warmup)
        del f64
    term = {}
    if "term" in legs:
        for prec in ("f32", "f64"):
            pt = sde_sim_rs.Plan(uni, times, S, INIT, "sobol", "runge-kutta", precision=prec, outputs=("terminal",),
                                 payoff=PAYOFF, **common)
            w, k = _time_plan(pt, None, args.steps, args.warmup)
            m = pt.moments()
            term[prec] = {"ms": k, "path_steps_per_s": S * N / (k * 1e-3), "payoff_mean": m["payoff"][0],
                          "payoff_var": m["payoff"][1], "S_mean": m["S"][0], "code_hash": pt.code_hash}
        torch.cuda.synchronize()

    batch = None
    if "batch" in legs:
        from sde_sim_rs import models
        insts = models.batched_instances(n=4096, seed=11)
        bt = sde_sim_rs.Batch(insts, precision="f32")
        ppl, blk, tune_ms = bt.autotune()
        bt.capture()
        stream = torch.cuda.current_stream()
        for _ in range(args.warmup):
            bt.replay(stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = max(1, args.steps // 4)
        e0.record(stream)
        for _ in range(reps):
            bt.replay(stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        info = bt.info()
        batch = {"workload": "4096 batched GBM/OU/CIR instances (seed 11), Euler, f32, terminal moments",
                 "path_steps": bt.path_steps, "ms_per_launch": ms, "path_steps_per_s": bt.path_steps / (ms * 1e-3),
                 "graph": True, "paths_per_lane": ppl, "block_size": blk, "grid": info["grid"],
                 "work_items": info["work_items"]}

    adj = None
    if "adjoint" in legs:
        from sde_sim_rs.adjoint import Basket, basket_config
        c = basket_config(seed=3)
        AS, AN = 1 << 22, 365
        bk = Basket(8, AN, AS, horizon=1.0, strike=100.0, width=0.5, seed=3, precision="f32",
                    storage=args.adjoint_storage)
        times = {}
        for mode in (False, True):
            for _ in range(max(1, args.warmup)):
                bk.evaluate(c["x0"], c["mu"], c["sigma"], c["rho"], adjoint=mode)
            reps = max(1, args.steps // 4)
            t0 = time.perf_counter()
            for _ in range(reps):
                r = bk.evaluate(c["x0"], c["mu"], c["sigma"], c["rho"], adjoint=mode)
            times[mode] = (time.perf_counter() - t0) * 1e3 / reps
        adj = {"workload": "8-asset correlated GBM basket, softplus call K=100 w=0.5, 365 Euler steps, 2^22 paths, f32",
               "greeks": 52, "adjoint_ms": times[True], "price_ms": times[False],
               "adjoint_over_price": times[True] / times[False],
               "path_steps_per_s_adjoint": AS * AN / (times[True] * 1e-3), "price": r["price"],
               "storage": bk.storage,
               "state_bytes": (AN + 1) * 8 * AS * 4 if bk.storage == "tape" else ((AN + 18) // 19) * 8 * AS * 4}
        del bk
        if adj["storage"] == "tape":
            # The checkpoint form (every 19th state, the rest recomputed in the reverse sweep),
            # timed the same way for comparison.
            bk = Basket(8, AN, AS, horizon=1.0, strike=100.0, width=0.5, seed=3, precision="f32", storage="checkpoints")
            adj["adjoint_ms_checkpoints"] = _time_eval(
                lambda: bk.evaluate(c["x0"], c["mu"], c["sigma"], c["rho"], adjoint=True), args)
            del bk

    lmm = None
    if "lmm" in legs:
        from sde_sim_rs.lmm import ForwardRateModel, forward_rate_config
        c = forward_rate_config(seed=5)
        LS, LN = 1 << 20, 160
        fm = ForwardRateModel(steps=LN, paths=LS, expiry_index=4, strike=0.03, scramble_seed=5, precision="f32",
                              storage=args.adjoint_storage)
        times = {}
        for mode in (False, True):
            for _ in range(max(1, args.warmup)):
                fm.evaluate(c["L0"], c["vol"], c["loadings"], adjoint=mode)
            reps = max(1, args.steps // 4)
            t0 = time.perf_counter()
            for _ in range(reps):
                r = fm.evaluate(c["L0"], c["vol"], c["loadings"], adjoint=mode)
            times[mode] = (time.perf_counter() - t0) * 1e3 / reps
        lmm = {"workload": "40 forwards, 3 factors, 1y x 9y receiver swaption K=0.03, 160 log-Euler steps, "
                           "2^20 paths, Sobol-XOR dim 480, f32",
               "sensitivities": 40, "adjoint_ms": times[True], "price_ms": times[False],
               "forward_steps_per_s_price": LS * LN * 40 / (times[False] * 1e-3), "price": r["price"],
               "storage": fm.storage}
        del fm
        if lmm["storage"] == "tape":
            fm = ForwardRateModel(steps=LN, paths=LS, expiry_index=4, strike=0.03, scramble_seed=5, precision="f32",
                                  storage="checkpoints")
            lmm["adjoint_ms_checkpoints"] = _time_eval(
                lambda: fm.evaluate(c["L0"], c["vol"], c["loadings"], adjoint=True), args)
            del fm

    path_steps = S * N
    line = {"metric": "path-steps/s on 1 MI355X, FP32 vs FP64 plain at equal error; HBM GB/s %peak",
            "unit": "path-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (Sobol-XOR RQMC, seed 7)",
            "config": {"workload": "heston_rk_rqmc_full_filtration", "paths": S, "steps": N, "processes": 2,
                       "sobol_dims": 2 * N, "scramble": "xor", "scramble_seed": 7, "scheme": "runge-kutta",
                       "state": "f32", "estimator_sums": "f64", "parallelism": "single GPU"}}
    if "f32" in res:
        wall, kern = res["f32"]
        bytes_alg = (N + 1) * 2 * S * 4
        line.update({"value": path_steps / (wall * 1e-3), "ms_per_step": wall})
        ach = bytes_alg / (kern * 1e-3) / 1e9
        prof, why = (None, "counters are summarised for the 2^20 x 1024 workload only")
        if (S, N) == (1 << 20, 1024):
            prof, why = _profiled(PROFILE_FILT, res["f32_hash"], ["traffic_bytes_per_launch"])
        line["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": ach / HBM_PEAK_GBS,
                            "traffic": prof["traffic_bytes_per_launch"] if prof else None,
                            "kernel": "sdesim_rk_f32_sobolxor_filt", "kernel_code_hash": res["f32_hash"],
                            "kernel_ms": kern, "algorithmic_bytes": bytes_alg}
        if why:
            line["roofline"]["traffic_note"] = why
        fill_gbs = bytes_alg / (res["fill_ms"] * 1e-3) / 1e9
        line["roofline"].update({"fill_same_bytes_gbs": fill_gbs, "frac_of_fill": ach / fill_gbs})
    if "f64" in res:
        wall, kern = res["f64"]
        bytes_alg = (N + 1) * 2 * S * 8
        line["fp64_plain"] = {"path_steps_per_s": path_steps / (wall * 1e-3), "ms_per_step": wall, "kernel_ms": kern,
                              "hbm_gbs": bytes_alg / (kern * 1e-3) / 1e9}
        if "f32" in res:
            line["fp32_over_fp64"] = res["f64"][0] / res["f32"][0]
    if term:
        e = abs(term["f32"]["payoff_mean"] - term["f64"]["payoff_mean"]) / abs(term["f64"]["payoff_mean"])
        term["fp32_payoff_rel_err_vs_fp64"] = e
        term["fp32_over_fp64"] = term["f64"]["ms"] / term["f32"]["ms"]
        # VALU roofline of the FP32 terminal-only kernel: wave64 VALU instructions per path-step
        # from the committed counter summary, over the kernel's time here, against the MI355X peak
        # of one wave64 FP32 VALU instruction per 2 clocks per SIMD (1024 SIMDs at 2.4 GHz);
        # transcendentals run at a quarter of that rate and count 4 slots in the weighted form.
        prof, why = (None, "counters are summarised for the 2^20 x 1024 workload only")
        if (S, N) == (1 << 20, 1024):
            prof, why = _profiled(PROFILE_TERM, term["f32"]["code_hash"],
                                  ["valu_instructions_per_path_step", "trans_instructions_per_path_step"])
        if why:
            term["valu_roofline"] = None
            term["valu_roofline_note"] = why
        else:
            vi, tr = prof["valu_instructions_per_path_step"], prof["trans_instructions_per_path_step"]
            peak = 1024 * 2.4e9 / 2
            rate = vi * S * N / 64 / (term["f32"]["ms"] * 1e-3)
            term["valu_roofline"] = {"bound": "valu", "achieved": rate, "peak": peak,
                                     "unit": "wave64 VALU instructions/s", "frac": rate / peak,
                                     "frac_trans_weighted": rate * (vi + 3 * tr) / vi / peak,
                                     "valu_per_path_step": vi, "trans_per_path_step": tr,
                                     "kernel": "sdesim_rk_f32_sobolxor_term",
                                     "kernel_code_hash": term["f32"]["code_hash"]}
        line["terminal_only"] = term
    if "long" in res:
        line["long_layout"] = res["long"]
    if batch:
        line["batched"] = batch
    if adj:
        line["adjoint"] = adj
    if lmm:
        line["forward_rate_model"] = lmm
    if not args.no_cpu:
        # Every CPU this process may use, up to the 16 that one GPU's job is allotted on a shared
        # GPU host (a node's CPUs are divided among the jobs of its GPUs).
        model, online, allowed = host_cpu()
        threads = max(1, min(16, allowed))
        v, dts = cpu_baseline(args.cpu_paths, N, threads)
        vs, dts_s = cpu_baseline(args.cpu_paths, N, threads, scramble="shift")
        line["cpu_baseline"] = {"value": v, "unit": "path-steps/s", "cores": threads, "kind": "port",
                                "cpu_model": model, "cpus_online": online, "cpus_allowed": allowed,
                                "sample": "same workload, %d paths x %d steps, oracle restatement of the reference "
                                          "simulator (f64, full filtration, Sobol-XOR); median of 3 runs (%s s)"
                                          % (args.cpu_paths, N, ", ".join("%.1f" % d for d in dts)),
                                "reference_shift_mode": {
                                    "value": vs, "unit": "path-steps/s",
                                    "sample": "the same paths with the reference's own per-path f64 random shift of "
                                              "the Sobol points; median of 3 runs (%s s)"
                                              % ", ".join("%.1f" % d for d in dts_s)}}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
