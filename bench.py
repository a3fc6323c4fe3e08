"""Benchmark: candidate evaluations per second of the full six-stage chain on one MI355X.

One step = one ce_evaluate() call over the whole candidate block of the workload (inputs already
resident in HBM), i.e. simulation, linearised eigensolve, modal fit, margin sweep and scoring of
every candidate.  The default workload is the 4-state two-mass drive (65,536 candidates, T=5000,
F=400); --config c3 selects the 8-state, 262,144-candidate case.

Printed: one JSON line with the throughput of the headline mode, the FP64 mode beside it, the
roofline of the dominant kernel (HIP events on the launch stream) and the CPU oracle timed on the
host cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


F32_KERNELS = ("ce_sim_f32", "k_sim_f32", "k_sweep_mark32")


def kernel_flops(problem, info64, d):
    """Algorithmic flops per launch of each stage kernel, counted for the candidates that reach it
    (FMA = 2, add/mul/div = 1, each non-SKIP instruction = 1; transcendentals not counted)."""
    import numpy as np
    c = problem.config
    ns, nd, nw = int(c["statespace_dim"]), int(c["discrete_dim"]), int(c["nonlin_dim"])
    T, F = int(c["time"]), int(c["freq"])
    nt = ns + nd
    ninstr = int(c["dc_instructions"]) + int(c["u_instructions"]) + sum(c["nl_instructions"] or [0])
    n = info64.shape[0]
    sc = info64[:, d.info_scoring:d.info_scoring + 3]
    stable_lin = ~((sc[:, 0] >= float(c["unstable_err"])) & (info64[:, d.info_allmargin + 3] == 0))
    n_sim = int(stable_lin.sum())
    uniq = info64[:, d.info_modal + 3 * nt]
    fitted = (uniq > 0) & stable_lin
    u = uniq[fitted]
    nc = 2 * u
    tfit = d.fit_count
    per_step = 4 * ns + 5 + ns * (2 * ns + 4) + nd * (2 * nd + 2 * nw + 2) + ninstr
    sim = float(per_step) * T * n_sim
    solve = float((2 * nt) ** 3 / 3 + 4 * (2 * nt) ** 2) * int(fitted.sum())
    per_point = 8.0 / 3.0 * ns ** 3 + 8 * ns * ns + 8 * ns + 30
    swept = int((info64[:, d.info_allmargin + 3] == 1).sum())
    return {
        "ce_sim_f32": sim, "ce_sim_f64": sim, "k_sim_f32": sim, "k_sim_fp64": sim,
        "k_eig_lin": float(10 * nt ** 3 + 4 * ns * ns + 2 * ns * nd + 2 * ns) * n,
        "k_eig_nom": float(10 * (2 * ns) ** 3) * n_sim,
        "k_atb": float(np.sum(8.0 * u * tfit)),
        "k_fitcf": solve + float(np.sum(u * (u + 1) / 2)) * 2 * 40,
        "k_fit_fp64": float(np.sum(tfit * (nc * (nc + 1) + 2 * nc + 4 * u + 1) + 2 * T)) + solve,
        "k_sweep_mark32": F * per_point * swept, "k_sweep_fp64": F * per_point * swept,
    }


def cpu_baseline(problem, target_s=12.0, max_n=None):
    """Oracle (FP64 restatement, OpenMP over candidates) on a bounded sample of the workload."""
    import numpy as np
    from oracle import oracle
    threads = int(os.environ.get("CE_CPU_THREADS", "0")) or oracle.lib().ora_max_threads()
    n_total = problem.gains.shape[0]
    n = min(256, n_total)
    t = 0.0
    while True:
        sub = problem.subset(np.arange(n))
        t0 = time.perf_counter()
        oracle.evaluate(sub, threads=threads)
        t = time.perf_counter() - t0
        if t >= 0.5 * target_s / 4 or n >= n_total or (max_n and n >= max_n):
            break
        n = min(n_total, n * 4)
    # final timed run sized to about target_s
    n_final = int(min(n_total, max(n, n * target_s / max(t, 1e-3))))
    if max_n:
        n_final = min(n_final, max_n)
    sub = problem.subset(np.arange(n_final))
    # repeat the sample until about target_s of CPU work when the whole workload is quicker
    reps = 0
    t0 = time.perf_counter()
    while True:
        oracle.evaluate(sub, threads=threads)
        reps += 1
        t = time.perf_counter() - t0
        if t >= 0.8 * target_s or n_final < n_total:
            break
    return dict(value=reps * n_final / t, unit="candidate-evals/s", cores=threads, kind="port",
                sample="%d x first %d of %d candidates, FP64 oracle (all stages), %.1f s" % (
                    reps, n_final, n_total, t))


def bench_c4(args):
    """Batched uneven Monte Carlo: all plants of a round in one batched chain (grouped persistent
    launches, hipGraph replay, autotuned grids); every round uploads the perturbed plants."""
    import numpy as np
    import torch
    from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, workloads
    from controllerengine_amd.batch import Batch
    t0 = time.perf_counter()
    w = workloads.C4(n_plants=args.plants, rounds=max(args.rounds, 1))
    gen_s = time.perf_counter() - t0
    b = Batch(0)
    for p in w.plants:
        b.add(p, rnd=0)
    d_g, d_i = b.pack([p.gains for p in w.plants], [p.info for p in w.plants])
    d_o = torch.empty_like(d_i)
    n = w.candidates()
    res = {}
    modes = [("mixed", PRECISION_MIXED), ("fp64", PRECISION_FP64)]
    if args.precision != "auto":
        modes = [m for m in modes if m[0] == args.precision]
    for name, prec in modes:
        rounds = args.rounds if name == "mixed" else max(1, min(args.rounds, args.fp64_rounds))
        b.set_options(use_graph=True, autotune=True)
        for r in range(args.warmup):
            b.run_device(prec, d_g, d_i, d_o)
        torch.cuda.synchronize()
        b.set_options(use_graph=True, autotune=False)
        # per-kernel times of one round (events after every grouped launch, graph off)
        b.set_options(use_graph=False, autotune=False, profiling=True)
        b.run_device(prec, d_g, d_i, d_o)
        torch.cuda.synchronize()
        kt = b.kernel_times()
        b.set_options(use_graph=True, autotune=False, profiling=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in range(rounds):
            for j, p in enumerate(w.plants):
                b.update(j, p, rnd=r % w.rounds)
            b.run_device(prec, d_g, d_i, d_o)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        info = b.info()
        res[name] = dict(value=n * rounds / dt, s_per_round=dt / rounds, rounds=rounds, used_graph=info["used_graph"],
                         launches=info["launches"], kernel_ms={k: round(v, 3) for k, v in kt})
    head = modes[0][0]
    line = {
        "metric": "candidate evals/s, mixed vs FP64 plain at matched margins; p50/p99 job latency",
        "value": res[head]["value"], "unit": "candidate-evals/s", "n_gpus": 1, "steps": res[head]["rounds"],
        "warmup": args.warmup, "ms_per_step": 1000 * res[head]["s_per_round"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if head == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": "c4: %d plants of 2-8 states, %d candidates per round (log-uniform 256-16384 per "
                               "plant), T 1000-8000, F=400, serial rounds with +-10%% perturbations, one batch" %
                               (len(w.plants), n), "precision": head},
        "batch": {k: {kk: vv for kk, vv in v.items() if kk != "kernel_ms"} for k, v in res.items()},
        "kernel_ms": res[head]["kernel_ms"], "generator_s": gen_s,
    }
    if "fp64" in res and head != "fp64":
        line["fp64_value"] = res["fp64"]["value"]
        line["speedup_vs_fp64"] = res[head]["value"] / res["fp64"]["value"]
    b.close()
    print(json.dumps(line))


def bench_c5(args):
    """Serving loop: Poisson arrivals of c5 jobs at each rate for `duration` seconds, submitted to a
    ce_server (footprint admission, continuous batching); p50 / p99 latency and throughput."""
    import numpy as np
    from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, workloads
    from controllerengine_amd.server import Server
    rng = np.random.default_rng(4)
    pool = [workloads.c5_job(rng) for _ in range(args.pool)]
    rates = [float(r) for r in args.rates.split(",")]
    modes = [("mixed", PRECISION_MIXED), ("fp64", PRECISION_FP64)]
    if args.precision != "auto":
        modes = [m for m in modes if m[0] == args.precision]
    out = {}
    for name, prec in modes:
        srv = Server(prec, slots=args.slots)
        for j in pool:  # compiles each simulation structure once, outside the measurement
            srv.submit(j)
        srv.drain(600.0)
        out[name] = {}
        for rate in rates:
            srv.reset_stats()
            arr = np.random.default_rng(int(rate) * 7 + 4)
            t0 = time.perf_counter()
            t_next = 0.0
            ids = []
            cands = 0
            while True:
                t_next += arr.exponential(1.0 / rate)
                if t_next > args.duration:
                    break
                now = time.perf_counter() - t0
                if t_next > now:
                    time.sleep(t_next - now)
                j = pool[int(arr.integers(len(pool)))]
                ids.append(srv.submit(j))
                cands += j.gains.shape[0]
            srv.drain(600.0)
            st = srv.stats()
            for jid in ids:
                srv.poll(jid)
            out[name][str(int(rate))] = dict(jobs=int(st["jobs"]), p50_ms=st["p50_ms"], p99_ms=st["p99_ms"],
                                             mean_ms=st["mean_ms"], max_ms=st["max_ms"],
                                             evals_per_s=st["evals_per_s"], offered_evals_per_s=cands / args.duration,
                                             batches=int(st["batches"]),
                                             peak_jobs_in_flight=int(st["peak_jobs_in_flight"]),
                                             peak_batches_in_flight=int(st["peak_batches_in_flight"]))
            print("rate %g %s: %s" % (rate, name, out[name][str(int(rate))]), file=sys.stderr, flush=True)
        srv.close()
    head = modes[0][0]
    top = str(int(max(rates)))
    line = {
        "metric": "candidate evals/s, mixed vs FP64 plain at matched margins; p50/p99 job latency",
        "value": out[head][top]["evals_per_s"], "unit": "candidate-evals/s", "n_gpus": 1,
        "steps": out[head][top]["jobs"], "warmup": len(pool), "ms_per_step": out[head][top]["mean_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": "c5: Poisson job arrivals at %s jobs/s for %g s each; jobs: random stable plants "
                               "of 2-8 states, 0-4 NL, 256-65536 candidates, T 500-5000, F 128-512 (pool of %d)"
                               % (args.rates, args.duration, len(pool)), "precision": head},
        "serving": out,
        "p50_ms": out[head][top]["p50_ms"], "p99_ms": out[head][top]["p99_ms"],
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--n", type=int, default=0, help="candidates (default: the configuration's count)")
    ap.add_argument("--precision", default="auto", choices=["auto", "fp64", "mixed"])
    ap.add_argument("--prof-steps", type=int, default=3, help="extra steps with per-kernel events")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (HBM counter calibration)")
    ap.add_argument("--plants", type=int, default=512, help="c4: plants")
    ap.add_argument("--rounds", type=int, default=64, help="c4: rounds (mixed)")
    ap.add_argument("--fp64-rounds", type=int, default=4, help="c4: rounds timed for the FP64 comparison")
    ap.add_argument("--rates", default="20,50,100", help="c5: arrival rates (jobs/s)")
    ap.add_argument("--duration", type=float, default=10.0, help="c5: seconds of arrivals per rate")
    ap.add_argument("--pool", type=int, default=64, help="c5: distinct generated jobs")
    ap.add_argument("--slots", type=int, default=2, help="c5: batches in flight")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(args.copy_calibration * (1 << 18), dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 4}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the candidate chain has no cross-GPU step "
              "(a multi-GPU run would split the candidate ranges across GPUs); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    if args.config == "c4":
        bench_c4(args)
        return
    if args.config == "c5":
        bench_c5(args)
        return

    import numpy as np
    import torch
    from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, workloads
    from controllerengine_amd.engine import Engine

    gen = workloads.CONFIGS[args.config]
    problem = gen(n=args.n) if args.n else gen()
    n = problem.gains.shape[0]
    d = problem.derived()
    eng = Engine(0)
    cfg = problem.cfg_struct()
    plant = problem.plant_struct(0)
    g = eng.to_device(problem.gains)
    i_in = eng.to_device(problem.info)
    i_out = torch.empty_like(i_in)
    i64 = torch.empty(i_in.shape, dtype=torch.float64, device=i_in.device)

    modes = []
    have_mixed = True
    if args.precision in ("auto", "mixed"):
        try:
            eng.evaluate_device(cfg, plant, PRECISION_MIXED, g, i_in, i_out, None)
            torch.cuda.synchronize()
        except RuntimeError:
            have_mixed = False
    if args.precision == "fp64" or not have_mixed:
        modes = [("fp64", PRECISION_FP64)]
    elif args.precision == "mixed":
        modes = [("mixed", PRECISION_MIXED)]
    else:
        modes = [("mixed", PRECISION_MIXED), ("fp64", PRECISION_FP64)]

    results = {}
    for name, prec in modes:
        for _ in range(args.warmup):
            eng.evaluate_device(cfg, plant, prec, g, i_in, i_out, i64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            eng.evaluate_device(cfg, plant, prec, g, i_in, i_out, i64)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        info64 = i64.cpu().numpy()
        # per-kernel durations from HIP events recorded on the launch stream, separate passes
        eng.set_profiling(True)
        acc = {}
        for _ in range(args.prof_steps):
            eng.evaluate_device(cfg, plant, prec, g, i_in, i_out, i64)
            for k, ms in eng.kernel_times():
                acc.setdefault(k, []).append(ms)
        eng.set_profiling(False)
        kern = {k: sum(v) / len(v) for k, v in acc.items()}
        results[name] = dict(value=n * args.steps / dt, ms=1000 * dt / args.steps, kernels=kern, info64=info64)

    head = modes[0][0]
    r = results[head]
    fl = kernel_flops(problem, r["info64"], d)
    kern = {k: v for k, v in r["kernels"].items() if k != "start"}
    dom = max(kern, key=lambda k: kern[k]) if kern else None
    dom_ms = kern.get(dom)
    peak = FP32_PEAK_TFLOPS if dom in F32_KERNELS else FP64_PEAK_TFLOPS
    achieved = fl[dom] / (dom_ms * 1e-3) / 1e12 if dom in fl else None
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_%s_%s_%s.json" % (dom, args.config, head))
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    line = {
        "metric": "candidate evals/s, mixed vs FP64 plain at matched margins; p50/p99 job latency",
        "value": r["value"],
        "unit": "candidate-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "%s: %s, %d candidates, T=%d, F=%d, %s" % (
            args.config, {"c1": "mass-spring-damper 2-state", "c2": "two-mass drive 4-state, 2 NL",
                          "c3": "random stable 8-state, 4 NL"}[args.config], n, cfg.time, cfg.freq, head),
                   "precision": head},
        "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
        # compute-bound: the kernel issues vector FP32/FP64 instructions (no matrix op on this path);
        # MI355X's published FP32 and FP64 peaks are the same for vector and matrix instructions
        "roofline": {"bound": "mfma", "pipe": "valu", "kernel": dom, "launch_ms": dom_ms, "achieved": achieved,
                     "peak": peak,
                     "unit": "TFLOP/s", "frac": (achieved / peak) if achieved is not None else None,
                     "traffic": traffic, "flops_per_launch": fl.get(dom)},
    }
    # the same workload through host buffers (H2D of gains and INFO rows, D2H of the rows)
    prec_head = dict(modes)[head]
    eng.evaluate(problem, precision=prec_head)
    t0 = time.perf_counter()
    hs = max(1, min(args.steps, 3))
    for _ in range(hs):
        eng.evaluate(problem, precision=prec_head)
    line["host_buffers_value"] = n * hs / (time.perf_counter() - t0)
    if "fp64" in results and head != "fp64":
        line["fp64_value"] = results["fp64"]["value"]
        line["speedup_vs_fp64"] = r["value"] / results["fp64"]["value"]
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(problem, target_s=args.cpu_seconds)
    eng.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
