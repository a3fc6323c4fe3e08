"""Benchmark: candidate evaluations per second of the full six-stage chain on one MI355X.

One step = one ce_evaluate() call over the whole candidate block of the workload (inputs already
resident in HBM): simulation, linearised eigensolve, modal fit, margin sweep and scoring of every
candidate.  The default workload is the metric's comparison, mixed precision against the FP64
chain at matched margins, on the 8-state plant (262,144 candidates, T = 4000, F = 512, 4 NL).

Printed: one JSON line with
  value             mixed candidate-evals/s (the headline), fp64_value beside it;
  matched_margins   max |dGM|, |dPM|, |db|, overshoot / wn / yss errors, pass-flag mismatches,
                    pass-ranking identity, score-rank Spearman, mixed vs FP64 rows of the same run;
  roofline          the dominant kernel's algorithmic rate (HIP events on its launch stream);
  cpu_baseline      the CPU oracle (FP64 restatement) on the host cores, all threads and one core;
  c4, c5            sub-objects: the batched uneven Monte Carlo (512 plants) and the serving loop
                    (p50 / p99 job latency), each in both precisions (skip with --no-legs).
--config c1|c2|c3 selects the block workload; --config c4|c5 runs only that leg as the line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3     # published (packed FP32); this build issues non-packed FP32 (78.6)
HBM_PEAK_GBS = 8000.0
METRIC = "candidate evals/s, mixed vs FP64 plain at matched margins; p50/p99 job latency"

F32_KERNELS = ("ce_sim_f32", "ce_sim_f32_g", "ce_sims_f32", "ce_sims_f32_g", "k_sim_f32", "k_sweep_mark32")
WORKLOADS = {"c1": "mass-spring-damper 2-state", "c2": "two-mass drive 4-state, 2 NL",
             "c3": "random stable 8-state, 4 NL, Kalman observer"}


def kernel_flops(problem, info64, d, mixed, sim_slots=None):
    """Algorithmic flops per launch of each stage kernel, counted for the candidates that reach it
    (FMA = 2, add/mul/div = 1, each non-SKIP instruction = 1; transcendentals not counted).  The
    float simulation counts the compensated state of the reference (lo-part dot products and the
    11-operation TwoSum / Fast2Sum update of every state).  sim_slots: the candidates the simulation
    kernel itself runs (the one-lane kernel of a split simulation), else every simulated one."""
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
    if mixed:
        per_step += 4 * ns + 2 * ns * ns + 2 * nd * nd + 11 * (ns + nd)
    sim = float(per_step) * T * (n_sim if sim_slots is None else min(n_sim, sim_slots))
    solve = float((2 * nt) ** 3 / 3 + 4 * (2 * nt) ** 2) * int(fitted.sum())
    per_point = 8.0 / 3.0 * ns ** 3 + 8 * ns * ns + 8 * ns + 30
    swept = int((info64[:, d.info_allmargin + 3] == 1).sum())
    return {
        "ce_sim_f32": sim, "ce_sim_f32_g": sim, "ce_sim_f64": sim, "ce_sim_f64_g": sim, "k_sim_f32": sim,
        "ce_sims_f32": sim, "ce_sims_f64": sim, "ce_sims_f32_g": sim, "ce_sims_f64_g": sim,
        "k_sim_f64": sim,
        "k_eig_lin": float(10 * nt ** 3 + 4 * ns * ns + 2 * ns * nd + 2 * ns) * n,
        "k_eig_nom": float(10 * (2 * ns) ** 3) * n_sim,
        "k_atb": float(np.sum(8.0 * u * tfit)),
        "k_fitcf": solve + float(np.sum(u * (u + 1) / 2)) * 2 * 40,
        "k_fit_fp64": float(np.sum(tfit * (nc * (nc + 1) + 2 * nc + 4 * u + 1) + 2 * T)) + solve,
        "k_sweep_mark32": F * per_point * swept, "k_sweep_fp64": F * per_point * swept,
    }


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail


DRIFT_KEYS = ("max_dgm_db", "max_dpm_deg", "max_db", "max_dovershoot", "max_rel_dwn", "max_rel_dyss", "n_margin_rows",
              "passfail_mismatch", "pass_rank_identical", "n_pass", "margin_pass_mismatch", "n_margin_pass",
              "margin_pass_rank_identical", "margin_pass_rank_swaps", "stable_mismatch", "behaviour_rel_err_median",
              "behaviour_rel_err_p99", "score_rank_spearman")


def cpu_baseline(problem, target_s=12.0, one_core_s=4.0, rows64=None, rows_mixed=None, drift_n=4096):
    """The oracle's restatements of the reference chain (OpenMP over candidates) on a bounded sample
    of the workload, on every host thread the process may use and on one thread: the plain FP64
    chain and the ref-FP32 chain (the reference's own precision map).  With the GPU's FP64 and
    mixed rows given, the ref-FP32 rows of the first drift_n candidates are compared with the FP64
    rows beside the mixed rows of the same candidates: how far the reference's own arithmetic lies
    from FP64, next to how far the mixed chain does."""
    import numpy as np
    from oracle import oracle
    model, avail = host_info()
    threads = int(os.environ.get("CE_CPU_THREADS", "0")) or min(avail, oracle.lib().ora_max_threads())
    n_total = problem.gains.shape[0]

    def timed(fn, nthreads, budget):
        n = min(256, n_total)
        while True:
            sub = problem.subset(np.arange(n))
            t0 = time.perf_counter()
            fn(sub, threads=nthreads)
            t = time.perf_counter() - t0
            if t >= budget / 3 or n >= n_total:
                break
            n = min(n_total, int(n * max(2.0, min(8.0, budget / 3 / max(t, 1e-3)))))
        if t < budget / 3:  # whole workload is quicker than the budget: repeat it
            reps = max(1, int(budget / max(t, 1e-3)))
            t0 = time.perf_counter()
            for _ in range(reps):
                fn(sub, threads=nthreads)
            t = time.perf_counter() - t0
            return reps * n / t, "%d x %d candidates, %.1f s" % (reps, n, t)
        return n / t, "first %d of %d candidates, %.1f s" % (n, n_total, t)

    v_all, s_all = timed(oracle.evaluate, threads, target_s)
    v_one, s_one = timed(oracle.evaluate, 1, one_core_s)
    r_all, rs_all = timed(oracle.evaluate_ref32, threads, target_s / 2)
    r_one, rs_one = timed(oracle.evaluate_ref32, 1, one_core_s / 2)
    out = dict(value=v_all, unit="candidate-evals/s", cores=threads, kind="port",
               sample="FP64 oracle, all six stages: " + s_all, one_core_value=v_one, one_core_sample=s_one,
               ref32_value=r_all, ref32_sample="ref-FP32 oracle (the reference's precision map): " + rs_all,
               ref32_one_core_value=r_one, ref32_one_core_sample=rs_one,
               host_cpu=model, host_threads_available=avail)
    if rows64 is not None:
        from controllerengine_amd import metrics
        k = min(drift_n, n_total)
        sub = problem.subset(np.arange(k))
        r32, _ = oracle.evaluate_ref32(sub, threads=threads)
        drift = {"n": k, "ref32_vs_fp64": metrics.compare(sub, rows64[:k], r32.astype(np.float64))}
        if rows_mixed is not None:
            drift["mixed_vs_fp64"] = metrics.compare(sub, rows64[:k], rows_mixed[:k])
        out["precision_drift_sample"] = {
            key: ({kk: v[kk] for kk in DRIFT_KEYS if kk in v} if isinstance(v, dict) else v) for key, v in drift.items()}
    return out


def bench_c4(args, steps=None, fp64_steps=None, gen_rounds=None):
    """Batched uneven Monte Carlo: all plants of a round in one batched chain (grouped persistent
    launches, hipGraph replay, autotuned grids); every round uploads the perturbed plants.  The
    timed rounds cycle through gen_rounds generated perturbation sets (default: one per round)."""
    import torch
    from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, workloads
    from controllerengine_amd.batch import Batch
    rounds_m = steps or args.rounds
    rounds_d = fp64_steps or max(1, min(args.rounds, args.fp64_rounds))
    t0 = time.perf_counter()
    w = workloads.C4(n_plants=args.plants, rounds=gen_rounds or max(rounds_m, rounds_d, 1))
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
        rounds = rounds_m if name == "mixed" else rounds_d
        b.set_options(use_graph=True, autotune=True)
        for _ in range(max(1, args.warmup if steps is None else 1)):
            b.run_device(prec, d_g, d_i, d_o)
        torch.cuda.synchronize()
        b.set_options(use_graph=False, autotune=False, profiling=True)
        b.run_device(prec, d_g, d_i, d_o)
        torch.cuda.synchronize()
        kt = b.kernel_times()
        # the per-kernel pass ran without the graph (which dropped it): one untimed round captures it
        # again, so the timed rounds are graph replays only
        b.set_options(use_graph=True, autotune=False, profiling=False)
        b.run_device(prec, d_g, d_i, d_o)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in range(rounds):
            for j, p in enumerate(w.plants):
                b.update(j, p, rnd=r % w.rounds)
            b.run_device(prec, d_g, d_i, d_o)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        info = b.info()
        res[name] = dict(value=n * rounds / dt, ms_per_round=1000 * dt / rounds, rounds=rounds,
                         used_graph=info["used_graph"], launches=info["launches"], plan=b.plan(),
                         kernel_ms={k: round(v, 3) for k, v in kt})
    b.close()
    out = dict(workload="c4: %d plants of 2-8 states, %d candidates per round (log-uniform 256-16384 per plant), "
                        "T 1000-8000, F=400, serial rounds with +-10%% perturbations, one batch" % (len(w.plants), n),
               generator_s=gen_s, perturbation_sets=w.rounds, **{k: v for k, v in res.items()})
    if "fp64" in res and "mixed" in res:
        out["speedup_vs_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
    return out


def bench_c5(args, rates=None, duration=None, pool_n=None):
    """Serving loop: Poisson arrivals of c5 jobs at each rate for `duration` seconds, submitted to a
    ce_server (footprint admission, continuous batching); p50 / p99 latency and throughput."""
    import numpy as np
    from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, workloads
    from controllerengine_amd.server import Server
    rng = np.random.default_rng(4)
    pool = [workloads.c5_job(rng) for _ in range(pool_n or args.pool)]
    rates = rates or [float(r) for r in args.rates.split(",")]
    duration = duration or args.duration
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
                if t_next > duration:
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
                                             evals_per_s=st["evals_per_s"], offered_evals_per_s=cands / duration,
                                             batches=int(st["batches"]),
                                             peak_jobs_in_flight=int(st["peak_jobs_in_flight"]),
                                             peak_batches_in_flight=int(st["peak_batches_in_flight"]))
            print("c5 rate %g %s: %s" % (rate, name, out[name][str(int(rate))]), file=sys.stderr, flush=True)
        srv.close()
    return dict(workload="c5: Poisson job arrivals at %s jobs/s for %g s each; jobs: random stable plants of 2-8 "
                         "states, 0-4 NL, 256-65536 candidates, T 500-5000, F 128-512 (pool of %d)"
                         % (",".join("%g" % r for r in rates), duration, len(pool)), **out)


def bench_block(args):
    """The c1 / c2 / c3 candidate block: mixed and FP64 chains, matched-margin comparison of the
    rows of the same run, per-kernel event times and the roofline of the dominant kernel."""
    import torch
    from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, metrics, workloads
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
    modes = {"fp64": [("fp64", PRECISION_FP64)], "mixed": [("mixed", PRECISION_MIXED)]}.get(
        args.precision, [("mixed", PRECISION_MIXED), ("fp64", PRECISION_FP64)])
    results = {}
    for name, prec in modes:
        for _ in range(args.warmup):
            eng.evaluate_device(cfg, plant, prec, g, i_in, i_out, i64)
        # the simulation kernel of every timed step, between HIP events on its launch stream
        eng.set_sim_history(args.steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            eng.evaluate_device(cfg, plant, prec, g, i_in, i_out, i64)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        sim_hist = eng.sim_history()
        sim_kernel = eng.last_sim_kernel()
        split_cap = eng.last_sim_split()
        eng.set_sim_history(0)
        info64 = i64.cpu().numpy()
        # per-kernel durations from HIP events recorded on the launch stream, separate passes
        eng.set_profiling(True)
        acc = {}
        for _ in range(args.prof_steps):
            eng.evaluate_device(cfg, plant, prec, g, i_in, i_out, i64)
            for k, ms in eng.kernel_times():
                acc.setdefault(k, []).append(ms)
        eng.set_profiling(False)
        kern = {k: sum(v) / len(v) for k, v in acc.items() if k != "start"}
        results[name] = dict(value=n * args.steps / dt, ms=1000 * dt / args.steps, kernels=kern, info64=info64,
                             sim_ms=sum(sim_hist) / len(sim_hist) if sim_hist else None, sim_kernel=sim_kernel,
                             split_cap=split_cap)
        print("%s %s: %.4g candidate-evals/s, %.3f ms per step" % (args.config, name, results[name]["value"],
                                                                   results[name]["ms"]), file=sys.stderr, flush=True)
    head = modes[0][0]
    r = results[head]
    # split simulation: the one-lane kernel (the roofline's) runs the list slots below the split,
    # a 2-lane kernel the rest on another stream
    from controllerengine_amd.engine import sim_split_slots
    sc = r["info64"][:, d.info_scoring:d.info_scoring + 3]
    n_sim = int((~((sc[:, 0] >= float(problem.config["unstable_err"])) &
                   (r["info64"][:, d.info_allmargin + 3] == 0))).sum())
    sim_slots = sim_split_slots(n_sim, r["split_cap"]) if r["split_cap"] else None
    fl = kernel_flops(problem, r["info64"], d, head == "mixed", sim_slots)
    kern = r["kernels"]
    dom = max(kern, key=lambda k: kern[k]) if kern else None
    dom_ms = kern.get(dom)
    # the per-kernel passes run one stream; the timed steps overlap the simulation with the
    # second stream's work, so the simulation kernel is timed over the timed steps themselves
    ms_source = "per-kernel passes (one stream)"
    if dom is not None and dom.startswith(("ce_sim", "k_sim")) and r["sim_ms"] is not None:
        # the timed steps' form of the simulation (the gated form, where the chain starts it early)
        dom, dom_ms = r["sim_kernel"], r["sim_ms"]
        ms_source = "timed steps (HIP events on the launch stream)"
    peak = FP32_PEAK_TFLOPS if dom in F32_KERNELS else FP64_PEAK_TFLOPS
    achieved = fl[dom] / (dom_ms * 1e-3) / 1e12 if dom in fl else None
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_%s_%s_%s.json" % (dom, args.config, head))
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    line = {
        "metric": METRIC, "value": r["value"], "unit": "candidate-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if head == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": "%s: %s, %d candidates, T=%d, F=%d" % (args.config, WORKLOADS[args.config], n,
                                                                       cfg.time, cfg.freq),
                   "precision": head},
        "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
        # vector FP32 / FP64 instructions only (no matrix op on this path): the compute roofline
        # is the VALU peak of the kernel's precision
        "roofline": {"bound": "valu", "kernel": dom, "launch_ms": dom_ms, "launch_ms_source": ms_source,
                     "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": (achieved / peak) if achieved is not None else None,
                     "traffic": traffic, "flops_per_launch": fl.get(dom),
                     "simulated": n_sim, "kernel_candidates": sim_slots if sim_slots is not None else n_sim},
    }
    if "fp64" in results and head != "fp64":
        line["fp64_value"] = results["fp64"]["value"]
        line["fp64_ms_per_step"] = results["fp64"]["ms"]
        line["speedup_vs_fp64"] = r["value"] / results["fp64"]["value"]
        line["fp64_kernel_ms"] = {k: round(v, 4) for k, v in results["fp64"]["kernels"].items()}
        line["matched_margins"] = dict(
            tolerance={"dgm_db": 0.01, "dpm_deg": 0.01, "db": 1e-4},
            **metrics.compare(problem, results["fp64"]["info64"], r["info64"]))
    # the same workload through host buffers (H2D of gains and INFO rows, D2H of the rows)
    prec_head = dict(modes)[head]
    eng.evaluate(problem, precision=prec_head)
    t0 = time.perf_counter()
    hs = max(1, min(args.steps, 3))
    for _ in range(hs):
        eng.evaluate(problem, precision=prec_head)
    line["host_buffers_value"] = n * hs / (time.perf_counter() - t0)
    eng.close()
    if not args.no_cpu:
        rows64 = results["fp64"]["info64"] if "fp64" in results else None
        rows_mixed = results["mixed"]["info64"] if "mixed" in results else None
        line["cpu_baseline"] = cpu_baseline(problem, target_s=args.cpu_seconds, rows64=rows64, rows_mixed=rows_mixed)
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--n", type=int, default=0, help="candidates (default: the configuration's count)")
    ap.add_argument("--precision", default="auto", choices=["auto", "fp64", "mixed"])
    ap.add_argument("--prof-steps", type=int, default=3, help="extra steps with per-kernel events")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="block workload only (no c4 / c5 sub-objects)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
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

    if args.config in ("c4", "c5"):
        leg = bench_c4(args) if args.config == "c4" else bench_c5(args)
        head = "mixed" if "mixed" in leg else "fp64"
        if args.config == "c4":
            v, steps, ms = leg[head]["value"], leg[head]["rounds"], leg[head]["ms_per_round"]
        else:
            top = max(leg[head], key=lambda k: float(k))
            v, steps, ms = leg[head][top]["evals_per_s"], leg[head][top]["jobs"], leg[head][top]["mean_ms"]
        line = {"metric": METRIC, "value": v, "unit": "candidate-evals/s", "n_gpus": 1, "steps": steps,
                "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f64" if head == "fp64" else "f32+f64", "data": "synthetic",
                "config": {"workload": leg["workload"], "precision": head}, args.config: leg}
        if args.config == "c5":
            line["p50_ms"], line["p99_ms"] = leg[head][top]["p50_ms"], leg[head][top]["p99_ms"]
        print(json.dumps(line))
        return

    line = bench_block(args)
    if not args.no_legs and args.config == "c3":
        # short legs of the batched and serving forms, so that one run carries every figure
        # 16 timed rounds over the 4 perturbation sets of the 4-round workload: the first round's
        # host work (plant updates, descriptor build) has nothing to overlap with, and over 4 rounds
        # it alone added 5-6 ms to each
        line["c4"] = bench_c4(args, steps=16, fp64_steps=2, gen_rounds=4)
        c5 = bench_c5(args, rates=[100.0], duration=5.0, pool_n=16)
        line["c5"] = c5
        head = line["config"]["precision"]
        if head in c5:
            line["p50_ms"], line["p99_ms"] = c5[head]["100"]["p50_ms"], c5[head]["100"]["p99_ms"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
