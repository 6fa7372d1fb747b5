#!/usr/bin/env python3
"""Serving benchmark of the MI355X interior-point engine.

Workload (one "step"): the serving stream of model-predictive-control
problems -- `--requests` MPC QP/SOCPs (T ~ U{10..50}, nx = 12, nu = 4, one
SOC(3) per step, seeded MT19937) -- submitted to a continuous-batching server
that solves them with FP32 factors + FP64 iterative refinement to the 1e-8
gap.  Requests are submitted while earlier launches run on the device
(asynchronous progress), so the host's submission overlaps the GPU.
value = solves/s over the timed steps (every status; the rate and latency
of SOLVED requests are given apart, since the stream's definition makes most
horizons infeasible).

Printed beside it: the same stream served in FP64; latency under Poisson
arrivals at 0.5x / 1x / 1.2x the measured saturation rate over the whole
stream; the plain FP64 single-problem GPU path on a sample (the baseline of
the 5x target); the roofline of the iteration kernel; the CPU oracle on the
host's cores (16 worker processes, the GPU box's CPU share) over the whole
stream, whose results double as a parity check of every served request; and
the three large single problems (exp/pow logistic, PSD relaxation, random
sparse SOCP) on the grid engine in FP64 and FP32+IR.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
CPU_WORKERS = 16        # the host CPU share of one GPU on the box


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def alg_bytes(info, nnzA, nnzPf, m):
    """Algorithmic HBM bytes of one solve: (iterations + 1) passes of
    8 nnz(K) + s_f nnz(L) + 3 [2 (s_f + 4) nnz(L) + (1 + k_IR)(12 nnz(K_full) + 24 N)]
      + 12 (2 nnz(A) + nnz(P_full)) + 8 * 10 * m
    with s_f = 4 for FP32 factor passes and 8 for FP64 ones, k_IR the mean
    refinement steps per KKT solve."""
    it = info["iterations"] + 1
    f32 = min(1.0, info["fp32_factors"] / max(1, it))
    sf = 4.0 * f32 + 8.0 * (1.0 - f32)
    kir = info["ir_steps"] / max(1, info["kkt_solves"])
    N = info["kkt_dim"]
    nnzK = info["nnz_K"]
    nnzKf = 2 * nnzK - N
    nnzL = info["nnz_L"]
    per = (8 * nnzK + sf * nnzL + 3 * (2 * (sf + 4) * nnzL + (1 + kir) * (12 * nnzKf + 24 * N))
           + 12 * (2 * nnzA + nnzPf) + 8 * 10 * m)
    return it * per


def profiled_traffic(requests, method, kernel):
    """HBM bytes per launch of the serving kernel from the committed
    FETCH_SIZE/WRITE_SIZE summary of this workload (two rocprofv3 --pmc runs
    of `bench.py --warmup 0 --steps 1 --skip-extras`, scaled by the same
    counters of a device copy), or None for another workload"""
    name = kernel.split("::")[-1]
    path = os.path.join(ROOT, "profiles", f"{name}_mpc_stream{requests}_hbm_bytes.md")
    if method != "hipmixed" or not os.path.exists(path):
        return None, None
    total = 0.0
    for line in open(path):
        if line.startswith("| FETCH_SIZE |") or line.startswith("| WRITE_SIZE |"):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 3 and cells[2].endswith("GB"):
                total += float(cells[2].split()[0]) * 1e9
    return (total or None), os.path.relpath(path, ROOT)


def pct(a, p):
    return float(np.percentile(np.asarray(a), p)) if len(a) else float("nan")


def serve(srv, reqs, passes, arrivals=None, chunk=512):
    """Submit and drive until done: all at once (in chunks, each submitted
    while the previous launch runs) or at the given arrival times (seconds
    from the start).  `reqs` are prepared problems (cuclarabel_amd.prepare).
    Returns (wall seconds, results in request order, submit times)."""
    t0 = time.perf_counter()
    tickets = [None] * len(reqs)
    t_sub = np.zeros(len(reqs))
    nxt = 0
    while True:
        if arrivals is None:
            lim = min(len(reqs), nxt + chunk)
        else:
            lim = int(np.searchsorted(arrivals, time.perf_counter() - t0, side="right"))
        while nxt < lim:
            tickets[nxt] = srv.submit(reqs[nxt])
            t_sub[nxt] = time.perf_counter() - t0
            nxt += 1
        inflight = srv.progress(passes, wait=False)
        if inflight == 0 and nxt >= len(reqs):
            break
        if inflight == 0 and arrivals is not None:
            time.sleep(max(0.0, min(0.002, arrivals[nxt] - (time.perf_counter() - t0))))
    wall = time.perf_counter() - t0
    return wall, [srv.result(t) for t in tickets], t_sub


def split_by_status(results, wall):
    out = {}
    for st in sorted({r.status for r in results}):
        lat = [r.info["latency"] for r in results if r.status == st]
        out[st] = {"count": len(lat), "solves_per_s": round(len(lat) / wall, 2),
                   "p50_s": round(pct(lat, 50), 6), "p99_s": round(pct(lat, 99), 6)}
    return out


# ---- the CPU oracle on the host's cores (checker and baseline) -------------

def _cpu_chunk(args):
    """one worker: the oracle on requests [lo, hi) of the stream"""
    lo, hi, nreq = args
    from cuclarabel_amd import problems
    from oracle import pyoracle as O
    raw = list(problems.mpc_stream(5, nreq))[lo:hi]
    out, t_busy = [], 0.0
    for r in raw:
        t = time.perf_counter()
        o = O.solve(*r)
        t_busy += time.perf_counter() - t
        out.append((o["status"], o["iterations"], o["obj_val"]))
    return out, t_busy


def cpu_baseline(nreq, workers):
    """The oracle (C++ restatement of the reference's CPU/QDLDL path, one core
    per solve) over the whole stream on `workers` processes, forked before
    anything touches the GPU.  Returns (results per request, the baseline
    record)."""
    import multiprocessing as mp
    from oracle import pyoracle as O
    O.build()
    cpu_model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    bounds = np.linspace(0, nreq, workers + 1).astype(int)
    t = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        parts = pool.map(_cpu_chunk, [(int(bounds[i]), int(bounds[i + 1]), nreq) for i in range(workers)])
    wall = time.perf_counter() - t
    res = [x for p, _ in parts for x in p]
    busy = sum(b for _, b in parts)
    rec = {"value": round(nreq / wall, 2), "unit": "solves/s", "cores": workers, "kind": "port",
           "sample": f"all {nreq} requests of the stream, {workers} worker processes (one solve per core at a "
                     f"time; wall time includes each worker generating the stream), oracle = C++ restatement of "
                     f"the Clarabel CPU/QDLDL path",
           "cpu_model": cpu_model, "host_cpus": os.cpu_count(),
           "per_core_solves_per_s": round(nreq / busy, 2), "wall_s": round(wall, 3)}
    return res, rec


def parity(served, cpu_res, raw=None):
    """served results against the oracle's: status, iterations within 1,
    objective to 1e-7 relative and residuals <= 1e-8 when SOLVED.  A request
    whose iteration count alone is off is also held to the oracle's own
    spread: the iteration counts of 16 oracle runs on copies of the request
    whose q and b carry 1e-14 relative perturbations (round-off-level changes
    that any other summation order can make), +-1."""
    bad, dit, dobj, dres = 0, 0, 0.0, 0.0
    off = []
    for idx, (got, (st, it, obj)) in enumerate(zip(served, cpu_res)):
        ok_it = abs(got.iterations - it) <= 1
        ok = got.status == st
        dit = max(dit, abs(got.iterations - it))
        if st == "SOLVED":
            e = abs(got.obj_val - obj) / max(1.0, abs(obj))
            dobj = max(dobj, e)
            dres = max(dres, got.r_prim, got.r_dual)
            ok = ok and e <= 1e-7 and got.r_prim <= 1e-8 and got.r_dual <= 1e-8
        if ok and not ok_it:
            off.append((idx, got.iterations, it))
        bad += 0 if (ok and ok_it) else 1
    rec = {"sample": len(cpu_res), "mismatches": bad, "max_iter_diff": dit, "max_rel_obj_err": dobj,
           "max_residual": dres}
    if off and raw is not None:
        from oracle import pyoracle as O
        within = []
        for idx, got_it, it in off[:20]:
            P, q, A, b, cones = raw[idx]
            rng = np.random.RandomState(idx)
            its = [O.solve(P, q * (1 + 1e-14 * rng.standard_normal(q.size)), A,
                           b * (1 + 1e-14 * rng.standard_normal(b.size)), cones)["iterations"] for _ in range(16)]
            within.append({"request": idx, "iterations": got_it, "oracle": it, "oracle_envelope": [min(its), max(its)],
                           "within": min(its) - 1 <= got_it <= max(its) + 1})
        rec["iteration_outliers"] = within
        rec["mismatches_outside_oracle_envelope"] = bad - sum(w["within"] for w in within)
    return rec


# ---- the large single problems ---------------------------------------------

def large_problems(ca, problems, which=("logistic", "psd", "socp"), methods=("hip", "hipmixed"), autotune_seconds=30.0):
    """The three full-size problems of the other workloads on the grid engine:
    time, iterations, status, factorization rate; each in FP64 and FP32+IR.
    Their checks: the oracle's stored results (logistic, PSD) or the KKT
    conditions evaluated here (random SOCP, beyond the oracle's reach)."""
    # the oracle's stored results, under the engine's dynamic-regularization
    # rule where those are stored (the "engine_rule" record)
    golden = {k: {**v, **v.get("engine_rule", {})}
              for k, v in json.load(open(os.path.join(ROOT, "tests", "golden", "large_cases.json"))).items()}
    specs = {
        "logistic": ("exp/pow logistic regression, 5000 samples x 200 features, 15000 exp + 2000 pow cones",
                     golden["logistic_5000x200"]),
        "psd": ("SDP relaxation, 500 PSD(10) cones + 1000 nonnegative rows, n = 3000",
                golden["psd_relaxation_3000x500"]),
        "socp": ("random sparse SOCP, n = 20000, m = 30000 (5000 zero, 15000 nonnegative, 2000 SOC(5)), "
                 "8 entries per row", None),
    }
    out = {}
    for name in which:
        desc, ref = specs[name]
        prob = getattr(problems, ref["generator"])(*ref["args"]) if ref else problems.large_socp(2)
        rec = {"workload": desc}
        for method in methods:
            log(f"large {name} {method}")
            t = time.perf_counter()
            s = ca.Solver(*prob, ca.Settings(direct_solve_method=method, verbose=autotune_seconds > 0))
            t_setup = time.perf_counter() - t
            tune = s.autotune(max_seconds=autotune_seconds) if autotune_seconds > 0 else None
            if tune:
                tune["schedule"] = {k: v for k, v in vars(s.schedule()).items()
                                    if k in ("grid_blocks", "tile_min_cols", "dense_min_cols")}
                log(f"  autotune: {tune}")
            t = time.perf_counter()
            sol = s.solve()
            t_solve = time.perf_counter() - t
            s.close()
            info = sol.info
            ph = info["phase_seconds"]
            f_rate = info["factor_flops"] * (sol.iterations + 1) / max(ph[3], 1e-9)
            peak = FP64_PEAK_TFLOPS if method == "hip" else FP32_PEAK_TFLOPS
            r = {"status": sol.status, "iterations": sol.iterations, "setup_s": round(t_setup, 3), "autotune": tune,
                 "solve_s": round(t_solve, 3), "phase_s": [round(x, 3) for x in ph],
                 "kkt_dim": info["kkt_dim"], "nnz_L": info["nnz_L"], "factor_flops": info["factor_flops"],
                 "fp32_factors": info["fp32_factors"], "kkt_solves": info["kkt_solves"], "ir_steps": info["ir_steps"],
                 "factorization": {"tflops": round(f_rate / 1e12, 3), "peak_tflops": peak,
                                   "frac": round(f_rate / (peak * 1e12), 4)}}
            if ref:
                r["oracle_stored"] = {"status": ref["status"], "iterations": ref["iterations"],
                                      "iterations_envelope": ref.get("iterations_envelope")}
                r["rel_obj_err_vs_oracle"] = abs(sol.obj_val - ref["obj_val"]) / max(1.0, abs(ref["obj_val"]))
            else:
                P, q, A, b, _ = prob
                r["kkt_check"] = {
                    "primal": float(np.abs(A @ sol.x + sol.s - b).max() / (1 + np.abs(b).max())),
                    "dual": float(np.abs(P @ sol.x + q + A.T @ sol.z).max() / (1 + np.abs(q).max())),
                    "gap": abs(sol.obj_val - sol.obj_val_dual) / max(1.0, abs(sol.obj_val))}
            rec[method] = r
            log(f"  {sol.status} {sol.iterations} it, setup {t_setup:.2f}s, solve {t_solve:.2f}s, "
                f"factorization {f_rate / 1e12:.2f} TFLOP/s")
        out[name] = rec
    return out


def tuned_server(ca, args, method, reqs, tune_reqs):
    """A server for the stream; the autotuner times candidate schedules on
    `tune_reqs` (another seed's stream of the same distribution, queued, then
    served under the chosen schedule), before any clock starts.  Returns
    (server, tuning record)."""
    # verbose: the tuner prints a line per trial (nothing else reads it here)
    st = ca.Settings(direct_solve_method=method, serve_engine=args.engine, verbose=args.autotune_seconds > 0)
    srv = ca.Server(hbm_budget_bytes=int(args.hbm_budget_gb * (1 << 30)), max_slots=args.max_slots, settings=st)
    if args.autotune_seconds <= 0:
        return srv, None
    tickets = [srv.submit(r) for r in tune_reqs]
    tune = srv.autotune(max_seconds=args.autotune_seconds)
    while srv.progress(args.passes, wait=False):
        pass
    for t in tickets:
        srv.result(t)
    classes = sorted({int(np.floor(np.log2(r.n + r.m))) for r in reqs})
    tune["schedules"] = {f"n+m in [{1 << c}, {2 << c})": {
        k: v for k, v in vars(srv.schedule(1 << c)).items() if k != "use_tuned"} for c in classes}
    log(f"autotune {method}: {tune}")
    return srv, tune


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--requests", type=int, default=10000)
    ap.add_argument("--method", default=None, choices=["hip", "hipmixed"],
                    help="default: hipmixed for the served stream, hip for --only-large")
    ap.add_argument("--max-slots", type=int, default=4096)
    ap.add_argument("--engine", default="auto", choices=["auto", "workgroup", "lanes", "resident"])
    ap.add_argument("--passes", type=int, default=4)
    ap.add_argument("--hbm-budget-gb", type=float, default=64.0)
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--autotune-seconds", type=float, default=0.0,
                    help="time given to the server's autotuner before the warmup, on another seed's stream of the "
                         "same distribution (0, the default: the built-in schedule)")
    ap.add_argument("--large-autotune-seconds", type=float, default=30.0,
                    help="time given to the grid engine's autotuner on each large problem (0: none)")
    ap.add_argument("--skip-extras", action="store_true", help="only the timed serving steps")
    ap.add_argument("--skip-large", action="store_true", help="leave out the three large single problems")
    ap.add_argument("--only-large", default=None, choices=["logistic", "psd", "socp"],
                    help="only solve one full-size single problem on the grid engine and print its figures "
                         "(logistic: exp/pow logistic 5000 x 200; psd: 500 PSD(10) cones, n = 3000; socp: random "
                         "sparse SOCP, n = 20000, m = 30000)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a device-to-device copy of 256 MiB (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the serving path would shard by request, one "
              "independent server per GPU); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda:0").fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel()}), flush=True)
        return

    import cuclarabel_amd as ca
    from cuclarabel_amd import problems

    if args.only_large:
        res = large_problems(ca, problems, (args.only_large,), (args.method or "hip",),
                             autotune_seconds=args.large_autotune_seconds)
        print(json.dumps(res[args.only_large]), flush=True)
        return

    args.method = args.method or "hipmixed"
    # the CPU leg first: its worker processes are forked before the GPU is touched
    cpu_res, cpu_rec = None, None
    if not args.skip_extras:
        workers = min(CPU_WORKERS, os.cpu_count() or 1)
        log(f"cpu baseline: oracle on {args.requests} requests, {workers} workers")
        cpu_res, cpu_rec = cpu_baseline(args.requests, workers)
        log(f"cpu baseline: {cpu_rec['value']} solves/s")

    t_gen = time.perf_counter()
    raw = list(problems.mpc_stream(5, args.requests))
    sizes = [(r[2].nnz, 2 * r[0].nnz - r[0].shape[0], r[2].shape[0]) for r in raw]  # nnzA, nnz(P full), m
    reqs = [ca.prepare(*r) for r in raw]  # client-side conversion to the ABI's arrays, before the clock
    log(f"generated {len(reqs)} requests in {time.perf_counter() - t_gen:.1f}s")

    tune_reqs = []
    if args.autotune_seconds > 0:  # the tuning sample: the same distribution, another seed
        tune_reqs = [ca.prepare(*r) for r in problems.mpc_stream(6, args.requests)]
    srv, tune = tuned_server(ca, args, args.method, reqs, tune_reqs)
    for w in range(args.warmup):
        wall, _, _ = serve(srv, reqs, args.passes)
        log(f"warmup {w}: {len(reqs)} requests in {wall:.2f}s")
    s0 = srv.stats()
    walls, lat, results = [], [], None
    for k in range(args.steps):
        wall, results, _ = serve(srv, reqs, args.passes)
        walls.append(wall)
        lat += [r.info["latency"] for r in results]
        log(f"step {k}: {len(reqs)} requests in {wall:.3f}s ({len(reqs) / wall:.0f} solves/s)")
    s1 = srv.stats()
    total = sum(walls)
    value = args.requests * args.steps / total
    launches = s1["launches"] - s0["launches"]
    ktime = s1["kernel_time_s"] - s0["kernel_time_s"]
    bytes_total = args.steps * sum(alg_bytes(r.info, *sz) for r, sz in zip(results, sizes))
    achieved = bytes_total / ktime / 1e9 if ktime > 0 else 0.0
    kernel = {"lanes": "clh::lane_kernel", "workgroup": "clh::ipm_kernel"}.get(args.engine, "clh::resident_kernel")
    traffic, traffic_src = profiled_traffic(args.requests, args.method, kernel)
    iters = [r.iterations for r in results]
    switched = sum(r.info["switched_fp64"] for r in results)
    fp32_share = sum(r.info["fp32_factors"] for r in results) / max(1, sum(r.iterations + 1 for r in results))
    ir_per_solve = sum(r.info["ir_steps"] for r in results) / max(1, sum(r.info["kkt_solves"] for r in results))

    out = {
        "metric": "solves/s + p50/p99 latency at 1e-8 gap: served FP32+IR vs plain FP64 GPU",
        "value": round(value, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # :hipmixed starts each request on FP32 factors and moves it to FP64
        # factors when refinement stalls; fp32_factor_share says how much of
        # the factorization work stayed FP32
        "dtype": "f32 factor + f64 refinement, f64 factors after a stall" if args.method == "hipmixed" else "f64",
        "data": "synthetic (seeded MT19937 MPC stream)",
        "config": {
            "workload": f"serving stream of {args.requests} MPC QP/SOCPs, T~U{{10..50}}, nx=12, nu=4, "
                        f"SOC(3) per step, all submitted at the start of a step (saturating), tol 1e-8",
            "method": args.method, "max_slots": args.max_slots, "passes_per_launch": args.passes,
            "engine": {"lanes": "one 64-lane workgroup per batch of 64 same-shape requests",
                       "workgroup": "one 256-thread workgroup per request slot, rounds of graph-replayed launches"}.get(
                args.engine, "device-resident request loop: one persistent launch of 256-thread workgroups (4 per "
                             "CU) taking queued requests from host memory, each solved to completion by one workgroup"),
            "status_mix_note": "x_init ~ N(0,1) against state boxes +-U(1,5) with A = I + 0.05 N(0,1): most "
                               "horizons are infeasible; every status is counted in value",
        },
        "autotune": tune,
        "latency_p50_s": round(pct(lat, 50), 6),
        "latency_p99_s": round(pct(lat, 99), 6),
        "by_status": split_by_status(results, total / args.steps),
        "iterations_mean": round(float(np.mean(iters)), 3),
        "fp64_switches": switched,
        "fp32_factor_share": round(fp32_share, 4),
        "ir_steps_per_kkt_solve": round(ir_per_solve, 3),
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": traffic,
            "traffic_source": (f"stored: {traffic_src}" if traffic_src else None),
            "kernel": kernel,
            "launches": launches,
            "avg_launch_ms": round(1000 * ktime / max(1, launches), 4),
            "alg_bytes_per_launch": int(bytes_total / max(1, launches)),
        },
    }

    if not args.skip_extras:
        out["parity"] = parity(results, cpu_res, raw)
        log(f"parity: {out['parity']}")
        out["cpu_baseline"] = cpu_rec
        # Poisson arrivals at fractions of the measured saturation rate, over
        # the whole stream: achieved rate over the middle 80% of completions
        rng = np.random.RandomState(55)
        rates = {}
        for frac in (0.5, 1.0, 1.2):
            lam = frac * value
            arr = np.cumsum(rng.exponential(1.0 / lam, len(reqs)))
            wall, res, t_sub = serve(srv, reqs, args.passes, arrivals=arr)
            l = np.array([r.info["latency"] for r in res])
            done_t = np.sort(t_sub + l)
            a, b = done_t[len(done_t) // 10], done_t[(9 * len(done_t)) // 10]
            solved = [x for x, r in zip(l, res) if r.status == "SOLVED"]
            rates[str(frac)] = {"offered_solves_per_s": round(lam, 1),
                                "achieved_solves_per_s": round(len(reqs) / wall, 1),
                                "steady_state_solves_per_s": round(float(0.8 * len(reqs) / max(b - a, 1e-9)), 1),
                                "p50_s": round(pct(l, 50), 6), "p99_s": round(pct(l, 99), 6),
                                "solved_p50_s": round(pct(solved, 50), 6), "solved_p99_s": round(pct(solved, 99), 6)}
            log(f"poisson {frac}x: {rates[str(frac)]}")
        out["latency_at_rate"] = rates
        # the same stream served with FP64 factors
        srv64, tune64 = tuned_server(ca, args, "hip", reqs, tune_reqs)
        serve(srv64, reqs, args.passes)
        wall64, res64, _ = serve(srv64, reqs, args.passes)
        l64 = [r.info["latency"] for r in res64]
        out["served_fp64"] = {"autotune": tune64,
                              "solves_per_s": round(len(reqs) / wall64, 2), "p50_s": round(pct(l64, 50), 6),
                              "p99_s": round(pct(l64, 99), 6), "parity": parity(res64, cpu_res, raw)}
        srv64.close()
        log(f"served fp64: {out['served_fp64']['solves_per_s']} solves/s")
        # plain FP64 single-problem GPU path (the baseline of the 5x target)
        plat = []
        for r in raw[: args.plain_sample]:
            s = ca.Solver(*r, ca.Settings(direct_solve_method="hip"))
            t = time.perf_counter()
            s.solve()
            plat.append(time.perf_counter() - t)
            s.close()
        plain = len(plat) / sum(plat)
        out["plain_fp64_gpu"] = {"solves_per_s": round(plain, 2), "p50_s": round(pct(plat, 50), 6),
                                 "p99_s": round(pct(plat, 99), 6), "sample": len(plat)}
        out["served_over_plain_fp64"] = round(value / plain, 2)
        log(f"plain fp64: {plain:.1f} solves/s")
        if not args.skip_large:
            out["large_problems"] = large_problems(ca, problems, autotune_seconds=args.large_autotune_seconds)
    srv.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
