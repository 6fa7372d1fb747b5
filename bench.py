#!/usr/bin/env python3
"""Serving benchmark of the MI355X interior-point engine.

Workload (one "step"): the serving stream of model-predictive-control
problems -- `--requests` MPC QP/SOCPs (T ~ U{10..50}, nx = 12, nu = 4, one
SOC(3) per step), all submitted at once to a continuous-batching server that
solves them with FP32 factors + FP64 iterative refinement to the 1e-8 gap.
value = solves/s over the timed steps.  Printed beside it: p50/p99 latency,
the same stream under Poisson arrivals at 0.5x / 1x / 1.2x the measured
saturation rate, the plain FP64 single-problem GPU path on a sample (the
baseline of the 5x target), parity of a sample against the CPU oracle, the
roofline of the iteration kernel, and the oracle timed on host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


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


def profiled_traffic(requests, method, engine):
    """HBM bytes per launch of the interior-point kernel from the committed
    FETCH_SIZE/WRITE_SIZE summary of this workload (rocprofv3 runs of
    `bench.py --warmup 0 --skip-extras`), or None for another workload"""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"ipm_kernel_mpc_stream{requests}_hbm_bytes.md")
    if method != "hipmixed" or engine != "auto" or not os.path.exists(path):
        return None
    total = 0.0
    for line in open(path):
        if line.startswith("| FETCH_SIZE |") or line.startswith("| WRITE_SIZE |"):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 3 and cells[2].endswith("GB"):
                total += float(cells[2].split()[0]) * 1e9
    return total or None


def pct(a, p):
    return float(np.percentile(np.asarray(a), p)) if len(a) else float("nan")


def serve(srv, reqs, passes, arrivals=None, log=None):
    """Submit (all at once, or at the given arrival times) and drive until done.
    `reqs` are prepared problems (cuclarabel_amd.prepare).  Returns (wall
    seconds, results in request order)."""
    t0 = time.perf_counter()
    tickets = [None] * len(reqs)
    nxt = 0
    if arrivals is None:
        for i, r in enumerate(reqs):
            tickets[i] = srv.submit(r)
        nxt = len(reqs)
    t_sub = time.perf_counter() - t0
    while True:
        if arrivals is not None:
            now = time.perf_counter() - t0
            while nxt < len(reqs) and arrivals[nxt] <= now:
                tickets[nxt] = srv.submit(reqs[nxt])
                nxt += 1
        inflight = srv.progress(passes)
        if inflight == 0 and nxt >= len(reqs):
            break
        if inflight == 0 and arrivals is not None:
            time.sleep(max(0.0, min(0.002, arrivals[nxt] - (time.perf_counter() - t0))))
    t_run = time.perf_counter() - t0
    res = [srv.result(t) for t in tickets]
    wall = time.perf_counter() - t0
    if log:
        log(f"submit {t_sub:.3f}s, serve loop {t_run - t_sub:.3f}s, results {wall - t_run:.3f}s")
    return wall, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--requests", type=int, default=10000)
    ap.add_argument("--method", default=None, choices=["hip", "hipmixed"],
                    help="default: hipmixed for the served stream, hip for --only-large")
    ap.add_argument("--max-slots", type=int, default=4096)
    ap.add_argument("--engine", default="auto", choices=["auto", "workgroup", "lanes"])
    ap.add_argument("--passes", type=int, default=4)
    ap.add_argument("--hbm-budget-gb", type=float, default=64.0)
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--parity-sample", type=int, default=48)
    ap.add_argument("--rate-requests", type=int, default=1500)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--skip-extras", action="store_true")
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

    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)
    if args.only_large:
        args.method = args.method or "hip"
        if args.only_large == "logistic":
            prob = problems.logistic_exp_pow(3, 5000, 200, 2000)
        elif args.only_large == "psd":
            ref = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                              "large_cases.json")))["psd_relaxation_3000x500"]
            prob = getattr(problems, ref["generator"])(*ref["args"])
        else:
            prob = problems.large_socp(2)
        log(f"large problem {args.only_large}: setup")
        t = time.perf_counter()
        solver = ca.Solver(*prob, ca.Settings(direct_solve_method=args.method))
        t_setup = time.perf_counter() - t
        log("solve")
        t = time.perf_counter()
        sol = solver.solve()
        t_solve = time.perf_counter() - t
        ph = sol.info["phase_seconds"]
        f_rate = sol.info["factor_flops"] * (sol.iterations + 1) / max(ph[3], 1e-9)
        print(json.dumps({"workload": args.only_large, "method": args.method, "status": sol.status,
                          "iterations": sol.iterations, "obj_val": sol.obj_val, "kkt_dim": sol.info["kkt_dim"],
                          "nnz_L": sol.info["nnz_L"], "factor_flops": sol.info["factor_flops"],
                          "setup_s": round(t_setup, 3), "solve_s": round(t_solve, 3),
                          "phase_s": [round(x, 3) for x in ph],
                          "kkt_solves": sol.info["kkt_solves"], "ir_steps": sol.info["ir_steps"],
                          "fp32_factors": sol.info["fp32_factors"], "switched_fp64": sol.info["switched_fp64"],
                          "factorization_tflops": round(f_rate / 1e12, 3)}), flush=True)
        solver.close()  # device memory released before the interpreter (and a profiler) shuts down
        return
    args.method = args.method or "hipmixed"
    t_gen = time.perf_counter()
    raw = list(problems.mpc_stream(5, args.requests))
    sizes = [(r[2].nnz, 2 * r[0].nnz - r[0].shape[0], r[2].shape[0]) for r in raw]  # nnzA, nnz(P full), m
    reqs = [ca.prepare(*r) for r in raw]  # client-side conversion to the ABI's arrays, before the clock
    log(f"generated {len(reqs)} requests in {time.perf_counter() - t_gen:.1f}s")

    st = ca.Settings(direct_solve_method=args.method, serve_engine=args.engine)
    srv = ca.Server(hbm_budget_bytes=int(args.hbm_budget_gb * (1 << 30)), max_slots=args.max_slots, settings=st)
    warm = reqs[: min(len(reqs), 2 * args.max_slots)]
    for w in range(args.warmup):
        wall, _ = serve(srv, warm, args.passes)
        log(f"warmup {w}: {len(warm)} requests in {wall:.2f}s")
    s0 = srv.stats()
    walls, lat, results = [], [], None
    for k in range(args.steps):
        wall, results = serve(srv, reqs, args.passes, log=log)
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
    status_mix = {}
    for r in results:
        status_mix[r.status] = status_mix.get(r.status, 0) + 1
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
        "dtype": "f32 factor + f64 refinement" if args.method == "hipmixed" else "f64",
        "data": "synthetic (seeded MT19937 MPC stream)",
        "config": {
            "workload": f"serving stream of {args.requests} MPC QP/SOCPs, T~U{{10..50}}, nx=12, nu=4, "
                        f"SOC(3) per step, saturating arrivals, tol 1e-8",
            "method": args.method, "max_slots": args.max_slots, "passes_per_launch": args.passes,
            "engine": ("one 64-lane workgroup per batch of 64 same-shape requests" if args.engine == "lanes"
                       else "one 256-thread workgroup per request slot"),
        },
        "latency_p50_s": round(pct(lat, 50), 6),
        "latency_p99_s": round(pct(lat, 99), 6),
        "iterations_mean": round(float(np.mean(iters)), 3),
        "status_mix": status_mix,
        "fp64_switches": switched,
        "fp32_factor_share": round(fp32_share, 4),
        "ir_steps_per_kkt_solve": round(ir_per_solve, 3),
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": profiled_traffic(args.requests, args.method, args.engine),
            "kernel": "clh::lane_kernel" if args.engine == "lanes" else "clh::ipm_kernel", "launches": launches,
            "avg_launch_ms": round(1000 * ktime / max(1, launches), 4),
            "alg_bytes_per_launch": int(bytes_total / max(1, launches)),
        },
    }

    if not args.skip_extras:
        # Poisson arrivals at fractions of the measured saturation rate
        rng = np.random.RandomState(55)
        rates = {}
        sub = reqs[: args.rate_requests]
        for frac in (0.5, 1.0, 1.2):
            lam = frac * value
            arr = np.cumsum(rng.exponential(1.0 / lam, len(sub)))
            wall, res = serve(srv, sub, args.passes, arrivals=arr)
            l = [r.info["latency"] for r in res]
            rates[str(frac)] = {"offered_solves_per_s": round(lam, 1), "achieved_solves_per_s": round(len(sub) / wall, 1),
                                "p50_s": round(pct(l, 50), 6), "p99_s": round(pct(l, 99), 6)}
            log(f"poisson {frac}x: {rates[str(frac)]}")
        out["latency_at_rate"] = rates
        # plain FP64 single-problem GPU path (the baseline of the 5x target)
        pl, plat = [], []
        for r in raw[: args.plain_sample]:
            s = ca.Solver(*r, ca.Settings(direct_solve_method="hip"))
            t = time.perf_counter()
            sol = s.solve()
            plat.append(time.perf_counter() - t)
            pl.append(sol)
        plain = len(plat) / sum(plat)
        out["plain_fp64_gpu"] = {"solves_per_s": round(plain, 2), "p50_s": round(pct(plat, 50), 6),
                                 "p99_s": round(pct(plat, 99), 6), "sample": len(plat)}
        out["served_over_plain_fp64"] = round(value / plain, 2)
        log(f"plain fp64: {plain:.1f} solves/s")
        # parity of served results against the oracle
        from oracle import pyoracle as O
        O.build()
        npar, bad, dit, dobj, dres = 0, 0, 0, 0.0, 0.0
        tcpu = []
        for r, got in zip(raw[: args.parity_sample], results[: args.parity_sample]):
            t = time.perf_counter()
            o = O.solve(*r)
            tcpu.append(time.perf_counter() - t)
            npar += 1
            ok = got.status == o["status"] and abs(got.iterations - o["iterations"]) <= 1
            dit = max(dit, abs(got.iterations - o["iterations"]))
            if o["status"] == "SOLVED":
                e = abs(got.obj_val - o["obj_val"]) / max(1.0, abs(o["obj_val"]))
                dobj = max(dobj, e)
                dres = max(dres, got.r_prim, got.r_dual)
                ok = ok and e <= 1e-7 and got.r_prim <= 1e-8 and got.r_dual <= 1e-8
            bad += 0 if ok else 1
        out["parity"] = {"sample": npar, "mismatches": bad, "max_iter_diff": dit, "max_rel_obj_err": dobj,
                         "max_residual": dres}
        # the oracle (CPU restatement of the reference CPU path) on host cores
        per = float(np.mean(tcpu)) if tcpu else 0.05
        ncpu = int(max(1, min(len(reqs), args.cpu_seconds / max(per, 1e-4))))
        t = time.perf_counter()
        for r in raw[:ncpu]:
            O.solve(*r)
        cpu_rate = ncpu / (time.perf_counter() - t)
        out["cpu_baseline"] = {"value": round(cpu_rate, 2), "unit": "solves/s", "cores": 1, "kind": "port",
                               "sample": f"first {ncpu} requests of the same stream, one core, oracle "
                                         f"(C++ restatement of the Clarabel CPU/QDLDL path)"}
        # one large problem beside the stream: the exp/pow logistic regression
        # at full size (5000 samples x 200 features, 2000 power cones), which
        # the solver runs on the grid engine, and the oracle on one core
        big = problems.logistic_exp_pow(3, 5000, 200, 2000)
        t = time.perf_counter()
        bs = ca.Solver(*big, ca.Settings(direct_solve_method="hip"))
        t_setup = time.perf_counter() - t
        t = time.perf_counter()
        bsol = bs.solve()
        t_solve = time.perf_counter() - t
        t = time.perf_counter()
        bo = O.solve(*big)
        t_cpu = time.perf_counter() - t
        out["large_problem"] = {
            "workload": "exp/pow logistic regression, 5000 samples x 200 features, 15000 exp + 2000 pow cones",
            "kkt_dim": bsol.info["kkt_dim"], "nnz_L": bsol.info["nnz_L"], "factor_flops": bsol.info["factor_flops"],
            "gpu": {"engine": "grid", "status": bsol.status, "iterations": bsol.iterations,
                    "setup_s": round(t_setup, 3), "solve_s": round(t_solve, 3)},
            "oracle_one_core": {"status": bo["status"], "iterations": bo["iterations"], "solve_s": round(t_cpu, 3)},
            "rel_obj_diff": abs(bsol.obj_val - bo["obj_val"]) / max(1.0, abs(bo["obj_val"])),
        }
        log(f"large problem: gpu {t_solve:.2f}s ({bsol.iterations} it), oracle {t_cpu:.2f}s ({bo['iterations']} it)")
        # the full PSD workload (500 PSD cones of side 10, n = 3000, KKT
        # 31 500 rows) on the grid engine, against the oracle's stored result
        # (its single solve takes 1.6 hours on one core, too long to rerun here)
        ref = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                          "large_cases.json")))["psd_relaxation_3000x500"]
        psd = getattr(problems, ref["generator"])(*ref["args"])
        t = time.perf_counter()
        ps = ca.Solver(*psd, ca.Settings(direct_solve_method="hip"))
        t_setup = time.perf_counter() - t
        t = time.perf_counter()
        psol = ps.solve()
        t_solve = time.perf_counter() - t
        out["large_problem_psd"] = {
            "workload": "SDP relaxation, 500 PSD(10) cones + 1000 nonnegative rows, n = 3000",
            "kkt_dim": psol.info["kkt_dim"], "nnz_L": psol.info["nnz_L"], "factor_flops": psol.info["factor_flops"],
            "gpu": {"engine": "grid", "status": psol.status, "iterations": psol.iterations,
                    "setup_s": round(t_setup, 3), "solve_s": round(t_solve, 3),
                    "phase_s": [round(x, 3) for x in psol.info["phase_seconds"]]},
            "oracle_stored": {"status": ref["status"], "iterations": ref["iterations"]},
            "rel_obj_diff": abs(psol.obj_val - ref["obj_val"]) / max(1.0, abs(ref["obj_val"])),
        }
        # the factorization's FP64 rate: one numeric factorization per
        # iteration plus the initial one, over the factorization phase time
        # (device clock); against the FP64 matrix peak
        f_rate = psol.info["factor_flops"] * (psol.iterations + 1) / max(psol.info["phase_seconds"][3], 1e-9)
        out["large_problem_psd"]["factorization"] = {
            "tflops": round(f_rate / 1e12, 3), "peak_tflops": 78.6, "frac": round(f_rate / 78.6e12, 4)}
        log(f"large PSD problem: gpu {t_solve:.2f}s ({psol.iterations} it, {psol.status}), "
            f"oracle stored {ref['iterations']} it {ref['status']}")
        # the full random sparse SOCP (n = 20 000, m = 30 000, KKT 50 000 rows,
        # a nearly dense 18 754-column root: 2.2e12 flops per factorization).
        # The oracle would take hours per factorization, so the result is
        # checked by the KKT conditions evaluated on the host.
        Pc, qc, Ac, bc, kc = problems.large_socp(2)
        t = time.perf_counter()
        cs = ca.Solver(Pc, qc, Ac, bc, kc, ca.Settings(direct_solve_method="hip"))
        t_setup = time.perf_counter() - t
        t = time.perf_counter()
        csol = cs.solve()
        t_solve = time.perf_counter() - t
        f_rate = csol.info["factor_flops"] * (csol.iterations + 1) / max(csol.info["phase_seconds"][3], 1e-9)
        out["large_problem_socp"] = {
            "workload": "random sparse SOCP, n = 20000, m = 30000 (5000 zero, 15000 nonnegative, 2000 SOC(5)), "
                        "8 entries per row",
            "kkt_dim": csol.info["kkt_dim"], "nnz_L": csol.info["nnz_L"], "factor_flops": csol.info["factor_flops"],
            "gpu": {"engine": "grid", "status": csol.status, "iterations": csol.iterations,
                    "setup_s": round(t_setup, 3), "solve_s": round(t_solve, 3),
                    "phase_s": [round(x, 3) for x in csol.info["phase_seconds"]]},
            "kkt_check": {
                "primal": float(np.abs(Ac @ csol.x + csol.s - bc).max() / (1 + np.abs(bc).max())),
                "dual": float(np.abs(Pc @ csol.x + qc + Ac.T @ csol.z).max() / (1 + np.abs(qc).max())),
                "gap": abs(csol.obj_val - csol.obj_val_dual) / max(1.0, abs(csol.obj_val))},
            "factorization": {"tflops": round(f_rate / 1e12, 3), "peak_tflops": 78.6, "frac": round(f_rate / 78.6e12, 4)},
        }
        log(f"large SOCP: gpu {t_solve:.2f}s ({csol.iterations} it, {csol.status})")
        for h in (bs, ps, cs):
            h.close()
    srv.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
