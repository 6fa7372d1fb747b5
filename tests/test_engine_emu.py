"""The device engine's logic compiled serially for the CPU (one thread per
workgroup) against the oracle.  This is a unit test of the engine source only;
the GPU parity tests (test_gpu_*.py) run the HIP build."""
import numpy as np
import pytest

import cuclarabel_amd as ca
import deadline_cases
from cuclarabel_amd import problems

from helpers import (LOGISTIC_CASES, SDP_CASES, SOCP_CASES, SPARSE_SOC_CASES, assert_parity, build_emu, case_problem, data_update_case, apply_updates,
                     engine_cases, socp_case)


@pytest.fixture(scope="module")
def emu():
    return build_emu()


@pytest.mark.parametrize("case", engine_cases(), ids=lambda c: c["name"])
def test_reference_cases(emu, oracle, case):
    P, q, A, b, cones = case_problem(case)
    s = ca.Solver(P, q, A, b, cones, ca.Settings(), _library=emu)
    kw = {}
    if "update_b" in case:
        s.solve()
        s.update_b(case["update_b"])
        s.update_P(np.array(case["update_P"]))
        kw = dict(update_b=np.array(case["update_b"]), update_P=np.array(case["update_P"]))
    sol = s.solve()
    ref = oracle.solve(P, q, A, b, cones, **kw)
    assert_parity(sol, ref, name=case["name"], oracle=None if kw else oracle, problem=(P, q, A, b, cones))


def test_portfolio_and_trace(emu, oracle):
    P, q, A, b, cones = problems.portfolio(1)
    s = ca.Solver(P, q, A, b, cones, _library=emu)
    sol = s.solve()
    ref = oracle.solve(P, q, A, b, cones)
    assert_parity(sol, ref, iters=0)
    np.testing.assert_allclose(s.trace()[:, :2], ref["trace"][:, :2], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("method", ["hip", "hipmixed"])
def test_mpc_stream_sample(emu, oracle, method):
    for k, (P, q, A, b, cones) in enumerate(problems.mpc_stream(5, 6)):
        sol = ca.Solver(P, q, A, b, cones, ca.Settings(direct_solve_method=method), _library=emu).solve()
        ref = oracle.solve(P, q, A, b, cones)
        assert_parity(sol, ref, name=f"mpc{k}-{method}", oracle=oracle, problem=(P, q, A, b, cones))


@pytest.mark.parametrize("engine", ["workgroup", "resident"])
def test_server_alone_vs_mix_bitwise(emu, engine):
    reqs = list(problems.mpc_stream(7, 5)) + [problems.portfolio(3)]
    alone = []
    for P, q, A, b, cones in reqs:
        alone.append(ca.Solver(P, q, A, b, cones, _library=emu).solve())
    srv = ca.Server(hbm_budget_bytes=1 << 30, max_slots=3, settings=ca.Settings(serve_engine=engine), _library=emu)
    tickets = [srv.submit(*r) for r in reqs]
    srv.run(passes=3)
    for t, a in zip(tickets, alone):
        r = srv.result(t)
        assert r.status == a.status and r.iterations == a.iterations
        assert np.array_equal(r.x, a.x) and np.array_equal(r.z, a.z) and np.array_equal(r.s, a.s)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["active_slots"] == 0


def _serve(emu, reqs, engine, max_slots=256, passes=3, method="hip"):
    srv = ca.Server(hbm_budget_bytes=8 << 30, max_slots=max_slots,
                    settings=ca.Settings(direct_solve_method=method, serve_engine=engine), _library=emu)
    tickets = [srv.submit(*r) for r in reqs]
    srv.run(passes=passes)
    out = [srv.result(t) for t in tickets]
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["active_slots"] == 0
    return out


def test_lane_engine_reference_cases(emu, oracle):
    """Lane engine (one request per lane, lane-interleaved values) on every
    golden case, each its own shape and batch."""
    cases = [c for c in engine_cases() if "update_b" not in c]
    reqs = [case_problem(c) for c in cases]
    got = _serve(emu, reqs, "lanes")
    for c, r, g in zip(cases, reqs, got):
        assert_parity(g, oracle.solve(*r), name=c["name"], oracle=oracle, problem=r)


@pytest.mark.parametrize("method", ["hip", "hipmixed"])
def test_lane_engine_matches_workgroup_engine_bitwise(emu, method):
    """Serially both engines run the same arithmetic in the same order, so the
    lane engine's interleaved layout must reproduce the workgroup engine's
    results bit for bit: 70 same-shape requests (two batches, one full) plus
    other shapes in the mix."""
    rng = np.random.RandomState(11)
    reqs = [problems.mpc(0, T=12, rng=rng) for _ in range(70)] + list(problems.mpc_stream(9, 4))
    reqs.append(problems.portfolio(2))
    lanes = _serve(emu, reqs, "lanes", method=method)
    wg = _serve(emu, reqs, "workgroup", method=method)
    for k, (a, b) in enumerate(zip(lanes, wg)):
        assert a.status == b.status and a.iterations == b.iterations, k
        assert np.array_equal(a.x, b.x) and np.array_equal(a.s, b.s) and np.array_equal(a.z, b.z), k
        assert a.info["ir_steps"] == b.info["ir_steps"], k


def test_lane_engine_alone_vs_mix_bitwise(emu):
    """A request's result does not depend on the lanes or batches beside it."""
    rng = np.random.RandomState(12)
    reqs = [problems.mpc(0, T=10, rng=rng) for _ in range(5)] + [problems.portfolio(4)]
    mix = _serve(emu, reqs, "lanes", max_slots=4, passes=2)
    for r, m in zip(reqs, mix):
        a = _serve(emu, [r], "lanes")[0]
        assert a.status == m.status and a.iterations == m.iterations
        assert np.array_equal(a.x, m.x) and np.array_equal(a.z, m.z) and np.array_equal(a.s, m.s)


@pytest.mark.parametrize("method", ["hip", "hipmixed"])
@pytest.mark.parametrize("args", SDP_CASES[:2], ids=str)
def test_sdp_relaxation(emu, oracle, args, method):
    """PSD cones (NT scaling by Cholesky + SVD, skron Hs blocks) against the oracle."""
    prob = problems.psd_relaxation(*args)
    sol = ca.Solver(*prob, ca.Settings(direct_solve_method=method), _library=emu).solve()
    assert_parity(sol, oracle.solve(*prob), name=str(args), oracle=oracle, problem=prob)


@pytest.mark.parametrize("method", ["hip", "hipmixed"])
@pytest.mark.parametrize("args", LOGISTIC_CASES[:2], ids=str)
def test_logistic_exp_pow(emu, oracle, args, method):
    prob = problems.logistic_exp_pow(*args)
    sol = ca.Solver(*prob, ca.Settings(direct_solve_method=method), _library=emu).solve()
    assert_parity(sol, oracle.solve(*prob), name=str(args), oracle=oracle, problem=prob)


@pytest.mark.parametrize("method", ["hip", "hipmixed"])
@pytest.mark.parametrize("args", SOCP_CASES[:2], ids=str)
def test_random_sparse_socp(emu, oracle, args, method):
    prob = socp_case(args)
    sol = ca.Solver(*prob, ca.Settings(direct_solve_method=method), _library=emu).solve()
    assert_parity(sol, oracle.solve(*prob), name=str(args), oracle=oracle, problem=prob)


@pytest.mark.parametrize("engine", ["workgroup", "grid"])
@pytest.mark.parametrize("method", ["hip", "hipmixed"])
@pytest.mark.parametrize("args", SPARSE_SOC_CASES[:2], ids=str)
def test_sparse_soc_expansion(emu, oracle, args, method, engine):
    """Second-order cones of dimension 6..20, each expanded into the sparse
    (d, u, v) form with two extra KKT rows (src/kktsolvers/
    kktsolver_directldl_gpu.jl:490-585), on both engines' code paths."""
    prob = socp_case(args)
    st = ca.Settings(direct_solve_method=method, serve_engine=engine)
    s = ca.Solver(*prob, st, _library=emu)
    sol = s.solve()
    nsp = sum(1 for c in prob[4] if isinstance(c, ca.SecondOrderConeT) and c.dim > 5)
    assert nsp > 0 and sol.info["kkt_dim"] == prob[2].shape[0] + prob[2].shape[1] + 2 * nsp
    assert_parity(sol, oracle.solve(*prob), name=str(args), oracle=oracle, problem=prob)


@pytest.mark.parametrize("method", ["hip", "hipmixed"])
@pytest.mark.parametrize("which", ["q", "A", "qA", "qAbP"])
def test_warm_resolve_data_updates(emu, oracle, which, method):
    """update_q! / update_A! (and with b, P) between two solves keep the first
    solve's equilibration (src/data_updating.jl:120-202): the second solve
    against the oracle's solve - update - solve of the same data."""
    prob, kw = data_update_case(which)
    print("updates", which, method, flush=True)
    s = ca.Solver(*prob, ca.Settings(direct_solve_method=method), _library=emu)
    first = s.solve()
    assert first.status == "SOLVED"
    apply_updates(s, kw)
    sol = s.solve()
    ref = oracle.solve(*prob, **kw)
    assert_parity(sol, ref, name=which)
    assert abs(sol.obj_val - first.obj_val) > 1e-6 * abs(first.obj_val)  # the update took effect


def test_same_problem_full_then_upper_triangular_P(emu, oracle):
    """P given in full and as its upper triangle describe the same problem; the
    per-shape cache must not hand the second the first one's entry map"""
    import scipy.sparse as sp
    A = sp.csc_matrix(np.array([[1.0, 1], [1, 0], [0, 1]]))
    cones = [ca.NonnegativeConeT(1), ca.ZeroConeT(1), ca.NonnegativeConeT(1)]
    full = sp.csc_matrix(np.array([[4.0, 1], [1, 2]]))
    upper = sp.csc_matrix(np.array([[4.0, 1], [0, 2]]))
    a = ca.Solver(full, np.ones(2), A, np.ones(3), cones, _library=emu).solve()
    b = ca.Solver(upper, np.ones(2), A, np.ones(3), cones, _library=emu).solve()
    ref = oracle.solve(upper, np.ones(2), A, np.ones(3), cones)
    for s in (a, b):
        assert_parity(s, ref, iters=0)
    np.testing.assert_array_equal(a.x, b.x)


GRID_CASES = [("portfolio", "hip"), ("mpc12", "hipmixed"), ("socp200", "hip"), ("socp200", "hipmixed"),
              ("logistic", "hip"), ("sdp", "hipmixed")]


def _grid_problem(name):
    if name == "portfolio":
        return problems.portfolio(1)
    if name == "mpc12":
        return problems.mpc(0, T=12)
    if name == "socp200":
        return socp_case(SOCP_CASES[0])
    if name == "logistic":
        return problems.logistic_exp_pow(*LOGISTIC_CASES[0])
    return problems.psd_relaxation(*SDP_CASES[0])


@pytest.mark.parametrize("name,method", GRID_CASES, ids=lambda v: str(v))
def test_grid_engine_serial(emu, oracle, name, method):
    """The grid engine's code paths (the top levels of the elimination tree
    factored one supernode at a time by the whole grid: split update lists
    summed in order, the blocked right-looking dense factorization of the
    panel, the wide triangular solves) run by one thread, against the oracle."""
    prob = _grid_problem(name)
    sol = ca.Solver(*prob, ca.Settings(direct_solve_method=method, serve_engine="grid"), _library=emu).solve()
    assert_parity(sol, oracle.solve(*prob), name=name, oracle=oracle, problem=prob)


@pytest.mark.parametrize("engine", ["workgroup", "grid"])
@pytest.mark.parametrize("name", ["mpc12", "socp200", "sdp"])
def test_relaxed_supernodes(emu, oracle, monkeypatch, engine, name):
    """Relaxed supernodes (a column joining its child's panel although their
    structures differ: the panel holds explicit zeros) must leave the
    solution as the oracle's and take fewer supernodes than the fundamental
    ones (CLH_RELAX=0 forces the fundamental supernodes)."""
    prob = _grid_problem(name)
    st = ca.Settings(direct_solve_method="hip", serve_engine=engine)
    monkeypatch.setenv("CLH_RELAX", "0")
    plain = ca.Solver(*prob, st, _library=emu).solve()
    monkeypatch.setenv("CLH_RELAX", "1")
    sol = ca.Solver(*prob, st, _library=emu).solve()
    assert sol.info["n_supernodes"] < plain.info["n_supernodes"]
    assert_parity(sol, oracle.solve(*prob), name=name, oracle=oracle, problem=prob)


@pytest.mark.parametrize("engine", ["workgroup", "grid"])
def test_time_limit_zero(emu, oracle, engine):
    deadline_cases.check_solver_zero_limit(emu, oracle, engine)


def test_time_limit_sweep(emu):
    deadline_cases.check_solver_limit_sweep(emu, "workgroup")


@pytest.mark.parametrize("engine", ["workgroup", "resident"])
def test_server_deadlines_and_cancel(emu, oracle, engine):
    deadline_cases.check_server_deadlines_and_cancel(emu, oracle, engine)



def test_grid_abort_path(emu, oracle, monkeypatch):
    deadline_cases.check_grid_abort_path(emu, oracle, monkeypatch)
