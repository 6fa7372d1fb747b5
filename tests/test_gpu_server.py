"""Continuous batching on the GPU: every served result is the same whether
the request is served alone or in a mix (bitwise) and matches the oracle;
requests of different shapes overlap on the device and are admitted as slots
and memory free up.  Both serving engines are covered: the lane engine (up to
64 same-shape requests per 64-lane workgroup) and the workgroup engine (whose
served results also equal the single-problem Solver's, which runs the same
kernel)."""
import numpy as np
import pytest

import cuclarabel_amd as ca
from cuclarabel_amd import problems

from helpers import assert_parity

pytestmark = pytest.mark.gpu


def serve(reqs, st, max_slots=4, budget=2 << 30, passes=2):
    srv = ca.Server(hbm_budget_bytes=budget, max_slots=max_slots, settings=st)
    tickets = [srv.submit(*r) for r in reqs]
    srv.run(passes=passes)
    out = [srv.result(t) for t in tickets]
    return out, srv.stats()


def same(a, b):
    return (a.status == b.status and a.iterations == b.iterations and np.array_equal(a.x, b.x)
            and np.array_equal(a.s, b.s) and np.array_equal(a.z, b.z))


@pytest.mark.parametrize("engine", ["lanes", "workgroup", "resident"])
@pytest.mark.parametrize("method", ["hip", "hipmixed"])
def test_alone_vs_mix_and_oracle(oracle, method, engine):
    st = ca.Settings(direct_solve_method=method, serve_engine=engine)
    reqs = list(problems.mpc_stream(21, 10)) + [problems.portfolio(2)]
    print("solving alone", method, engine, flush=True)
    if engine in ("workgroup", "resident"):  # the same slot code as the single-problem solver
        alone = [ca.Solver(*r, ca.Settings(direct_solve_method=method, serve_engine="workgroup")).solve()
                 for r in reqs]
    else:
        alone = [serve([r], st)[0][0] for r in reqs]
    print("serving", flush=True)
    got, s = serve(reqs, st)
    for k, (g, a, r) in enumerate(zip(got, alone, reqs)):
        assert same(g, a), k
        assert_parity(g, oracle.solve(*r), name=f"req{k}", oracle=oracle, problem=r)
    assert s["completed"] == len(reqs) and s["active_slots"] == 0 and s["queued"] == 0
    assert s["launches"] >= 1 and s["p99_latency_s"] >= s["p50_latency_s"] > 0


@pytest.mark.parametrize("method", ["hip", "hipmixed"])
def test_lane_batches_full_and_partial(oracle, method):
    """70 requests of one shape fill one 64-lane batch and part of a second;
    every lane matches the oracle and a sample matches its alone result."""
    rng = np.random.RandomState(31)
    reqs = [problems.mpc(0, T=20, rng=rng) for _ in range(70)]
    st = ca.Settings(direct_solve_method=method, serve_engine="lanes")
    print("serving 70 same-shape requests", method, flush=True)
    got, s = serve(reqs, st, max_slots=256, passes=4)
    assert s["completed"] == 70
    for k, (g, r) in enumerate(zip(got, reqs)):
        assert_parity(g, oracle.solve(*r), name=f"lane{k}", oracle=oracle, problem=r)
    for k in (0, 37, 69):
        assert same(serve([reqs[k]], st)[0][0], got[k]), k


def test_admission_by_hbm_budget():
    """A budget that fits one slot at a time still completes every request;
    one larger than the budget is rejected with an error."""
    reqs = list(problems.mpc_stream(23, 3))
    srv = ca.Server(hbm_budget_bytes=24 << 20, max_slots=8)
    tickets = [srv.submit(*r) for r in reqs]
    srv.run(passes=8)
    for t in tickets:
        assert srv.result(t).status in ("SOLVED", "PRIMAL_INFEASIBLE")
    tiny = ca.Server(hbm_budget_bytes=1 << 16, max_slots=2)
    t = tiny.submit(*reqs[0])
    tiny.run()
    with pytest.raises(RuntimeError):
        tiny.result(t)


def test_lane_admission_by_hbm_budget():
    """A budget that holds one lane batch at a time serves two shapes in turn."""
    rng = np.random.RandomState(33)
    reqs = [problems.mpc(0, T=10, rng=rng) for _ in range(3)] + [problems.mpc(0, T=11, rng=rng) for _ in range(3)]
    st = ca.Settings(serve_engine="lanes")
    got, s = serve(reqs, st, max_slots=16, budget=30 << 20, passes=4)
    assert s["completed"] == 6 and s["rejected"] == 0
    for g, r in zip(got, reqs):
        assert g.status in ("SOLVED", "PRIMAL_INFEASIBLE")
