"""Continuous batching on the GPU: every served result equals the same request
solved alone (bitwise) and matches the oracle; requests of different shapes
overlap in the slot table and are admitted as slots free up."""
import numpy as np
import pytest

import cuclarabel_amd as ca
from cuclarabel_amd import problems

from helpers import assert_parity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["hip", "hipmixed"])
def test_alone_vs_mix_and_oracle(oracle, method):
    st = ca.Settings(direct_solve_method=method)
    reqs = list(problems.mpc_stream(21, 10)) + [problems.portfolio(2)]
    print("solving alone", method, flush=True)
    alone = [ca.Solver(*r, st).solve() for r in reqs]
    srv = ca.Server(hbm_budget_bytes=2 << 30, max_slots=4, settings=st)
    print("serving", flush=True)
    tickets = [srv.submit(*r) for r in reqs]
    srv.run(passes=2)
    for k, (t, a, r) in enumerate(zip(tickets, alone, reqs)):
        got = srv.result(t)
        assert got.status == a.status and got.iterations == a.iterations, k
        assert np.array_equal(got.x, a.x) and np.array_equal(got.s, a.s) and np.array_equal(got.z, a.z), k
        assert_parity(got, oracle.solve(*r), name=f"req{k}", oracle=oracle, problem=r)
    s = srv.stats()
    assert s["completed"] == len(reqs) and s["active_slots"] == 0 and s["queued"] == 0
    assert s["launches"] > 1 and s["p99_latency_s"] >= s["p50_latency_s"] > 0


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
