"""Time limits, cancellation and waiting (src/info.jl:182-183: a solve whose
total time passes settings.time_limit ends with MAX_TIME; the server applies
the limit per request from its submission and can cancel a request), shared
by the serial-build and the GPU tests: `lib` is the library to run on (None:
the HIP library)."""
import numpy as np
import pytest

import cuclarabel_amd as ca
from cuclarabel_amd import problems


def _kw(lib):
    return {} if lib is None else {"_library": lib}


def check_solver_zero_limit(lib, oracle, engine="auto"):
    """time_limit = 0: the first termination check ends the solve, as in the
    oracle (status MAX_TIME, no iterations, the post-processed start point)"""
    prob = problems.portfolio(1)
    sol = ca.Solver(*prob, ca.Settings(time_limit=0.0, serve_engine=engine), **_kw(lib)).solve()
    ref = oracle.solve(*prob, time_limit=0.0)
    assert ref["status"] == "MAX_TIME" and ref["iterations"] == 0
    assert sol.status == "MAX_TIME" and sol.iterations == 0


def check_solver_limit_sweep(lib, engine="auto"):
    """A limit below the solve's own time stops it early with MAX_TIME (or an
    ALMOST_* status its last iterate earns under the reduced tolerances);
    longer limits take at least as many iterations; no limit solves it.  The
    second solve after an update counts no setup time (src/info.jl:242-260)."""
    prob = problems.mpc(11, T=50)
    full = ca.Solver(*prob, ca.Settings(serve_engine=engine), **_kw(lib)).solve()
    assert full.status in ("SOLVED", "PRIMAL_INFEASIBLE")
    per_iter = full.solve_time / max(1, full.iterations)
    last = -1
    for frac in (0.0, 0.3, 0.6):
        s = ca.Solver(*prob, ca.Settings(time_limit=1e9, serve_engine=engine), **_kw(lib))
        s.solve()
        # same data again: the second solve's limit is the whole budget
        s2 = ca.Solver(*prob, ca.Settings(time_limit=frac * full.solve_time, serve_engine=engine), **_kw(lib))
        s2.solve()
        s2.update_b(prob[3])
        sol = s2.solve()
        assert sol.iterations >= last
        last = sol.iterations
        if sol.iterations < full.iterations:
            assert sol.status in ("MAX_TIME", "ALMOST_SOLVED", "ALMOST_PRIMAL_INFEASIBLE"), sol.status
        # it stopped within a few iterations of its limit
        assert sol.solve_time <= frac * full.solve_time + 5 * per_iter + 0.05
    assert last <= full.iterations


def check_server_deadlines_and_cancel(lib, oracle, engine="workgroup"):
    """Per-request limits and cancellation in the server: with time_limit = 0
    every request ends MAX_TIME at its first check; a queued request cancelled
    is answered at once with no iterations; a resident one cancelled after a
    pass ends early; the others are unaffected and match the oracle; wait()
    drives the device until a ticket finishes."""
    reqs = list(problems.mpc_stream(23, 4))
    srv = ca.Server(hbm_budget_bytes=1 << 30, max_slots=4, settings=ca.Settings(time_limit=0.0, serve_engine=engine),
                    **_kw(lib))
    tickets = [srv.submit(*r) for r in reqs]
    srv.run(passes=2)
    for t in tickets:
        r = srv.result(t)
        assert r.status == "MAX_TIME" and r.iterations == 0
    srv.close()

    srv = ca.Server(hbm_budget_bytes=1 << 30, max_slots=2, settings=ca.Settings(serve_engine=engine), **_kw(lib))
    tickets = [srv.submit(*r) for r in reqs]
    # the first two are admitted (one pass on the launch engine; the
    # resident loop's workgroups take them as the call returns), two queued
    srv.progress(passes=1, wait=engine != "resident")
    assert srv.cancel(tickets[3])     # queued: answered at once
    assert srv.poll(tickets[3])
    q = srv.result(tickets[3])
    assert q.status == "MAX_TIME" and q.iterations == 0 and not np.any(q.x)
    # resident: ends at its next termination check (the serial CPU build runs
    # a resident launch to its end at once, so there it has finished)
    hit = srv.cancel(tickets[1])
    assert hit or (engine == "resident" and lib is not None)
    assert srv.wait(tickets[2], passes=2)
    srv.run(passes=2)
    c = srv.result(tickets[1])
    ref1 = oracle.solve(*reqs[1])
    finished = c.status == ref1["status"] and abs(c.iterations - ref1["iterations"]) <= 1
    if hit and not (engine == "resident" and finished):
        # (a resident request may finish between the host's look and the
        # cancel word: then its result is the complete one)
        assert c.iterations < ref1["iterations"]
        assert c.status in ("MAX_TIME", "ALMOST_SOLVED", "ALMOST_PRIMAL_INFEASIBLE"), c.status
    else:
        assert finished
    for k in (0, 2):
        r = srv.result(tickets[k])
        ref = oracle.solve(*reqs[k])
        assert r.status == ref["status"] and abs(r.iterations - ref["iterations"]) <= 1
    with pytest.raises(ValueError):
        srv.cancel(tickets[0])        # collected: no longer known
    st = srv.stats()
    assert st["cancelled"] == 1 + int(hit) and st["completed"] == 3
    srv.close()


def check_grid_abort_path(lib, oracle, monkeypatch):
    """The grid engine's abort word, set by the host before the launch
    (CLH_ABORT_AFTER_S=0): the launch does nothing, the solve reports an
    error instead of a result, and the same solver then solves normally
    (the grid's barrier state is reset)."""
    prob = problems.portfolio(1)
    s = ca.Solver(*prob, ca.Settings(serve_engine="grid"), **_kw(lib))
    monkeypatch.setenv("CLH_ABORT_AFTER_S", "0")
    with pytest.raises(RuntimeError, match="aborted"):
        s.solve()
    monkeypatch.delenv("CLH_ABORT_AFTER_S")
    sol = s.solve()
    ref = oracle.solve(*prob)
    assert sol.status == ref["status"] and abs(sol.iterations - ref["iterations"]) <= 1
