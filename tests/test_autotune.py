"""The autotuner: schedules chosen by timing candidates on the engine, kept
per size class in a cache file, and results that still match the oracle
under whatever schedule was chosen.  CPU tests run the tuner end to end on
the serial build of the engine; the GPU tests run it on the MI355X."""
import numpy as np
import pytest

import cuclarabel_amd as ca
from cuclarabel_amd import problems

from helpers import SOCP_CASES, assert_parity, build_emu, socp_case


@pytest.fixture(scope="module")
def emu():
    return build_emu()


def _mpc_sample(count, seed=3):
    return [problems.mpc(seed + i, T=10 + (i % 3) * 3) for i in range(count)]


def test_server_autotune_emu(emu, oracle, tmp_path, monkeypatch):
    cache = tmp_path / "schedules.txt"
    monkeypatch.setenv("CLH_TUNE_CACHE", str(cache))
    reqs = _mpc_sample(8)
    srv = ca.Server(hbm_budget_bytes=1 << 30, max_slots=4, settings=ca.Settings(direct_solve_method="hipmixed"),
                    _library=emu)
    tickets = [srv.submit(*r) for r in reqs]
    rep = srv.autotune(max_seconds=30)
    assert rep["trials"] >= 3 and rep["classes"] >= 1 and rep["default_rate"] > 0 and rep["tuned_rate"] > 0
    st = srv.stats()
    assert st["completed"] == 0 and st["queued"] == len(reqs)  # the trials leave no trace in the statistics
    srv.run(passes=4)
    for t, r in zip(tickets, reqs):  # the sample, served after tuning under the chosen schedules
        assert_parity(srv.result(t), oracle.solve(*r), name="mpc")
    lines = cache.read_text().splitlines()
    assert lines and all("|serve|f32|" in l and len(l.split("\t")[1].split()) == 9 for l in lines)
    # a new server with use_tuned serves each class with the cached schedule
    n_plus_m = reqs[0][2].shape[0] + reqs[0][2].shape[1]
    tuned = srv.schedule(n_plus_m)
    srv2 = ca.Server(hbm_budget_bytes=1 << 30, max_slots=4, _library=emu,
                     settings=ca.Settings(direct_solve_method="hipmixed", schedule=ca.Schedule(use_tuned=True)))
    got = srv2.schedule(n_plus_m)
    assert (got.relax_max_cols, got.relax_zero_pct, got.group_lanes) == (
        tuned.relax_max_cols, tuned.relax_zero_pct, tuned.group_lanes)


def test_server_autotune_needs_a_sample(emu):
    srv = ca.Server(hbm_budget_bytes=1 << 30, max_slots=4, _library=emu)
    with pytest.raises(ValueError, match="submit a sample"):
        srv.autotune(5)


def test_solver_autotune_emu(emu, oracle, tmp_path, monkeypatch):
    monkeypatch.setenv("CLH_TUNE_CACHE", str(tmp_path / "schedules.txt"))
    prob = socp_case(SOCP_CASES[0])
    s = ca.Solver(*prob, ca.Settings(serve_engine="grid"), _library=emu)
    rep = s.autotune(max_seconds=60)
    assert rep["trials"] >= 2 and rep["tuned_rate"] >= rep["default_rate"]
    sc = s.schedule()
    assert sc.grid_blocks > 0 and sc.tile_min_cols in (64, 128, 256) and sc.dense_min_cols in (64, 128, 256)
    assert_parity(s.solve(), oracle.solve(*prob), name="socp")
    with pytest.raises(ValueError, match="before the first"):
        s.autotune(5)
    s2 = ca.Solver(*prob, ca.Settings(serve_engine="grid", schedule=ca.Schedule(use_tuned=True)), _library=emu)
    got = s2.schedule()
    assert (got.grid_blocks, got.tile_min_cols, got.dense_min_cols) == (
        sc.grid_blocks, sc.tile_min_cols, sc.dense_min_cols)


def test_workgroup_solver_has_nothing_to_tune(emu):
    s = ca.Solver(*problems.portfolio(1), _library=emu)
    rep = s.autotune(5)
    assert rep["trials"] == 0


@pytest.mark.parametrize("sched", [dict(relax_max_cols=-1), dict(relax_max_cols=32, relax_zero_pct=10),
                                   dict(group_lanes=16), dict(group_lanes=64)])
def test_schedule_knobs_keep_parity(emu, oracle, sched):
    """every schedule the tuner may pick gives a result that matches the
    oracle; no relaxation gives more supernodes than the default"""
    prob = problems.mpc(11, T=14)
    ref = oracle.solve(*prob)
    base = ca.Solver(*prob, ca.Settings(), _library=emu).solve()
    got = ca.Solver(*prob, ca.Settings(schedule=ca.Schedule(**sched)), _library=emu).solve()
    assert_parity(got, ref, name=str(sched))
    if sched.get("relax_max_cols") == -1:
        assert got.info["n_supernodes"] > base.info["n_supernodes"]


def test_invalid_schedule_rejected(emu):
    with pytest.raises(ValueError, match="group_lanes"):
        ca.Solver(*problems.portfolio(1), ca.Settings(schedule=ca.Schedule(group_lanes=24)), _library=emu)


@pytest.mark.gpu
def test_gpu_server_autotune(oracle, tmp_path, monkeypatch):
    monkeypatch.setenv("CLH_TUNE_CACHE", str(tmp_path / "schedules.txt"))
    reqs = list(problems.mpc_stream(31, 600))
    srv = ca.Server(hbm_budget_bytes=16 << 30, max_slots=1024, settings=ca.Settings(direct_solve_method="hipmixed"))
    tickets = [srv.submit(*r) for r in reqs]
    print("autotune on", len(reqs), "requests", flush=True)
    rep = srv.autotune(max_seconds=40)
    print(rep, flush=True)
    assert rep["trials"] >= 3 and rep["classes"] >= 2
    srv.run(passes=4)
    res = [srv.result(t) for t in tickets]
    for r, p in list(zip(res, reqs))[::20]:
        assert_parity(r, oracle.solve(*p), name="mpc")


@pytest.mark.gpu
def test_gpu_solver_autotune(oracle, tmp_path, monkeypatch):
    monkeypatch.setenv("CLH_TUNE_CACHE", str(tmp_path / "schedules.txt"))
    prob = problems.large_socp(5, n=1500, m_zero=400, m_nn=1100, n_soc=150)
    s = ca.Solver(*prob, ca.Settings(serve_engine="grid"))
    rep = s.autotune(max_seconds=60)
    print(rep, s.schedule(), flush=True)
    assert rep["trials"] >= 2
    assert_parity(s.solve(), oracle.solve(*prob), name="socp1500")
