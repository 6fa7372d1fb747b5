"""Time limits (MAX_TIME), cancellation and waiting on the GPU: the solver
on both single-problem engines and the server (tests/deadline_cases.py)."""
import pytest

import deadline_cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("engine", ["workgroup", "grid"])
def test_time_limit_zero(oracle, engine):
    print("time limit 0", engine, flush=True)
    deadline_cases.check_solver_zero_limit(None, oracle, engine)


@pytest.mark.parametrize("engine", ["workgroup", "grid"])
def test_time_limit_sweep(engine):
    print("time limit sweep", engine, flush=True)
    deadline_cases.check_solver_limit_sweep(None, engine)


@pytest.mark.parametrize("engine", ["workgroup", "resident"])
def test_server_deadlines_and_cancel(oracle, engine):
    print("server deadlines and cancel", engine, flush=True)
    deadline_cases.check_server_deadlines_and_cancel(None, oracle, engine)


def test_grid_abort_path(oracle, monkeypatch):
    """the host sets the abort word before the launch: every workgroup
    returns at once, and the next solve on the same handle is normal"""
    print("grid abort path", flush=True)
    deadline_cases.check_grid_abort_path(None, oracle, monkeypatch)
