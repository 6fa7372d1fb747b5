"""The reference's JSON problem format (src/json.jl): the checks of
test/UnitTests/test_json.jl (write, read back, solve both; read with custom
settings), a document in the reference's layout (tests/golden/
json_reference_layout.json, written by make_json_layout.py) parsed and
solved against the oracle, and round trips of every supported cone."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import cuclarabel_amd as ca
from cuclarabel_amd import json_io, problems

from helpers import assert_parity, build_emu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "json_reference_layout.json")


@pytest.fixture(scope="module")
def emu():
    return build_emu()


def test_write_read_solve(emu, tmp_path):
    """test/UnitTests/test_json.jl:1-36"""
    P = sp.csc_matrix(np.array([[4.0, 1], [1, 2]]))
    c = np.array([1.0, 1])
    A = sp.csc_matrix(np.array([[1.0, 1], [1, 0], [0, 1]]))
    b = np.array([1.0, 1, 1])
    cones = [ca.NonnegativeConeT(1), ca.ZeroConeT(1), ca.NonnegativeConeT(1)]
    solver = ca.Solver(P, c, A, b, cones, _library=emu)
    f = str(tmp_path / "p.json")
    json_io.write_to_file(solver, f)
    solver2 = json_io.read_from_file(f, _library=emu)
    s1, s2 = solver.solve(), solver2.solve()
    np.testing.assert_allclose(s1.x, s2.x, atol=1e-10)
    assert s1.status == s2.status == "SOLVED"
    settings = ca.Settings(max_iter=1)
    solver3 = json_io.read_from_file(f, settings, _library=emu)
    assert solver3.solve().status == "MAX_ITERATIONS"


def test_written_layout(tmp_path, emu):
    """what write_to_file writes: the keys in order, every reference Settings
    field, 0-based CSC with the upper triangle of P, single-key cones"""
    P = sp.csc_matrix(np.array([[4.0, 1], [1, 2]]))
    A = sp.csc_matrix(np.array([[1.0, 1], [1, 0], [0, 1]]))
    solver = ca.Solver(P, np.ones(2), A, np.ones(3), [ca.NonnegativeConeT(3)], _library=emu)
    f = str(tmp_path / "p.json")
    json_io.write_to_file(solver, f)
    d = json.load(open(f))
    assert list(d) == ["settings", "P", "q", "A", "b", "cones"]
    ref_fields = list(json.load(open(GOLDEN))["settings"])
    assert list(d["settings"]) == ref_fields
    assert d["settings"]["time_limit"] == 1.7976931348623157e308
    assert d["settings"]["direct_solve_method"] == "hip"
    assert d["P"] == {"m": 2, "n": 2, "colptr": [0, 1, 3], "rowval": [0, 0, 1], "nzval": [4.0, 1.0, 2.0]}
    assert d["A"]["colptr"] == [0, 2, 4] and d["A"]["rowval"] == [0, 1, 0, 2]
    assert d["cones"] == [{"NonnegativeConeT": 3}]


def test_reference_layout_document(emu, oracle):
    P, q, A, b, cones, st = json_io.problem_from_json(open(GOLDEN).read())
    assert P.shape == (4, 4) and A.shape == (14, 4)
    assert [type(c).__name__ for c in cones] == ["ZeroConeT", "NonnegativeConeT", "SecondOrderConeT",
                                                 "ExponentialConeT", "PowerConeT"]
    assert cones[4].alpha == 0.4
    assert st.time_limit == float("inf") and st.direct_solve_method == "hip" and st.verbose is True
    assert st.static_regularization_proportional == 4.930380657631324e-32
    sol = ca.Solver(P, q, A, b, cones, ca.Settings(), _library=emu).solve()
    ref = oracle.solve(P, q, A, b, cones)
    assert_parity(sol, ref, name="json", oracle=oracle, problem=(P, q, A, b, cones))
    assert sol.status == "SOLVED"


def test_round_trip_all_cones():
    P, q, A, b, cones = problems.psd_relaxation(1, 12, 2, 3, 4)
    cones = cones + [ca.ExponentialConeT(), ca.PowerConeT(0.35), ca.SecondOrderConeT(4), ca.ZeroConeT(2)]
    m = sum(ca._numel(c) for c in cones)
    A = sp.vstack([A, sp.random(m - A.shape[0], A.shape[1], density=0.3, random_state=1)]).tocsc()
    b = np.r_[b, np.ones(m - b.size)]
    st = ca.Settings(direct_solve_method="hipmixed", max_iter=77, tol_gap_abs=3e-9)
    text = json_io.problem_to_json(P, q, A, b, cones, st)
    P2, q2, A2, b2, cones2, st2 = json_io.problem_from_json(text)
    assert (abs(sp.triu(P) - P2)).max() == 0 and (abs(A - A2)).max() == 0
    assert np.array_equal(q, q2) and np.array_equal(b, b2)
    assert [(type(c).__name__, c.dim, c.param) for c in cones] == [(type(c).__name__, c.dim, c.param) for c in cones2]
    assert st2.direct_solve_method == "hipmixed" and st2.max_iter == 77 and st2.tol_gap_abs == 3e-9
    assert st2.static_regularization_constant == pytest.approx(np.sqrt(np.finfo(np.float32).eps))


def test_unsupported_cone_rejected():
    d = json.load(open(GOLDEN))
    d["cones"][-1] = {"GenPowerConeT": [[0.5, 0.5], 1]}
    with pytest.raises(ValueError, match="GenPowerConeT"):
        json_io.problem_from_json(json.dumps(d))
