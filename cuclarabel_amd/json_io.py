"""The reference's JSON problem format (src/json.jl:25-209): a file that
Clarabel.jl's read_from_file loads, and that write_to_file writes.

Layout (keys in this order):
    "settings": every field of Clarabel's Settings (src/settings.jl:70-148),
                symbols as strings, +-Inf stored as +-floatmax;
    "P", "A":   {"m", "n", "colptr", "rowval", "nzval"}, 0-based CSC; P holds
                its upper triangle, as the reference's problem data does;
    "q", "b":   arrays;
    "cones":    one single-key object per cone: {"ZeroConeT": dim},
                {"NonnegativeConeT": dim}, {"SecondOrderConeT": dim},
                {"ExponentialConeT": []}, {"PowerConeT": alpha},
                {"PSDTriangleConeT": side}; {"GenPowerConeT": [alpha, dim2]}
                is part of the format but not of this solver and is rejected.
"""
from __future__ import annotations

import json
import sys
from collections import OrderedDict
from typing import Optional

import numpy as np
import scipy.sparse as sp

from . import (ExponentialConeT, NonnegativeConeT, PowerConeT, PSDTriangleConeT, SecondOrderConeT, Settings, Solver,
               ZeroConeT, _numel)

FLOATMAX = sys.float_info.max

# the reference's Settings fields in declaration order, with the values this
# solver writes for the fields it has no counterpart of (no presolve, no
# chordal decomposition: the GPU path runs without them)
_REFERENCE_FIELDS = [
    "max_iter", "time_limit", "verbose", "max_step_fraction",
    "tol_gap_abs", "tol_gap_rel", "tol_feas", "tol_infeas_abs", "tol_infeas_rel", "tol_ktratio",
    "reduced_tol_gap_abs", "reduced_tol_gap_rel", "reduced_tol_feas", "reduced_tol_infeas_abs",
    "reduced_tol_infeas_rel", "reduced_tol_ktratio",
    "equilibrate_enable", "equilibrate_max_iter", "equilibrate_min_scaling", "equilibrate_max_scaling",
    "linesearch_backtrack_step", "min_switch_step_length", "min_terminate_step_length",
    "max_threads", "direct_kkt_solver", "direct_solve_method",
    "static_regularization_enable", "static_regularization_constant", "static_regularization_proportional",
    "dynamic_regularization_enable", "dynamic_regularization_eps", "dynamic_regularization_delta",
    "iterative_refinement_enable", "iterative_refinement_reltol", "iterative_refinement_abstol",
    "iterative_refinement_max_iter", "iterative_refinement_stop_ratio",
    "presolve_enable", "chordal_decomposition_enable", "chordal_decomposition_merge_method",
    "chordal_decomposition_compact", "chordal_decomposition_complete_dual", "neighborhood", "soc_threshold",
]
_FIXED = {"max_threads": 0, "direct_kkt_solver": True, "presolve_enable": False,
          "chordal_decomposition_enable": False, "chordal_decomposition_merge_method": "clique_graph",
          "chordal_decomposition_compact": True, "chordal_decomposition_complete_dual": True,
          "neighborhood": 1e-6}

_CONES = {"ZeroConeT": ZeroConeT, "NonnegativeConeT": NonnegativeConeT, "SecondOrderConeT": SecondOrderConeT,
          "PSDTriangleConeT": PSDTriangleConeT}


def _sanitize(v):
    if isinstance(v, float) and np.isinf(v):
        return FLOATMAX if v > 0 else -FLOATMAX
    return v


def _desanitize(v):
    if isinstance(v, float) and abs(v) == FLOATMAX:
        return float("inf") if v > 0 else float("-inf")
    return v


def _lower_csc(M, shape):
    M = sp.csc_matrix(M, shape=shape)
    M.sort_indices()
    return OrderedDict([("m", int(shape[0])), ("n", int(shape[1])), ("colptr", [int(v) for v in M.indptr]),
                        ("rowval", [int(v) for v in M.indices]), ("nzval", [float(v) for v in M.data])])


def _parse_csc(d):
    return sp.csc_matrix((np.asarray(d["nzval"], np.float64), np.asarray(d["rowval"], np.int64),
                          np.asarray(d["colptr"], np.int64)), shape=(int(d["m"]), int(d["n"])))


def _lower_cone(c):
    if isinstance(c, ExponentialConeT):
        return OrderedDict([("ExponentialConeT", [])])
    if isinstance(c, PowerConeT):
        return OrderedDict([("PowerConeT", float(c.alpha))])
    return OrderedDict([(type(c).__name__, int(c.dim))])


def _parse_cone(d):
    if len(d) != 1:
        raise ValueError("a cone entry has exactly one key")
    (key, val), = d.items()
    if key == "ExponentialConeT":
        return ExponentialConeT()
    if key == "PowerConeT":
        return PowerConeT(float(val))
    if key in _CONES:
        return _CONES[key](int(val))
    raise ValueError(f"cone type {key} is not supported by this solver")


def _lower_settings(s: Settings):
    out = OrderedDict()
    for f in _REFERENCE_FIELDS:
        if f in _FIXED:
            v = _FIXED[f]
        elif f == "static_regularization_constant" and s.static_regularization_constant is None:
            v = 1e-8 if str(s.direct_solve_method).lstrip(":") == "hip" else float(np.sqrt(np.finfo(np.float32).eps))
        else:
            v = getattr(s, f)
            if f == "direct_solve_method":
                v = str(v).lstrip(":")
        out[f] = _sanitize(v)
    return out


def _parse_settings(d) -> Settings:
    s = Settings()
    for key, val in d.items():
        val = _desanitize(val)
        if key == "direct_solve_method":
            s.direct_solve_method = "hipmixed" if str(val).lstrip(":") in ("hipmixed", "cudssmixed") else "hip"
        elif hasattr(s, key) and key not in ("serve_engine", "schedule"):
            cur = getattr(Settings(), key)
            if isinstance(cur, bool):
                val = bool(val)
            elif isinstance(cur, int) and not isinstance(cur, bool):
                val = int(val)
            elif val is not None and not isinstance(val, str):
                val = float(val)
            setattr(s, key, val)
        # fields of the reference this solver has no counterpart of are ignored
    return s


def problem_to_json(P, q, A, b, cones, settings: Optional[Settings] = None) -> str:
    """the JSON text of a problem (write_to_file's content)"""
    q = np.asarray(q, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    n, m = q.size, b.size
    if sum(_numel(c) for c in cones) != m:
        raise ValueError("cone dimensions do not match the number of rows of A")
    Pu = sp.triu(sp.csc_matrix(P, shape=(n, n)), format="csc")
    data = OrderedDict([("settings", _lower_settings(settings or Settings())), ("P", _lower_csc(Pu, (n, n))),
                        ("q", [float(v) for v in q]), ("A", _lower_csc(A, (m, n))), ("b", [float(v) for v in b]),
                        ("cones", [_lower_cone(c) for c in cones])])
    return json.dumps(data)


def problem_from_json(text: str):
    """(P, q, A, b, cones, settings) from the JSON text"""
    d = json.loads(text)
    P = _parse_csc(d["P"])
    A = _parse_csc(d["A"])
    q = np.asarray(d["q"], np.float64)
    b = np.asarray(d["b"], np.float64)
    cones = [_parse_cone(c) for c in d["cones"]]
    settings = _parse_settings(d.get("settings", {}))
    return P, q, A, b, cones, settings


def write_to_file(solver: Solver, filename: str) -> None:
    """write_to_file(solver, filename) (src/json.jl:25-57): the problem data
    the solver was built with (unscaled; P as its upper triangle) and its
    settings."""
    p = solver._p
    P = sp.csc_matrix((p.Px, p.Pi, p.Pp), shape=(p.n, p.n))
    A = sp.csc_matrix((p.Ax, p.Ai, p.Ap), shape=(p.m, p.n))
    with open(filename, "w") as f:
        f.write(problem_to_json(P, p.q, A, p.b, p.cones, solver.settings))


def read_from_file(filename: str, settings: Optional[Settings] = None, _library=None) -> Solver:
    """read_from_file(filename, settings = nothing) (src/json.jl:67-90): a
    Solver of the file's problem, with the file's settings unless given."""
    with open(filename) as f:
        P, q, A, b, cones, file_settings = problem_from_json(f.read())
    return Solver(P, q, A, b, cones, settings or file_settings, _library=_library)
