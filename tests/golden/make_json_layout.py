"""Writes json_reference_layout.json: a problem document in the layout the
reference's write_to_file produces (src/json.jl:118-162, JSON.lower of
JsonProblemData): every Settings field (symbols as strings, Inf as floatmax),
0-based CSC matrices with P upper-triangular, single-key cone objects.

The problem: min x'Px/2 + q'x over x in R^4 with x1+x2+x3+x4 = 1, x >= -3,
(x1, x2, x3) in SOC(3), (x4, 1, x1 + 2) in K_exp, (x2 + 2, x3 + 2, x4) in
K_pow(0.4).  Run from the repository root: python tests/golden/make_json_layout.py"""
import json
import os
from collections import OrderedDict

settings = OrderedDict([
    ("max_iter", 200), ("time_limit", 1.7976931348623157e308), ("verbose", True), ("max_step_fraction", 0.99),
    ("tol_gap_abs", 1e-8), ("tol_gap_rel", 1e-8), ("tol_feas", 1e-8), ("tol_infeas_abs", 1e-8),
    ("tol_infeas_rel", 1e-8), ("tol_ktratio", 1e-6), ("reduced_tol_gap_abs", 5e-5), ("reduced_tol_gap_rel", 5e-5),
    ("reduced_tol_feas", 1e-4), ("reduced_tol_infeas_abs", 5e-12), ("reduced_tol_infeas_rel", 5e-5),
    ("reduced_tol_ktratio", 1e-4), ("equilibrate_enable", True), ("equilibrate_max_iter", 10),
    ("equilibrate_min_scaling", 1e-4), ("equilibrate_max_scaling", 1e4), ("linesearch_backtrack_step", 0.8),
    ("min_switch_step_length", 1e-2), ("min_terminate_step_length", 1e-4), ("max_threads", 0),
    ("direct_kkt_solver", True), ("direct_solve_method", "cudss"), ("static_regularization_enable", True),
    ("static_regularization_constant", 1e-8), ("static_regularization_proportional", 4.930380657631324e-32),
    ("dynamic_regularization_enable", True), ("dynamic_regularization_eps", 1e-13),
    ("dynamic_regularization_delta", 2e-7), ("iterative_refinement_enable", True),
    ("iterative_refinement_reltol", 1e-12), ("iterative_refinement_abstol", 1e-12),
    ("iterative_refinement_max_iter", 10), ("iterative_refinement_stop_ratio", 5.0), ("presolve_enable", False),
    ("chordal_decomposition_enable", False), ("chordal_decomposition_merge_method", "clique_graph"),
    ("chordal_decomposition_compact", True), ("chordal_decomposition_complete_dual", True),
    ("neighborhood", 1e-6), ("soc_threshold", 5)])
P = OrderedDict([("m", 4), ("n", 4), ("colptr", [0, 1, 3, 4, 5]), ("rowval", [0, 0, 1, 2, 3]),
                 ("nzval", [2.0, 0.5, 1.0, 1.0, 0.5])])
q = [1.0, -1.0, 0.5, -0.25]
rows, b = [[1, 1, 1, 1]], [1.0]                       # zero cone: sum x = 1
for j in range(4):                                    # nonnegative: 3 + x_j >= 0
    r = [0] * 4; r[j] = -1; rows.append(r); b.append(3.0)
for j in range(3):                                    # SOC: (x1, x2, x3)
    r = [0] * 4; r[j] = -1; rows.append(r); b.append(0.0)
rows += [[0, 0, 0, -1], [0, 0, 0, 0], [-1, 0, 0, 0]]  # exp: (x4, 1, x1 + 2)
b += [0.0, 1.0, 2.0]
rows += [[0, -1, 0, 0], [0, 0, -1, 0], [0, 0, 0, -1]]  # pow: (x2 + 2, x3 + 2, x4)
b += [2.0, 2.0, 0.0]
colptr, rowval, nzval = [0], [], []
for j in range(4):
    for i, r in enumerate(rows):
        if r[j]:
            rowval.append(i); nzval.append(float(r[j]))
    colptr.append(len(rowval))
A = OrderedDict([("m", len(rows)), ("n", 4), ("colptr", colptr), ("rowval", rowval), ("nzval", nzval)])
cones = [{"ZeroConeT": 1}, {"NonnegativeConeT": 4}, {"SecondOrderConeT": 3}, {"ExponentialConeT": []},
         {"PowerConeT": 0.4}]
doc = OrderedDict([("settings", settings), ("P", P), ("q", q), ("A", A), ("b", b), ("cones", cones)])
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "json_reference_layout.json"), "w") as f:
    json.dump(doc, f)
