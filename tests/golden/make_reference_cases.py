"""Writes reference_cases.json: the deterministic known-answer problems of the
reference's own test suite and the two notebook runs of the CUDA path,
transcribed as data (problem matrices, cones, expected status / solution /
objective, the tolerance the reference test uses).

Run:  python tests/golden/make_reference_cases.py
Nothing here reads the reference at run time; the numbers below were typed
from the files cited in each case's "source" field.
"""
import json
import os

import numpy as np

Z, NN, SOC, EXP, POW, PSD = 0, 1, 2, 3, 4, 5


def dense(M):
    return np.asarray(M, float).tolist()


cases = []


def add(name, source, P, q, A, b, cones, status, x=None, obj=None, tol=1e-3, **extra):
    d = dict(name=name, source=source, P=dense(P), q=list(map(float, q)), A=dense(A), b=list(map(float, b)),
             cones=cones, status=status, tol=tol)
    if x is not None:
        d["x"] = list(map(float, x))
    if obj is not None:
        d["obj"] = float(obj)
    d.update(extra)
    cases.append(d)


# --- test/OptTests/basic_qp.jl -------------------------------------------
add("qp_univariate", "test/OptTests/basic_qp.jl:44-62", [[1.0]], [0.0], [[1.0]], [1.0], [[NN, 1]],
    "SOLVED", x=[0.0], obj=0.0)
A0 = np.array([[1.0, 1], [1, 0], [0, 1]])
l0 = np.array([1.0, 0, 0])
u0 = np.array([1.0, 0.7, 0.7])
Aqp = np.vstack([-A0, A0])
bqp = np.concatenate([-l0, u0])
add("qp_feasible", "test/OptTests/basic_qp.jl:6-19,64-75", [[4.0, 1], [1, 2]], [1.0, 1], Aqp, bqp,
    [[NN, 3], [NN, 3]], "SOLVED", x=[0.3, 0.7], obj=1.8800000298331538)
b2 = bqp.copy(); b2[0] = -1; b2[3] = -1
add("qp_primal_infeasible", "test/OptTests/basic_qp.jl:77-89", [[4.0, 1], [1, 2]], [1.0, 1], Aqp, b2,
    [[NN, 3], [NN, 3]], "PRIMAL_INFEASIBLE")
add("qp_dual_infeasible", "test/OptTests/basic_qp.jl:22-31,91-101", [[1.0, 1], [1, 1]], [1.0, -1],
    [[1.0, 1], [1, 0]], [1.0, 1], [[NN, 2]], "DUAL_INFEASIBLE")
add("qp_dual_infeasible_nonqsd", "test/OptTests/basic_qp.jl:103-116", [[1.0, 1], [1, 1]], [1.0, -1],
    [[1.0, 1]], [1.0], [[NN, 1]], "DUAL_INFEASIBLE")

# --- test/OptTests/basic_lp.jl -------------------------------------------
Alp = np.vstack([np.eye(3), -np.eye(3)]) * 2
add("lp_feasible", "test/OptTests/basic_lp.jl:6-16,26-37", np.zeros((3, 3)), [3.0, -2, 1], Alp, np.ones(6),
    [[NN, 3], [NN, 3]], "SOLVED", x=[-0.5, 0.5, -0.5], obj=-3.0)
b3 = np.ones(6); b3[0] = -1; b3[3] = -1
add("lp_primal_infeasible", "test/OptTests/basic_lp.jl:39-51", np.zeros((3, 3)), [3.0, -2, 1], Alp, b3,
    [[NN, 3], [NN, 3]], "PRIMAL_INFEASIBLE")
A4 = Alp.copy(); A4[3, 0] = 1.0
add("lp_dual_infeasible", "test/OptTests/basic_lp.jl:53-66", np.zeros((3, 3)), [1.0, 0, 0], A4, np.ones(6),
    [[NN, 3], [NN, 3]], "DUAL_INFEASIBLE")
A5 = Alp.copy(); A5[0, 0] = np.finfo(float).eps; A5[3, 0] = 0.0
add("lp_dual_infeasible_ill_conditioned", "test/OptTests/basic_lp.jl:68-82", np.zeros((3, 3)), [1.0, 0, 0], A5,
    np.ones(6), [[NN, 3], [NN, 3]], "DUAL_INFEASIBLE")

# --- test/OptTests/basic_eq_constrained.jl -------------------------------
add("eq_constrained", "test/OptTests/basic_eq_constrained.jl:14-28", np.eye(3), [0.0, 0, 0],
    [[0.0, 1, 1], [0, 1, -1]], [2.0, 0], [[Z, 2]], "SOLVED", x=[0.0, 1, 1])
add("eq_constrained_2", "test/OptTests/basic_eq_constrained.jl:30-44", np.eye(3), [1.0, 2, 3],
    [[1.0, 1, 1], [0, 1, -1]], [2.0, 0], [[Z, 2]], "SOLVED", x=[10 / 6, 1 / 6, 1 / 6])
Aeq = np.array([[0.0, 1, 1], [0, 1, -1]])
add("eq_constrained_redundant", "test/OptTests/basic_eq_constrained.jl:46-62", np.eye(3), [0.0, 0, 0],
    np.vstack([Aeq, Aeq]), [2.0, 0, 2, 0], [[Z, 2], [Z, 2]], "SOLVED", x=[0.0, 1, 1])
add("eq_constrained_primal_infeasible", "test/OptTests/basic_eq_constrained.jl:64-80", np.eye(3), [0.0, 0, 0],
    [[0.0, 1, 1], [0, 1, -1], [1, 2, -1], [2, -1, 3]], [1.0, 1, 1, 1], [[Z, 4]], "PRIMAL_INFEASIBLE")
P6 = np.eye(3); P6[0, 0] = 0
add("eq_constrained_dual_infeasible", "test/OptTests/basic_eq_constrained.jl:82-96", P6, [1.0, 1, 1],
    [[0.0, 1, 1], [0, 1, -1]], [2.0, 0], [[Z, 2]], "DUAL_INFEASIBLE")

# --- test/OptTests/basic_unconstrained.jl ---------------------------------
add("unconstrained", "test/OptTests/basic_unconstrained.jl:14-27", np.eye(3), [1.0, 2, -3], np.zeros((0, 3)), [],
    [], "SOLVED", x=[-1.0, -2, 3])
P7 = np.eye(3); P7[0, 0] = 0
add("unconstrained_dual_infeasible", "test/OptTests/basic_unconstrained.jl:29-44", P7, [1.0, 0, 0],
    np.zeros((0, 3)), [], [], "DUAL_INFEASIBLE")

# --- test/OptTests/basic_exp.jl -------------------------------------------
A1 = np.hstack([np.ones((1, 3)), np.zeros((1, 4))])
A2 = np.hstack([np.zeros((3, 2)), -np.eye(3), np.zeros((3, 2))])
A3 = np.zeros((3, 7)); A3[0, 0] = -1; A3[1, 2] = -1; A3[2, 4] = -1
add("exp_feasible", "test/OptTests/basic_exp.jl:6-62", np.eye(7) * 0.1, [1.0, 0.5, -2.0, -0.1, 1.0, 3.0, 0.0],
    np.vstack([A1, A2, A3]), [10.0, 0, 0, 0, 0, 0, 0], [[Z, 1], [NN, 3], [EXP, 3]], "SOLVED",
    x=[-9.425995201329599, 4.828561507482018, 14.59743362204262, 1.0000012112102774, 7.65314081561849,
       -29.99999978458479, -0.0], obj=-54.41243965302268)

# --- test/OptTests/basic_pow.jl -------------------------------------------
Ap = np.vstack([np.eye(6), [[1.0, 2, 0, 3, 0, 0]], [[0.0, 0, 0, 0, 1, 0]]])
add("pow_feasible", "test/OptTests/basic_pow.jl:6-54", np.zeros((6, 6)), [0.0, 0, -1, 0, 0, -1], -Ap,
    [0.0] * 6 + [-3.0, -1.0], [[POW, 3, 0.6], [POW, 3, 0.1], [Z, 1], [Z, 1]], "SOLVED", obj=-1.8458)

# --- test/OptTests/basic_sdp.jl -------------------------------------------
bsdp = [-3.0, 1.0, 4.0, 1.0, 2.0, 5.0]
refsol = [-3.0729833267361095, 0.3696004167288786, -0.022226685581313674, 0.31441213129613066,
          -0.026739700851545107, -0.016084530571308823]
add("sdp_feasible", "test/OptTests/basic_sdp.jl:6-48", np.eye(6), [0.0] * 6, np.eye(6), bsdp, [[PSD, 3]],
    "SOLVED", x=refsol, obj=4.840076866013861)
add("sdp_empty_cone", "test/OptTests/basic_sdp.jl:50-73", np.eye(6), [0.0] * 6, np.eye(6), bsdp,
    [[PSD, 3], [PSD, 0]], "SOLVED", x=refsol, obj=4.840076866013861)
add("sdp_primal_infeasible", "test/OptTests/basic_sdp.jl:75-89", np.eye(6), [0.0] * 6,
    np.vstack([np.eye(6), -np.eye(6)]), bsdp + [0.0] * 6, [[PSD, 3], [PSD, 3]], "PRIMAL_INFEASIBLE")

# --- src/python/call_in_python.ipynb: the two runs of the CUDA path -------
# printed rows: iter pcost dcost gap pres dres k/t mu step
traceA = [
    [0, 8.1242e-01, 2.9347e+00, 2.12e+00, 1.47e-16, 6.02e-01, 1.00e+00, 1.39e+00, None],
    [1, 7.5737e-01, 7.9652e-01, 3.92e-02, 7.07e-17, 3.79e-02, 4.87e-02, 8.03e-02, 9.62e-01],
    [2, 7.5007e-01, 7.4199e-01, 8.08e-03, 3.21e-16, 4.08e-04, 4.02e-03, 5.89e-03, 9.90e-01],
    [3, 7.5000e-01, 7.4982e-01, 1.77e-04, 3.19e-16, 4.10e-06, 8.02e-05, 1.15e-04, 9.90e-01],
    [4, 7.5000e-01, 7.5000e-01, 1.81e-06, 3.75e-16, 4.10e-08, 8.19e-07, 1.17e-06, 9.90e-01],
    [5, 7.5000e-01, 7.5000e-01, 1.81e-08, 9.86e-15, 4.10e-10, 8.19e-09, 1.17e-08, 9.90e-01],
    [6, 7.5000e-01, 7.5000e-01, 1.81e-10, 5.64e-16, 4.10e-12, 8.19e-11, 1.17e-10, 9.90e-01],
]
traceB = [
    [0, 6.6377e+00, 1.6932e+01, 1.55e+00, 2.85e-16, 5.91e-01, 1.00e+00, 4.42e+00, None],
    [1, 5.6713e+00, 5.0140e+00, 1.31e-01, 1.21e-16, 1.48e-02, 7.86e-02, 4.46e-01, 9.90e-01],
    [2, 5.5128e+00, 5.3393e+00, 3.25e-02, 5.83e-16, 1.77e-04, 1.70e-02, 6.06e-02, 9.90e-01],
    [3, 5.5002e+00, 5.4807e+00, 3.56e-03, 4.30e-16, 1.91e-06, 1.98e-03, 5.52e-03, 9.90e-01],
    [4, 5.5000e+00, 5.4994e+00, 1.01e-04, 2.43e-15, 1.94e-08, 5.69e-05, 1.52e-04, 9.90e-01],
    [5, 5.5000e+00, 5.5000e+00, 1.07e-06, 1.70e-15, 1.94e-10, 6.02e-07, 1.60e-06, 9.90e-01],
    [6, 5.5000e+00, 5.5000e+00, 1.07e-08, 9.06e-15, 1.94e-12, 6.03e-09, 1.60e-08, 9.90e-01],
    [7, 5.5000e+00, 5.5000e+00, 1.07e-10, 3.17e-16, 1.96e-14, 6.03e-11, 1.60e-10, 9.90e-01],
]
PA = [[2.0, 1, 0], [1, 2, 0], [0, 0, 2]]
AA = [[1.0, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1]]
add("notebook_A", "src/python/call_in_python.ipynb cell 2 (problem :89-108, log :60-71)", PA, [0.0, -1, -1], AA,
    [1.0, 0, 0, 0], [[Z, 1], [SOC, 3]], "SOLVED", x=[1.0, 0.0, 0.5], obj=0.75, iterations=6, trace=traceA)
add("notebook_B", "src/python/call_in_python.ipynb cells 4-5 (updates :186-214, log :154-166, x :226)", PA,
    [0.0, -1, -1], AA, [1.0, 0, 0, 0], [[Z, 1], [SOC, 3]], "SOLVED", x=[2.0, -1.52972138e-10, 9.99999999e-01],
    obj=5.5, iterations=7, trace=traceB, update_b=[2.0, 1, 1, 1],
    update_P=[[3.0, 0.5, 0], [0.5, 2, 0], [0, 0, 1]], update_P_without_cscale=1)

out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_cases.json")
with open(out, "w") as f:
    json.dump(cases, f, indent=1)
print("wrote", out, len(cases), "cases")
