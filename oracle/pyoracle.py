"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The cuclarabel_amd package never imports it.

Cones are given as a sequence of objects with attributes ``code``, ``dim`` and
``param`` (the package's cone classes have them) or as tuples
``(code, dim[, param])`` with codes 0 zero, 1 nonneg, 2 soc, 3 exp, 4 pow,
5 psd (dim = side of the matrix for psd).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

STATUS = [
    "UNSOLVED", "SOLVED", "PRIMAL_INFEASIBLE", "DUAL_INFEASIBLE", "ALMOST_SOLVED",
    "ALMOST_PRIMAL_INFEASIBLE", "ALMOST_DUAL_INFEASIBLE", "MAX_ITERATIONS", "MAX_TIME",
    "NUMERICAL_ERROR", "INSUFFICIENT_PROGRESS",
]


class OrcSettings(C.Structure):
    _fields_ = [
        ("max_iter", C.c_int32), ("time_limit", C.c_double), ("verbose", C.c_int32),
        ("max_step_fraction", C.c_double),
        ("tol_gap_abs", C.c_double), ("tol_gap_rel", C.c_double), ("tol_feas", C.c_double),
        ("tol_infeas_abs", C.c_double), ("tol_infeas_rel", C.c_double), ("tol_ktratio", C.c_double),
        ("reduced_tol_gap_abs", C.c_double), ("reduced_tol_gap_rel", C.c_double),
        ("reduced_tol_feas", C.c_double), ("reduced_tol_infeas_abs", C.c_double),
        ("reduced_tol_infeas_rel", C.c_double), ("reduced_tol_ktratio", C.c_double),
        ("equilibrate_enable", C.c_int32), ("equilibrate_max_iter", C.c_int32),
        ("equilibrate_min_scaling", C.c_double), ("equilibrate_max_scaling", C.c_double),
        ("linesearch_backtrack_step", C.c_double), ("min_switch_step_length", C.c_double),
        ("min_terminate_step_length", C.c_double),
        ("static_regularization_enable", C.c_int32), ("static_regularization_constant", C.c_double),
        ("static_regularization_proportional", C.c_double),
        ("dynamic_regularization_enable", C.c_int32), ("dynamic_regularization_eps", C.c_double),
        ("dynamic_regularization_delta", C.c_double),
        ("iterative_refinement_enable", C.c_int32), ("iterative_refinement_reltol", C.c_double),
        ("iterative_refinement_abstol", C.c_double), ("iterative_refinement_max_iter", C.c_int32),
        ("iterative_refinement_stop_ratio", C.c_double),
        ("update_P_without_cscale", C.c_int32),
        ("dynamic_regularization_keep_magnitude", C.c_int32),
    ]


class OrcResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("iterations", C.c_int32),
        ("obj_val", C.c_double), ("obj_val_dual", C.c_double),
        ("r_prim", C.c_double), ("r_dual", C.c_double),
        ("solve_time", C.c_double), ("setup_time", C.c_double),
        ("trace_len", C.c_int32), ("nnz_L", C.c_int64), ("flops_factor", C.c_double),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"], check=True)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.orc_default_settings.argtypes = [C.POINTER(OrcSettings)]
        L.orc_solve.argtypes = [C.c_int32, C.c_int32, ip, ip, dp, dp, ip, ip, dp, dp, C.c_int32, ip, ip, dp,
                                C.POINTER(OrcSettings), dp, dp, dp, C.POINTER(OrcResult), dp]
        L.orc_solve_update.argtypes = [C.c_int32, C.c_int32, ip, ip, dp, dp, ip, ip, dp, dp, C.c_int32, ip, ip,
                                       dp, C.POINTER(OrcSettings), dp, dp, dp, dp, dp, dp, dp,
                                       C.POINTER(OrcResult), dp]
        L.orc_wright_omega.restype = C.c_double
        L.orc_wright_omega.argtypes = [C.c_double]
        L.orc_exp_gradient_primal.argtypes = [dp, dp]
        L.orc_exp_dual_grad_H.argtypes = [dp, dp, dp]
        L.orc_pow_gradient_primal.argtypes = [dp, C.c_double, dp]
        L.orc_pow_dual_grad_H.argtypes = [dp, C.c_double, dp, dp]
        L.orc_soc_update_scaling.argtypes = [C.c_int32, dp, dp, dp, dp, dp, dp, dp, dp]
        L.orc_soc_step_length_component.restype = C.c_double
        L.orc_soc_step_length_component.argtypes = [C.c_int32, dp, dp, C.c_double]
        L.orc_psd_update_scaling.argtypes = [C.c_int32, dp, dp, dp, dp, dp, dp]
        L.orc_sym_eigvals.argtypes = [C.c_int32, dp, dp]
        L.orc_nonsym_higher_correction.argtypes = [C.c_int32, C.c_double, dp, dp, dp, dp]
        L.orc_nonsym_update_Hs.argtypes = [C.c_int32, C.c_double, dp, dp, C.c_double, C.c_int32, dp]
        L.orc_ldl_solve.argtypes = [C.c_int32, ip, ip, dp, ip, C.c_double, C.c_double, C.c_int32, dp, dp,
                                    C.POINTER(C.c_int64)]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def settings(**kw) -> OrcSettings:
    s = OrcSettings()
    lib().orc_default_settings(C.byref(s))
    for k, v in kw.items():
        if not hasattr(s, k):
            raise KeyError(k)
        setattr(s, k, v)
    return s


def _cone_arrays(cones):
    t, d, p = [], [], []
    for c in cones:
        if isinstance(c, tuple):
            t.append(int(c[0]))
            d.append(int(c[1]) if len(c) > 1 else 3)
            p.append(float(c[2]) if len(c) > 2 else 0.0)
        else:
            t.append(int(c.code))
            d.append(int(c.dim))
            p.append(float(getattr(c, "param", 0.0) or 0.0))
    return (np.asarray(t, np.int32), np.asarray(d, np.int32), np.asarray(p, np.float64))


def _csc(M, shape):
    M = sp.csc_matrix(M, shape=shape)
    M.sort_indices()
    return (np.asarray(M.indptr, np.int32), np.asarray(M.indices, np.int32),
            np.asarray(M.data, np.float64))


def solve(P, q, A, b, cones, update_P=None, update_b=None, update_q=None, update_A=None, **kw):
    """Run the oracle.  Returns a dict with x, s, z, status, iterations,
    obj_val, obj_val_dual, r_prim, r_dual, solve_time, trace (iters+1 x 8).
    With update_* given: solve, apply the updates (P and A as matrices of the
    original patterns, q and b as vectors) and solve again; the result is the
    second solve's."""
    q = np.ascontiguousarray(q, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    n, m = q.size, b.size
    Pp, Pi, Px = _csc(P, (n, n))
    Ap, Ai, Ax = _csc(A, (m, n))
    ct, cd, cp = _cone_arrays(cones)
    st = settings(**kw)
    x = np.zeros(n)
    s = np.zeros(m)
    z = np.zeros(m)
    r = OrcResult()
    trace = np.zeros(8 * (int(st.max_iter) + 64))
    if update_P is None and update_b is None and update_q is None and update_A is None:
        lib().orc_solve(n, m, _ip(Pp), _ip(Pi), _dp(Px), _dp(q), _ip(Ap), _ip(Ai), _dp(Ax), _dp(b),
                        len(ct), _ip(ct), _ip(cd), _dp(cp), C.byref(st), _dp(x), _dp(s), _dp(z), C.byref(r),
                        _dp(trace))
    else:
        Pn = None
        if update_P is not None:
            # values aligned with the original P pattern
            Pn_mat = sp.csc_matrix(update_P, shape=(n, n))
            Pn_mat.sort_indices()
            if not (np.array_equal(Pn_mat.indptr, Pp) and np.array_equal(Pn_mat.indices, Pi)):
                raise ValueError("update_P must match the sparsity pattern of P")
            Pn = np.ascontiguousarray(Pn_mat.data, np.float64)
        bn = None if update_b is None else np.ascontiguousarray(update_b, np.float64)
        qn = None if update_q is None else np.ascontiguousarray(update_q, np.float64)
        An = None
        if update_A is not None:
            An_mat = sp.csc_matrix(update_A, shape=(m, n))
            An_mat.sort_indices()
            if not (np.array_equal(An_mat.indptr, Ap) and np.array_equal(An_mat.indices, Ai)):
                raise ValueError("update_A must match the sparsity pattern of A")
            An = np.ascontiguousarray(An_mat.data, np.float64)
        opt = lambda a: _dp(a) if a is not None else None  # noqa: E731
        lib().orc_solve_update(n, m, _ip(Pp), _ip(Pi), _dp(Px), _dp(q), _ip(Ap), _ip(Ai), _dp(Ax), _dp(b),
                               len(ct), _ip(ct), _ip(cd), _dp(cp), C.byref(st),
                               opt(Pn), opt(bn), opt(qn), opt(An),
                               _dp(x), _dp(s), _dp(z), C.byref(r), _dp(trace))
    tl = r.trace_len
    return dict(x=x, s=s, z=z, status=STATUS[r.status], status_code=r.status, iterations=r.iterations,
                obj_val=r.obj_val, obj_val_dual=r.obj_val_dual, r_prim=r.r_prim, r_dual=r.r_dual,
                solve_time=r.solve_time, setup_time=r.setup_time, nnz_L=r.nnz_L, flops_factor=r.flops_factor,
                trace=trace[: 8 * tl].reshape(tl, 8).copy())


# ---- component wrappers -------------------------------------------------

def wright_omega(z: float) -> float:
    return lib().orc_wright_omega(float(z))


def exp_gradient_primal(s):
    s = np.ascontiguousarray(s, np.float64)
    g = np.zeros(3)
    lib().orc_exp_gradient_primal(_dp(s), _dp(g))
    return g


def exp_dual_grad_H(z):
    z = np.ascontiguousarray(z, np.float64)
    g = np.zeros(3)
    H = np.zeros(9)
    lib().orc_exp_dual_grad_H(_dp(z), _dp(g), _dp(H))
    return g, H.reshape(3, 3, order="F")


def pow_gradient_primal(s, alpha):
    s = np.ascontiguousarray(s, np.float64)
    g = np.zeros(3)
    lib().orc_pow_gradient_primal(_dp(s), float(alpha), _dp(g))
    return g


def pow_dual_grad_H(z, alpha):
    z = np.ascontiguousarray(z, np.float64)
    g = np.zeros(3)
    H = np.zeros(9)
    lib().orc_pow_dual_grad_H(_dp(z), float(alpha), _dp(g), _dp(H))
    return g, H.reshape(3, 3, order="F")


def soc_update_scaling(s, z):
    s = np.ascontiguousarray(s, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    d = s.size
    w, lam, u, v = (np.zeros(d) for _ in range(4))
    eta = C.c_double()
    dd = C.c_double()
    ok = lib().orc_soc_update_scaling(d, _dp(s), _dp(z), _dp(w), _dp(lam), C.byref(eta), C.byref(dd), _dp(u),
                                      _dp(v))
    return dict(ok=bool(ok), w=w, lam=lam, eta=eta.value, d=dd.value, u=u, v=v)


def soc_step_length_component(x, y, amax):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return lib().orc_soc_step_length_component(x.size, _dp(x), _dp(y), float(amax))


def psd_update_scaling(n, s, z):
    s = np.ascontiguousarray(s, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    t = n * (n + 1) // 2
    R = np.zeros(n * n)
    Rinv = np.zeros(n * n)
    lam = np.zeros(n)
    Hs = np.zeros(t * t)
    ok = lib().orc_psd_update_scaling(n, _dp(s), _dp(z), _dp(R), _dp(Rinv), _dp(lam), _dp(Hs))
    return dict(ok=bool(ok), R=R.reshape(n, n, order="F"), Rinv=Rinv.reshape(n, n, order="F"), lam=lam,
                Hs=Hs.reshape(t, t, order="F"))


def sym_eigvals(M):
    M = np.asfortranarray(M, np.float64)
    n = M.shape[0]
    e = np.zeros(n)
    lib().orc_sym_eigvals(n, _dp(M), _dp(e))
    return e


def nonsym_higher_correction(code, z, ds, v, alpha=0.0):
    """eta(ds, v) of an exponential (code 3) or power (code 4) cone at dual z"""
    z, ds, v = (np.ascontiguousarray(t, np.float64) for t in (z, ds, v))
    eta = np.zeros(3)
    lib().orc_nonsym_higher_correction(int(code), float(alpha), _dp(z), _dp(ds), _dp(v), _dp(eta))
    return eta


def nonsym_update_Hs(code, s, z, mu, strategy=0, alpha=0.0):
    """Hs (3x3) of an exponential (code 3) or power (code 4) cone"""
    s, z = (np.ascontiguousarray(t, np.float64) for t in (s, z))
    Hs = np.zeros(9)
    lib().orc_nonsym_update_Hs(int(code), float(alpha), _dp(s), _dp(z), float(mu), int(strategy), _dp(Hs))
    return Hs.reshape(3, 3, order="F")


def ldl_solve(K_triu, dsigns, b, eps=1e-13, delta=2e-7, keep_magnitude=False):
    n = K_triu.shape[0]
    Kp, Ki, Kx = _csc(K_triu, (n, n))
    ds = np.ascontiguousarray(dsigns, np.int32)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n)
    nnzL = C.c_int64()
    rc = lib().orc_ldl_solve(n, _ip(Kp), _ip(Ki), _dp(Kx), _ip(ds), eps, delta, int(keep_magnitude), _dp(b), _dp(x), C.byref(nnzL))
    if rc != 0:
        raise RuntimeError("LDL failed")
    return x, nnzL.value
