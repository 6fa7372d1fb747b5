"""The LDL' plugin (clh_ldl_*): the device solver of one quasi-definite matrix
at the reference's LDL solver seam (directldl_cudss.jl:33-57 refactor!/solve!,
the refinement of kktsolver_directldl_gpu.jl:349-428).

What the mathematics fixes, checked on KKT matrices K = [P + d1 I, A'; A, -W]
with the sparsity of MPC and portfolio problems:
  * solve: the normwise backward error of x against the regularized matrix
    K + eps * diag(dsigns) is at round-off level (FP64 factors) or FP32
    round-off (mixed), and x agrees with a dense LU solve;
  * solve_refine: refined against the unregularized K the residual meets the
    tolerance, also from FP32 factors;
  * inertia: the pivot signs count the eigenvalue signs of K (Sylvester);
  * dynamic regularization: a structurally singular K factors only with it,
    and the regularized pivots are counted;
  * new values on the same pattern reuse the symbolic analysis.
The emulated build runs the engine source serially on the CPU; the gpu-marked
tests call the HIP library through the same C ABI."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from cuclarabel_amd import problems
from cuclarabel_amd.ldl import LDLSolver

from helpers import build_emu


def kkt_matrix(kind, seed):
    """(K full symmetric CSR, dsigns) of the KKT system of a problem family."""
    rng = np.random.RandomState(seed)
    if kind == "mpc":
        P, q, A, b, cones = problems.mpc(seed, T=12 + 4 * seed)
    elif kind == "portfolio":
        P, q, A, b, cones = problems.portfolio(seed)
    elif kind == "socp":  # random 8 entries per row: fill that leaves a dense root of hundreds of columns
        P, q, A, b, cones = problems.large_socp(seed, n=800, m_zero=200, m_nn=600, n_soc=70)
    else:  # random sparse QP
        n, m = 40, 60
        R = sp.random(n, n, density=0.08, random_state=rng)
        P = (R @ R.T + sp.eye(n) * 0.1).tocsc()
        A = sp.random(m, n, density=0.1, random_state=rng).tocsc()
    n, m = P.shape[0], A.shape[0]
    W = sp.diags(rng.uniform(0.05, 20.0, m))
    K = sp.bmat([[sp.csc_matrix(P) + 1e-8 * sp.eye(n), A.T], [A, -W]]).tocsr()
    K = ((K + K.T) * 0.5).tocsr()  # full symmetric pattern, the triu of P mirrored
    K.sort_indices()
    dsigns = np.r_[np.ones(n), -np.ones(m)].astype(np.int32)
    return K, dsigns


KINDS = [("mpc", 0), ("mpc", 3), ("portfolio", 1), ("random", 5)]


def _backward_error(K, x, b):
    r = K @ x - b
    return np.abs(r).max() / (abs(K).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())


def check_solver(lib, kind, seed, method):
    K, dsigns = kkt_matrix(kind, seed)
    n = K.shape[0]
    rng = np.random.RandomState(100 + seed)
    b = rng.standard_normal(n)
    eps = 1e-8
    Kreg = (K + eps * sp.diags(dsigns.astype(float))).toarray()
    s = LDLSolver(K, dsigns, method=method, _library=lib)
    try:
        assert s.refactor(static_eps=eps)
        info = s.info()
        evals = np.linalg.eigvalsh(Kreg)
        assert info["n_positive"] == int((evals > 0).sum()) == int((dsigns > 0).sum())
        assert info["n_negative"] == int((evals < 0).sum())
        assert info["last_dynreg"] == 0
        assert info["nnz_L"] > 0 and info["n_supernodes"] >= 1 and info["n_levels"] >= 1
        x = s.solve(b)
        xd = np.linalg.solve(Kreg, b)
        be = _backward_error(Kreg, x, b)
        if method == "hip":
            assert be < 1e-14, be
            np.testing.assert_allclose(x, xd, rtol=1e-8, atol=1e-10 * np.abs(xd).max())
        else:
            assert be < 1e-5, be
        # refinement against the unregularized matrix
        xr = s.solve_refine(b, reltol=1e-12, abstol=1e-12, maxiter=20, stopratio=1.0)
        res = np.abs(K @ xr - b).max()
        info = s.info()
        assert info["last_refine_error"] == pytest.approx(res, rel=0.5, abs=1e-15)
        assert res <= 1e-12 + 1e-12 * np.abs(b).max() or method == "hipmixed" and res < 1e-10, res
        if method == "hipmixed":
            assert info["last_ir_steps"] >= 1
        # new values on the same pattern: a congruence S K S keeps the inertia
        S = sp.diags(rng.uniform(0.5, 2.0, n))
        K2 = (S @ K @ S).tocsr()
        K2.sort_indices()
        assert np.array_equal(K2.indices, K.indices) and np.array_equal(K2.indptr, K.indptr)
        assert s.refactor(K2.data, static_eps=0.0)
        assert s.info()["last_dynreg"] == 0
        x2 = s.solve(b)
        be2 = _backward_error(K2, x2, b)
        assert be2 < (1e-14 if method == "hip" else 1e-5), be2
    finally:
        s.close()


def _solver_for(kind, seed, method, lib):
    K, dsigns = kkt_matrix(kind, seed)
    return LDLSolver(K, dsigns, method=method, _library=lib)


def _with_diagonal(Kd):
    """CSR of a dense symmetric matrix that stores every diagonal entry, zeros too."""
    N = Kd.shape[0]
    r, c = np.nonzero(Kd)
    keep = r != c
    r = np.r_[r[keep], np.arange(N)]
    c = np.r_[c[keep], np.arange(N)]
    K = sp.coo_matrix((Kd[r, c], (r, c)), shape=Kd.shape).tocsr()
    K.sort_indices()
    return K


def check_dynamic_regularization(lib, method):
    # column 0 of A and row 0 of P empty: the first pivot is exactly zero
    n, m = 6, 4
    rng = np.random.RandomState(7)
    P = np.diag(np.r_[0.0, rng.uniform(1, 2, n - 1)])
    A = rng.standard_normal((m, n))
    A[:, 0] = 0
    K = _with_diagonal(np.block([[P, A.T], [A, -np.eye(m)]]))
    dsigns = np.r_[np.ones(n), -np.ones(m)].astype(np.int32)
    s = LDLSolver(K, dsigns, method=method, _library=lib)
    try:
        ok = s.refactor(static_eps=0.0, dynamic=True, dyn_eps=1e-13, dyn_delta=2e-7)
        assert ok
        assert s.info()["last_dynreg"] >= 1
        assert s.info()["n_positive"] == n and s.info()["n_negative"] == m
        ok = s.refactor(static_eps=0.0, dynamic=False)
        assert not ok
    finally:
        s.close()


def check_validation(lib):
    K, dsigns = kkt_matrix("random", 5)
    Kb = K.tolil()
    Kb[3, 3] = 0
    Kb = Kb.tocsr()
    Kb.eliminate_zeros()
    with pytest.raises(ValueError, match="diagonal"):
        LDLSolver(Kb, dsigns, _library=lib)
    with pytest.raises(ValueError, match="sign"):
        LDLSolver(K, np.zeros_like(dsigns), _library=lib)
    with pytest.raises(ValueError):
        LDLSolver(K, dsigns, method="cudss", _library=lib)
    with pytest.raises(ValueError):
        LDLSolver(sp.csr_matrix(np.ones((3, 4))), np.ones(3), _library=lib)


# ---- CPU: the engine source compiled serially -----------------------------

@pytest.fixture(scope="module")
def emu():
    return build_emu()


@pytest.mark.parametrize("method", ["hip", "hipmixed"])
@pytest.mark.parametrize("kind,seed", KINDS)
def test_emu_solve(emu, kind, seed, method):
    check_solver(emu, kind, seed, method)


@pytest.mark.parametrize("method", ["hip", "hipmixed"])
def test_emu_dynamic_regularization(emu, method):
    check_dynamic_regularization(emu, method)


@pytest.mark.parametrize("method", ["hip", "hipmixed"])
@pytest.mark.parametrize("kind,seed", KINDS[:3])
def test_emu_solve_grid_path(emu, monkeypatch, kind, seed, method):
    """the grid engine's LDL code paths (CLH_LDL_GRID forces them for small
    matrices), serially"""
    monkeypatch.setenv("CLH_LDL_GRID", "1")
    check_solver(emu, kind, seed, method)


def test_emu_validation(emu):
    check_validation(emu)


def test_emu_one_based_indices(emu):
    import ctypes as C
    K, dsigns = kkt_matrix("random", 5)
    h = C.c_void_p()
    rp = np.ascontiguousarray(K.indptr + 1, dtype=np.int32)
    cv = np.ascontiguousarray(K.indices + 1, dtype=np.int32)
    ip = C.POINTER(C.c_int32)
    rc = emu.clh_ldl_create(C.byref(h), K.shape[0], rp.ctypes.data_as(ip), cv.ctypes.data_as(ip), 1,
                            dsigns.ctypes.data_as(ip), 0)
    assert rc == 0, emu.clh_last_error()
    emu.clh_ldl_destroy(h)


# ---- GPU: the HIP library -------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("method", ["hip", "hipmixed"])
@pytest.mark.parametrize("kind,seed", KINDS)
def test_gpu_solve(kind, seed, method):
    print("ldl", kind, seed, method, flush=True)
    check_solver(None, kind, seed, method)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["hip", "hipmixed"])
def test_gpu_dense_root_on_the_grid(method):
    """A KKT matrix whose factor ends in a dense root of at least 128 columns
    (over 1e8 flops): the LDL operations run over a cooperative grid filling
    the GPU, the root by the whole-grid blocked factorization on the matrix
    cores (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32 tiles).  Backward
    error, agreement with a dense solve and inertia as for the small cases."""
    print("ldl dense root on the grid", method, flush=True)
    s = _solver_for("socp", 2, method, None)
    try:
        info = s.info()
        assert info["grid_blocks"] >= 256, info
        assert info["factor_flops"] >= 1e8
    finally:
        s.close()
    check_solver(None, "socp", 2, method)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["hip", "hipmixed"])
def test_gpu_dense_root_of_several_super_blocks(method):
    """A KKT matrix (random sparse SOCP, n = 2000, KKT 5000 rows) whose factor
    ends in a dense root of about 2000 columns: several 512-column
    super-blocks, so the factorization builds one diagonal-block inverse per
    super-block and the solves substitute block by block across them
    (forward through LDS-staged blocks of L, backward by wavefront dot
    products, each super-block's inverse read from its transposed copy).
    Backward error of the solve against the regularized matrix, agreement
    with a sparse LU solve, and the refined residual against K."""
    P, q, A, b, cones = problems.large_socp(2, n=2000, m_zero=500, m_nn=1500, n_soc=200)
    n, m = P.shape[0], A.shape[0]
    rng = np.random.RandomState(11)
    W = sp.diags(rng.uniform(0.05, 20.0, m))
    K = sp.bmat([[sp.csc_matrix(P) + 1e-8 * sp.eye(n), A.T], [A, -W]]).tocsr()
    K = ((K + K.T) * 0.5).tocsr()
    K.sort_indices()
    dsigns = np.r_[np.ones(n), -np.ones(m)].astype(np.int32)
    print("ldl dense root of several super-blocks", method, flush=True)
    s = LDLSolver(K, dsigns, method=method)
    try:
        eps = 1e-8
        assert s.refactor(static_eps=eps)
        info = s.info()
        assert info["grid_blocks"] >= 256 and info["factor_flops"] >= 1e10, info
        assert info["n_positive"] == n and info["n_negative"] == m
        b = rng.standard_normal(n + m)
        Kreg = (K + eps * sp.diags(dsigns.astype(float))).tocsc()
        x = s.solve(b)
        be = _backward_error(Kreg, x, b)
        if method == "hip":
            assert be < 1e-14, be
            xd = spla.spsolve(Kreg, b)
            np.testing.assert_allclose(x, xd, rtol=1e-7, atol=1e-9 * np.abs(xd).max())
        else:
            assert be < 1e-5, be
        xr = s.solve_refine(b, reltol=1e-12, abstol=1e-12, maxiter=20, stopratio=1.0)
        res = np.abs(K @ xr - b).max()
        assert res <= 1e-10 * (1 + np.abs(b).max()), res
    finally:
        s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["hip", "hipmixed"])
def test_gpu_dynamic_regularization(method):
    print("ldl dynreg", method, flush=True)
    check_dynamic_regularization(None, method)
