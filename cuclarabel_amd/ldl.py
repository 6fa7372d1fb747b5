"""The device LDL' solver on its own, at the seam of the reference's LDL solver
protocol (src/kktsolvers/direct-ldl/directldl_defaults.jl:1-38; the GPU
solvers src/kktsolvers/gpu/directldl_cudss.jl:5-58 and
directldl_mixed_cudss.jl:5-78):

    s = LDLSolver(K, dsigns, method="hip")   # ldlsolver_constructor: symbolic analysis once
    s.refactor(K.data, static_eps=1e-8)      # refactor!
    x = s.solve(b)                           # solve!
    x = s.solve_refine(b, K.data)            # the refinement of kktsolver_directldl_gpu.jl:349-428

K is a full symmetric scipy CSR matrix (the reference's GPU shape :full);
refactor() may take new values in K's CSR order.  The values and vectors are
copied to device buffers the solver owns; the *_dev methods take raw device
pointers instead (the C ABI's calling convention).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import scipy.sparse as sp

from . import _lib

_METHODS = {"hip": 0, "hipmixed": 1}


class _DeviceBuffer:
    def __init__(self, L, nbytes: int):
        self._L = L
        self.ptr = C.c_void_p()
        self.nbytes = max(8, nbytes)
        if L.clh_device_alloc(C.byref(self.ptr), self.nbytes) != 0:
            raise RuntimeError(L.clh_last_error().decode())

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if self._L.clh_memcpy_to_device(self.ptr, a.ctypes.data, a.nbytes) != 0:
            raise RuntimeError(self._L.clh_last_error().decode())

    def download(self, n: int, dtype=np.float64) -> np.ndarray:
        out = np.empty(n, dtype=dtype)
        if self._L.clh_memcpy_to_host(out.ctypes.data, self.ptr, out.nbytes) != 0:
            raise RuntimeError(self._L.clh_last_error().decode())
        return out

    def free(self):
        if self.ptr:
            self._L.clh_device_free(self.ptr)
            self.ptr = C.c_void_p()


class LDLSolver:
    def __init__(self, K, dsigns, method: str = "hip", _library=None):
        if method not in _METHODS:
            raise ValueError(f"method must be one of {sorted(_METHODS)}")
        self._L = _library if _library is not None else _lib.lib()
        K = sp.csr_matrix(K)
        K.sort_indices()
        n = K.shape[0]
        if K.shape != (n, n):
            raise ValueError("K must be square")
        self.n = n
        self.nnz = K.nnz
        dsigns = np.ascontiguousarray(dsigns, dtype=np.int32)
        if dsigns.shape != (n,):
            raise ValueError("dsigns must have one entry per row")
        rowptr = np.ascontiguousarray(K.indptr, dtype=np.int32)
        colval = np.ascontiguousarray(K.indices, dtype=np.int32)
        self._h = C.c_void_p()
        self._bufs = []
        rc = self._L.clh_ldl_create(C.byref(self._h), n, _lib_ptr(rowptr), _lib_ptr(colval), 0, _lib_ptr(dsigns),
                                    _METHODS[method])
        if rc != 0:
            raise ValueError(self._L.clh_last_error().decode())
        self._vals = self._buf(8 * self.nnz)
        self._x = self._buf(8 * n)
        self._b = self._buf(8 * n)
        self._vals.upload(np.asarray(K.data, dtype=np.float64))

    def _buf(self, nbytes):
        b = _DeviceBuffer(self._L, nbytes)
        self._bufs.append(b)
        return b

    def _check(self, rc):
        if rc < 0:
            raise ValueError(self._L.clh_last_error().decode())
        return rc == 0

    # refactor!(s): returns False on a zero or non-finite pivot, like the reference's Bool
    def refactor(self, vals: Optional[np.ndarray] = None, static_eps: float = 0.0, dynamic: bool = True,
                 dyn_eps: float = 1e-13, dyn_delta: float = 2e-7) -> bool:
        if vals is not None:
            vals = np.asarray(vals, dtype=np.float64)
            if vals.shape != (self.nnz,):
                raise ValueError("vals must hold one value per stored entry of K")
            self._vals.upload(vals)
        return self.refactor_dev(self._vals.ptr, static_eps, dynamic, dyn_eps, dyn_delta)

    def refactor_dev(self, vals_dev, static_eps=0.0, dynamic=True, dyn_eps=1e-13, dyn_delta=2e-7) -> bool:
        return self._check(self._L.clh_ldl_refactor(self._h, vals_dev, static_eps, 1 if dynamic else 0, dyn_eps,
                                                    dyn_delta, None))

    # solve!(s, x, b)
    def solve(self, b: np.ndarray) -> np.ndarray:
        b = np.asarray(b, dtype=np.float64)
        if b.shape != (self.n,):
            raise ValueError("b must have length n")
        self._b.upload(b)
        self.solve_dev(self._x.ptr, self._b.ptr)
        return self._x.download(self.n)

    def solve_dev(self, x_dev, b_dev):
        self._check(self._L.clh_ldl_solve(self._h, x_dev, b_dev, None))

    def solve_refine(self, b: np.ndarray, vals: Optional[np.ndarray] = None, reltol: float = 1e-12,
                     abstol: float = 1e-12, maxiter: int = 10, stopratio: float = 5.0) -> np.ndarray:
        """x = K^{-1} b refined against the unregularized values (those of the
        last refactor() unless vals is given); the defaults are the reference's
        iterative_refinement_* settings (src/settings.jl:128-131)."""
        b = np.asarray(b, dtype=np.float64)
        if b.shape != (self.n,):
            raise ValueError("b must have length n")
        if vals is not None:
            vals = np.asarray(vals, dtype=np.float64)
            if vals.shape != (self.nnz,):
                raise ValueError("vals must hold one value per stored entry of K")
            self._vals.upload(vals)
        self._b.upload(b)
        ok = self.solve_refine_dev(self._vals.ptr, self._x.ptr, self._b.ptr, reltol, abstol, maxiter, stopratio)
        if not ok:
            raise FloatingPointError(self._L.clh_last_error().decode())
        return self._x.download(self.n)

    def solve_refine_dev(self, vals_dev, x_dev, b_dev, reltol=1e-12, abstol=1e-12, maxiter=10, stopratio=5.0):
        return self._check(self._L.clh_ldl_solve_refine(self._h, vals_dev, x_dev, b_dev, reltol, abstol, maxiter,
                                                        stopratio, None))

    def info(self) -> dict:
        inf = _lib.ClhLdlInfo()
        self._check(self._L.clh_ldl_get_info(self._h, C.byref(inf)))
        return {f: getattr(inf, f) for f, _ in inf._fields_}

    def close(self):
        for b in self._bufs:
            b.free()
        self._bufs = []
        if self._h:
            self._L.clh_ldl_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _lib_ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))
