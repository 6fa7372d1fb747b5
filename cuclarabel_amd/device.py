"""Device arrays in and out of the solver.

The counterpart of the reference's Python bridge (src/python/py2jl.jl:1-26
wraps CuPy device pointers as CuVector / CuSparseMatrixCSR for the Julia
solver; src/python/jl2py.py hands the solver's CuVectors back as CuPy arrays
without a copy).  Here the problem comes as device arrays -- any object that
exposes its device pointer through ``__cuda_array_interface__`` (CuPy, Numba,
PyTorch) or ``data_ptr()`` (PyTorch) -- and the solution goes into
caller-owned device arrays (device to device).  :class:`DeviceArray` is a
device allocation of the library's own for callers without a GPU array
library; :class:`DeviceVector` names a raw device buffer so that CuPy or
PyTorch can wrap it.  Nothing here imports a GPU array library.

    P = DeviceCSC(indptr, indices, data, (n, n))      # device arrays, 0-based int32 / float64
    s = DeviceSolver(P, q_dev, A, b_dev, cones, Settings(direct_solve_method="hip"))
    sol = s.solve()                                   # DefaultSolution (host copies, as Solver)
    s.solution_to(x_dev, s_dev, z_dev)                # device-to-device
    s.update_q(q2_dev); s.solve()                     # update_q! from device memory
"""
from __future__ import annotations

import ctypes as C
from types import SimpleNamespace
from typing import Optional, Sequence

import numpy as np

from . import Settings, Solver, _Cone, _lib

_TYPESTR = {np.dtype(np.float64): "<f8", np.dtype(np.int32): "<i4"}


def device_pointer(a, dtype, length: int) -> int:
    """The device address of `a`, checked to hold `length` contiguous
    elements of `dtype` (float64 or int32)."""
    dtype = np.dtype(dtype)
    iface = getattr(a, "__cuda_array_interface__", None)
    if iface is not None:
        if iface.get("typestr") != _TYPESTR[dtype]:
            raise ValueError(f"device array of type {iface.get('typestr')}, expected {_TYPESTR[dtype]}")
        shape = tuple(iface["shape"])
        if int(np.prod(shape)) != length:
            raise ValueError(f"device array of {int(np.prod(shape))} elements, expected {length}")
        strides = iface.get("strides")
        if strides is not None and len(shape) == 1 and strides[0] != dtype.itemsize:
            raise ValueError("device array must be contiguous")
        return int(iface["data"][0])
    if hasattr(a, "data_ptr"):  # a PyTorch tensor without the interface
        if not a.is_contiguous() or a.numel() != length or a.element_size() != dtype.itemsize:
            raise ValueError(f"device tensor must be contiguous with {length} elements of {dtype}")
        if ("float" in str(a.dtype)) != (dtype.kind == "f"):
            raise ValueError(f"device tensor of type {a.dtype}, expected {dtype}")
        return int(a.data_ptr())
    raise TypeError("expected a device array (__cuda_array_interface__ or data_ptr())")


class DeviceVector:
    """A float64 device vector by address, not owned (as jl2py.py hands a
    CuVector's pointer to CuPy); ``__cuda_array_interface__`` lets CuPy or
    PyTorch wrap it and the solver's device entry points take it."""

    def __init__(self, ptr: int, length: int):
        self.ptr, self.length = int(ptr), int(length)

    @property
    def __cuda_array_interface__(self):
        return {"shape": (self.length,), "typestr": "<f8", "data": (self.ptr, False), "version": 3,
                "strides": None}


class DeviceArray:
    """A one-dimensional array in device memory allocated by the library
    (clh_device_alloc), for callers without a GPU array library of their own
    (float64 or int32); ``__cuda_array_interface__`` lets the device entry
    points and CuPy or PyTorch take it."""

    def __init__(self, length: int, dtype=np.float64, _library=None):
        self._L = _library if _library is not None else _lib.lib()
        self.dtype = np.dtype(dtype)
        if self.dtype not in _TYPESTR:
            raise ValueError("DeviceArray holds float64 or int32")
        self.length = int(length)
        self._ptr = C.c_void_p()
        _lib.check(self._L, self._L.clh_device_alloc(C.byref(self._ptr), max(1, self.length) * self.dtype.itemsize))

    @classmethod
    def from_host(cls, a, dtype=None, _library=None) -> "DeviceArray":
        h = np.ascontiguousarray(a, dtype=dtype if dtype is not None else np.asarray(a).dtype).ravel()
        d = cls(h.size, h.dtype, _library)
        if h.size:
            _lib.check(d._L, d._L.clh_memcpy_to_device(d._ptr, h.ctypes.data_as(C.c_void_p), h.nbytes))
        return d

    def to_host(self) -> np.ndarray:
        h = np.empty(self.length, self.dtype)
        if self.length:
            _lib.check(self._L, self._L.clh_memcpy_to_host(h.ctypes.data_as(C.c_void_p), self._ptr, h.nbytes))
        return h

    @property
    def __cuda_array_interface__(self):
        return {"shape": (self.length,), "typestr": _TYPESTR[self.dtype], "data": (int(self._ptr.value), False),
                "version": 3, "strides": None}

    def free(self):
        if getattr(self, "_ptr", None) is not None and self._ptr.value:
            self._L.clh_device_free(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        self.free()


class DeviceCSC:
    """A CSC matrix in device memory: column pointers and row indices (0-based
    int32) and values (float64), as py2jl.jl's cupy_to_cucsrmat takes them."""

    def __init__(self, indptr, indices, data, shape):
        self.indptr, self.indices, self.data = indptr, indices, data
        self.shape = (int(shape[0]), int(shape[1]))


class DeviceSolver(Solver):
    """Clarabel.Solver(P,q,A,b,cones,settings) from device arrays
    (clh_solver_create_device).  The cones and settings are host objects."""

    def __init__(self, P: DeviceCSC, q, A: DeviceCSC, b, cones: Sequence[_Cone], settings: Optional[Settings] = None,
                 nnz_P: Optional[int] = None, nnz_A: Optional[int] = None, _library=None):
        self._L = _library if _library is not None else _lib.lib()
        self.settings = settings or Settings()
        n, m = P.shape[1], A.shape[0]
        if P.shape != (n, n) or A.shape[1] != n:
            raise ValueError("P must be n x n and A m x n")
        vp = C.c_void_p
        # the pattern sizes: the device arrays' own lengths
        nnz_P = _length(P.indices) if nnz_P is None else int(nnz_P)
        nnz_A = _length(A.indices) if nnz_A is None else int(nnz_A)
        ptrs = [device_pointer(P.indptr, np.int32, n + 1), device_pointer(P.indices, np.int32, nnz_P),
                device_pointer(P.data, np.float64, nnz_P), device_pointer(q, np.float64, n),
                device_pointer(A.indptr, np.int32, n + 1), device_pointer(A.indices, np.int32, nnz_A),
                device_pointer(A.data, np.float64, nnz_A), device_pointer(b, np.float64, m)]
        from . import _numel
        self.cones = list(cones)
        if sum(_numel(c) for c in self.cones) != m:
            raise ValueError("cone dimensions do not match the number of rows of A")
        ct = np.array([c.code for c in self.cones], np.int32)
        cd = np.array([c.dim for c in self.cones], np.int32)
        cp = np.array([float(c.param) for c in self.cones], np.float64)
        self._h = C.c_void_p()
        cs = self.settings.to_c()
        _lib.check(self._L, self._L.clh_solver_create_device(
            C.byref(self._h), n, m, *[vp(p) for p in ptrs], len(ct), ct.ctypes.data_as(C.POINTER(C.c_int32)),
            cd.ctypes.data_as(C.POINTER(C.c_int32)), cp.ctypes.data_as(C.POINTER(C.c_double)), C.byref(cs)))
        self._n, self._m, self._nnz_P, self._nnz_A = n, m, nnz_P, nnz_A
        self._p = SimpleNamespace(n=n, m=m)  # what Solver.solve reads
        self.solution = None

    def solution_to(self, x=None, s=None, z=None, stream=None):
        """Copy the last solution into device arrays (device to device)."""
        vp = C.c_void_p
        px = vp(device_pointer(x, np.float64, self._n)) if x is not None else None
        ps = vp(device_pointer(s, np.float64, self._m)) if s is not None else None
        pz = vp(device_pointer(z, np.float64, self._m)) if z is not None else None
        _lib.check(self._L, self._L.clh_get_solution_device(self._h, px, ps, pz, stream))

    def update_P(self, Px):
        _lib.check(self._L, self._L.clh_update_P_device(self._h, C.c_void_p(device_pointer(Px, np.float64, self._nnz_P))))

    def update_q(self, q):
        _lib.check(self._L, self._L.clh_update_q_device(self._h, C.c_void_p(device_pointer(q, np.float64, self._n))))

    def update_A(self, Ax):
        _lib.check(self._L, self._L.clh_update_A_device(self._h, C.c_void_p(device_pointer(Ax, np.float64, self._nnz_A))))

    def update_b(self, b):
        _lib.check(self._L, self._L.clh_update_b_device(self._h, C.c_void_p(device_pointer(b, np.float64, self._m))))


def _length(a) -> int:
    iface = getattr(a, "__cuda_array_interface__", None)
    if iface is not None:
        return int(np.prod(iface["shape"]))
    if hasattr(a, "numel"):
        return int(a.numel())
    raise TypeError("expected a device array (__cuda_array_interface__ or numel())")
