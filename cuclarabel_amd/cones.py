"""The composite cone's device operations on their own, for a caller that
drives its own interior-point loop: the reference's GPU composite cone
(src/gpucones/compositecone_type_gpu.jl:9-216) and its operations
(src/gpucones/coneops_compositecone_gpu.jl) through clh_cones_* of the C ABI.

    K = CompositeCone([NonnegativeConeT(3), SecondOrderConeT(4)])
    K.update_scaling(s, z, mu)          # update_scaling!
    y = K.mul_Hs(x)                     # mul_Hs!
    a = K.step_length(dz, ds, z, s)     # step_length

Vectors are numpy arrays of length m, copied to and from device buffers the
object owns; the *_dev entry points of the C ABI take device pointers.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import Settings, _Cone, _lib, _numel
from .ldl import _DeviceBuffer


class CompositeCone:
    def __init__(self, cones: Sequence[_Cone], settings: Optional[Settings] = None, _library=None):
        self._L = _library if _library is not None else _lib.lib()
        self.cones = list(cones)
        ct = np.array([c.code for c in self.cones], np.int32)
        cd = np.array([c.dim for c in self.cones], np.int32)
        cp = np.array([float(c.param) for c in self.cones], np.float64)
        cs = (settings or Settings()).to_c()
        self._h = C.c_void_p()
        rc = self._L.clh_cones_create(C.byref(self._h), len(ct), ct.ctypes.data_as(C.POINTER(C.c_int32)),
                                      cd.ctypes.data_as(C.POINTER(C.c_int32)),
                                      cp.ctypes.data_as(C.POINTER(C.c_double)), C.byref(cs))
        if rc != 0:
            raise ValueError(self._L.clh_last_error().decode())
        inf = _lib.ClhConesInfo()
        self._L.clh_cones_get_info(self._h, C.byref(inf))
        self.info = {f: getattr(inf, f) for f, _ in inf._fields_}
        self.m = self.info["m"]
        self._bufs = []
        self._v = [self._buf(8 * self.m) for _ in range(4)]
        self._o = [self._buf(8 * self.m) for _ in range(3)]
        self._hs = self._buf(8 * max(1, self.info["hs_len"]))
        self._sc = self._buf(16)

    def _buf(self, nbytes):
        b = _DeviceBuffer(self._L, nbytes)
        self._bufs.append(b)
        return b

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self._L.clh_last_error().decode())

    def _up(self, *vecs):
        ptrs = []
        for buf, v in zip(self._v, vecs):
            v = np.ascontiguousarray(v, np.float64)
            if v.shape != (self.m,):
                raise ValueError(f"vectors must have length {self.m}")
            buf.upload(v)
            ptrs.append(buf.ptr)
        return ptrs

    def _set_out(self, *vecs):
        ptrs = []
        for buf, v in zip(self._o, vecs):
            buf.upload(np.ascontiguousarray(v, np.float64))
            ptrs.append(buf.ptr)
        return ptrs

    def _scalars(self, k=1):
        return self._sc.download(k)

    @property
    def numel(self):
        return sum(_numel(c) for c in self.cones)

    # ---- operations (names of the reference's functions) ------------------
    def update_scaling(self, s, z, mu, scaling="primal_dual") -> bool:
        ps, pz = self._up(s, z)
        self._check(self._L.clh_cones_update_scaling(self._h, ps, pz, float(mu), 0 if scaling == "primal_dual" else 1,
                                                     self._sc.ptr, None))
        return bool(self._scalars()[0] > 0)

    def set_identity_scaling(self):
        self._check(self._L.clh_cones_set_identity_scaling(self._h, None))

    def get_Hs(self) -> np.ndarray:
        self._check(self._L.clh_cones_get_Hs(self._h, self._hs.ptr, None))
        return self._hs.download(self.info["hs_len"])

    def mul_Hs(self, x) -> np.ndarray:
        (px,) = self._up(x)
        self._check(self._L.clh_cones_mul_Hs(self._h, self._o[0].ptr, px, None))
        return self._o[0].download(self.m)

    def affine_ds(self, s) -> np.ndarray:
        (ps,) = self._up(s)
        self._check(self._L.clh_cones_affine_ds(self._h, self._o[0].ptr, ps, None))
        return self._o[0].download(self.m)

    def combined_ds_shift(self, step_z, step_s, sigma_mu):
        """(shift, step_z, step_s): step_z and step_s as the operation leaves them."""
        _, pz, ps = self._set_out(np.zeros(self.m), step_z, step_s)
        self._check(self._L.clh_cones_combined_ds_shift(self._h, self._o[0].ptr, pz, ps, float(sigma_mu), None))
        return tuple(b.download(self.m) for b in self._o)

    def ds_from_dz_offset(self, ds, z) -> np.ndarray:
        pds, pz = self._up(ds, z)
        self._check(self._L.clh_cones_ds_from_dz_offset(self._h, self._o[0].ptr, pds, pz, None))
        return self._o[0].download(self.m)

    def step_length(self, dz, ds, z, s, alpha_max=1.0) -> float:
        p = self._up(dz, ds, z, s)
        self._check(self._L.clh_cones_step_length(self._h, *p, float(alpha_max), self._sc.ptr, None))
        return float(self._scalars()[0])

    def step_length_noalloc(self, dz, ds, z, s, alpha_max=1.0) -> float:
        """step_length through the entry point whose reduction scratch is the
        library's static device array (clh_cones_step_length_noalloc)"""
        p = self._up(dz, ds, z, s)
        self._check(self._L.clh_cones_step_length_noalloc(self._h, *p, float(alpha_max), self._sc.ptr, None))
        return float(self._scalars()[0])

    def compute_barrier(self, z, s, dz, ds, alpha) -> float:
        p = self._up(z, s, dz, ds)
        self._check(self._L.clh_cones_compute_barrier(self._h, *p, float(alpha), self._sc.ptr, None))
        return float(self._scalars()[0])

    def margins(self, z):
        (pz,) = self._up(z)
        self._check(self._L.clh_cones_margins(self._h, pz, self._sc.ptr, None))
        a, b = self._scalars(2)
        return float(a), float(b)

    def scaled_unit_shift(self, z, alpha, primal=True) -> np.ndarray:
        (pz,) = self._set_out(z)
        self._check(self._L.clh_cones_scaled_unit_shift(self._h, pz, float(alpha), 1 if primal else 0, None))
        return self._o[0].download(self.m)

    def unit_initialization(self):
        """(z, s)"""
        self._check(self._L.clh_cones_unit_initialization(self._h, self._o[0].ptr, self._o[1].ptr, None))
        return self._o[0].download(self.m), self._o[1].download(self.m)

    def rectify_equilibration(self, e):
        """(delta, any_changed)"""
        (pe,) = self._up(e)
        changed = C.c_int32()
        self._check(self._L.clh_cones_rectify_equilibration(self._h, self._o[0].ptr, pe, C.byref(changed), None))
        return self._o[0].download(self.m), bool(changed.value)

    def close(self):
        for b in self._bufs:
            b.free()
        self._bufs = []
        if self._h:
            self._L.clh_cones_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
