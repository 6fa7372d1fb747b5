"""cuclarabel_amd -- the interior-point hot path of CuClarabel on MI355X.

Python mirror of Clarabel.jl's API for direct_solve_method = :hip / :hipmixed
(src/solver.jl:5-70, src/settings.jl:70-148, src/cones/cone_api.jl:24-56):

    solver = Solver(P, q, A, b, cones, Settings(direct_solve_method="hip"))
    sol = solver.solve()          # solve!
    solver.update_b(b2); solver.solve()

Everything runs through libclarabel_hip.so (include/clarabel_hip.h); the
package has no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, fields
from typing import List, Optional, Sequence

import numpy as np
import scipy.sparse as sp

from . import _lib

__all__ = [
    "ZeroConeT", "NonnegativeConeT", "SecondOrderConeT", "ExponentialConeT", "PowerConeT", "PSDTriangleConeT",
    "Settings", "Schedule", "Solver", "Server", "DefaultSolution", "STATUS",
]

STATUS = [
    "UNSOLVED", "SOLVED", "PRIMAL_INFEASIBLE", "DUAL_INFEASIBLE", "ALMOST_SOLVED",
    "ALMOST_PRIMAL_INFEASIBLE", "ALMOST_DUAL_INFEASIBLE", "MAX_ITERATIONS", "MAX_TIME",
    "NUMERICAL_ERROR", "INSUFFICIENT_PROGRESS",
]


# ---- cones (src/cones/cone_api.jl:24-56) ----------------------------------
class _Cone:
    code = -1
    param = 0.0

    def __repr__(self):
        return f"{type(self).__name__}({self.dim})"


class ZeroConeT(_Cone):
    code = 0

    def __init__(self, dim: int):
        self.dim = int(dim)


class NonnegativeConeT(_Cone):
    code = 1

    def __init__(self, dim: int):
        self.dim = int(dim)


class SecondOrderConeT(_Cone):
    code = 2

    def __init__(self, dim: int):
        self.dim = int(dim)


class ExponentialConeT(_Cone):
    code = 3

    def __init__(self):
        self.dim = 3


class PowerConeT(_Cone):
    code = 4

    def __init__(self, alpha: float):
        self.dim = 3
        self.param = float(alpha)
        self.alpha = self.param


class PSDTriangleConeT(_Cone):
    code = 5

    def __init__(self, dim: int):
        self.dim = int(dim)  # side of the matrix; numel = dim (dim + 1) / 2


def _numel(c: _Cone) -> int:
    return c.dim * (c.dim + 1) // 2 if c.code == 5 else c.dim


# ---- schedule (no counterpart in the reference) -----------------------------
@dataclass
class Schedule:
    """The engine's launch and tile choices (include/clarabel_hip.h,
    clh_schedule); 0 keeps the built-in rule.  Server.autotune and
    Solver.autotune choose them on the device; use_tuned takes the cached
    choices of this device for the problem's size class."""
    relax_max_cols: int = 0     # relaxed supernodes: most columns per panel (128; -1: none)
    relax_zero_pct: int = 0     # ... most explicit zeros, percent (25)
    group_lanes: int = 0        # lanes per supernode group (16, 32, 64; 0: by panel heights)
    blocks_per_cu: int = 0      # serving: resident workgroups per CU (0: as many as fit)
    grid_blocks: int = 0        # grid engine: workgroups of the launch (0: by factorization cost)
    tile_min_cols: int = 0      # grid engine: supernodes this wide take tiled updates (128)
    dense_min_cols: int = 0     # grid engine: panels this wide are factored over the grid (128)
    use_tuned: bool = False

    def to_c(self) -> _lib.ClhSchedule:
        c = _lib.ClhSchedule()
        for f in fields(self):
            setattr(c, f.name, int(getattr(self, f.name)))
        return c

    @staticmethod
    def from_c(c: _lib.ClhSchedule) -> "Schedule":
        return Schedule(**{f.name: (bool(getattr(c, f.name)) if f.name == "use_tuned" else int(getattr(c, f.name)))
                           for f in fields(Schedule)})


def _report(r: _lib.ClhTuneReport) -> dict:
    return {"default_rate": r.default_rate, "tuned_rate": r.tuned_rate, "trials": r.trials, "classes": r.classes,
            "seconds": r.seconds, "gain": (r.tuned_rate / r.default_rate) if r.default_rate > 0 else None}


# ---- settings (src/settings.jl:70-148) -------------------------------------
@dataclass
class Settings:
    max_iter: int = 200
    time_limit: float = float("inf")
    verbose: bool = False
    max_step_fraction: float = 0.99
    tol_gap_abs: float = 1e-8
    tol_gap_rel: float = 1e-8
    tol_feas: float = 1e-8
    tol_infeas_abs: float = 1e-8
    tol_infeas_rel: float = 1e-8
    tol_ktratio: float = 1e-6
    reduced_tol_gap_abs: float = 5e-5
    reduced_tol_gap_rel: float = 5e-5
    reduced_tol_feas: float = 1e-4
    reduced_tol_infeas_abs: float = 5e-12
    reduced_tol_infeas_rel: float = 5e-5
    reduced_tol_ktratio: float = 1e-4
    equilibrate_enable: bool = True
    equilibrate_max_iter: int = 10
    equilibrate_min_scaling: float = 1e-4
    equilibrate_max_scaling: float = 1e4
    linesearch_backtrack_step: float = 0.8
    min_switch_step_length: float = 1e-2
    min_terminate_step_length: float = 1e-4
    static_regularization_enable: bool = True
    static_regularization_constant: Optional[float] = None  # None: 1e-8 (:hip) or sqrt(eps32) (:hipmixed)
    static_regularization_proportional: float = float(np.finfo(float).eps) ** 2
    dynamic_regularization_enable: bool = True
    dynamic_regularization_eps: float = 1e-13
    dynamic_regularization_delta: float = 2e-7
    # a wrongly signed pivot above delta keeps its magnitude (False: the
    # reference's rule, sign * delta, always)
    dynamic_regularization_keep_magnitude: bool = True
    iterative_refinement_enable: bool = True
    iterative_refinement_reltol: float = 1e-12
    iterative_refinement_abstol: float = 1e-12
    iterative_refinement_max_iter: int = 10
    iterative_refinement_stop_ratio: float = 5.0
    direct_solve_method: str = "hip"  # "hip" (FP64) or "hipmixed" (FP32 factor + FP64 refinement)
    soc_threshold: int = 5
    # :hipmixed: FP32 refinement stalled above this (relative to |rhs|) moves the solve to FP64 factors
    mixed_switch_reltol: float = 1e-10
    mixed_fp64_static_reg: float = 1e-8
    # serving: "auto", "workgroup" (one request per 256-thread workgroup) or
    # "lanes" (up to 64 same-shape requests per 64-lane workgroup); a single
    # Solver: "auto" (the grid from 1e8 factorization flops), "workgroup" or
    # "grid" (one problem over a cooperative grid filling the GPU)
    serve_engine: str = "auto"
    schedule: Schedule = field(default_factory=Schedule)

    def to_c(self) -> _lib.ClhSettings:
        s = _lib.ClhSettings()
        for f in fields(self):
            v = getattr(self, f.name)
            if f.name == "direct_solve_method":
                m = str(v).lstrip(":")
                if m not in ("hip", "hipmixed"):
                    raise ValueError("direct_solve_method must be 'hip' or 'hipmixed'")
                s.direct_solve_method = 0 if m == "hip" else 1
            elif f.name == "serve_engine":
                engines = {"auto": 0, "workgroup": 1, "lanes": 2, "grid": 3, "resident": 4}
                if v not in engines:
                    raise ValueError("serve_engine must be 'auto', 'workgroup', 'lanes', 'grid' or 'resident'")
                s.serve_engine = engines[v]
            elif f.name == "schedule":
                s.schedule = (v if isinstance(v, Schedule) else Schedule(**dict(v))).to_c()
            elif f.name == "static_regularization_constant":
                if v is None:
                    v = 1e-8 if str(self.direct_solve_method).lstrip(":") == "hip" else float(np.sqrt(np.finfo(np.float32).eps))
                s.static_regularization_constant = float(v)
            else:
                setattr(s, f.name, type(getattr(s, f.name))(v) if not isinstance(v, bool) else int(v))
        return s


@dataclass
class DefaultSolution:
    """src/types.jl:297-332"""
    x: np.ndarray
    z: np.ndarray
    s: np.ndarray
    status: str
    obj_val: float
    obj_val_dual: float
    setup_time: float
    solve_time: float
    iterations: int
    r_prim: float
    r_dual: float
    info: dict = field(default_factory=dict)


def _csc(M, shape):
    M = sp.csc_matrix(M, shape=shape)
    M.sort_indices()
    M.sum_duplicates()
    return (np.ascontiguousarray(M.indptr, np.int32), np.ascontiguousarray(M.indices, np.int32),
            np.ascontiguousarray(M.data, np.float64))


def _ptr(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class _Problem:
    def __init__(self, P, q, A, b, cones: Sequence[_Cone]):
        self.q = np.ascontiguousarray(q, np.float64).ravel()
        self.b = np.ascontiguousarray(b, np.float64).ravel()
        self.n, self.m = self.q.size, self.b.size
        self.Pp, self.Pi, self.Px = _csc(P, (self.n, self.n))
        self.Ap, self.Ai, self.Ax = _csc(A, (self.m, self.n))
        self.cones = list(cones)
        if sum(_numel(c) for c in self.cones) != self.m:
            raise ValueError("cone dimensions do not match the number of rows of A")
        self.ct = np.array([c.code for c in self.cones], np.int32)
        self.cd = np.array([c.dim for c in self.cones], np.int32)
        self.cp = np.array([float(c.param) for c in self.cones], np.float64)

    def args(self):
        i32 = C.c_int32
        return (self.n, self.m, _ptr(self.Pp, i32), _ptr(self.Pi, i32), _ptr(self.Px), _ptr(self.q),
                _ptr(self.Ap, i32), _ptr(self.Ai, i32), _ptr(self.Ax), _ptr(self.b), len(self.ct),
                _ptr(self.ct, i32), _ptr(self.cd, i32), _ptr(self.cp))


def prepare(P, q, A, b, cones: Sequence[_Cone]) -> "_Problem":
    """A problem converted once to the C ABI's arrays (0-based int32 CSC,
    contiguous FP64), for callers that submit the same data repeatedly or want
    the conversion outside a timed serving loop.  Server.submit accepts it in
    place of (P, q, A, b, cones)."""
    return _Problem(P, q, A, b, cones)


def _info_dict(info: _lib.ClhInfo) -> dict:
    d = {f[0]: getattr(info, f[0]) for f in info._fields_}
    if "phase_seconds" in d:
        d["phase_seconds"] = list(d["phase_seconds"])
    return d


class Solver:
    """Clarabel.Solver(P,q,A,b,cones,settings) for the :hip engine."""

    def __init__(self, P, q, A, b, cones: Sequence[_Cone], settings: Optional[Settings] = None, _library=None):
        self._L = _library if _library is not None else _lib.lib()
        self.settings = settings or Settings()
        self._p = _Problem(P, q, A, b, cones)
        self._h = C.c_void_p()
        cs = self.settings.to_c()
        _lib.check(self._L, self._L.clh_solver_create(C.byref(self._h), *self._p.args(), C.byref(cs)))
        self.solution: Optional[DefaultSolution] = None

    def solve(self) -> DefaultSolution:
        """solve!(solver) (src/solver.jl:231-422)"""
        L = self._L
        _lib.check(L, L.clh_solve(self._h, None))
        n, m = self._p.n, self._p.m
        x, s, z = np.zeros(n), np.zeros(m), np.zeros(m)
        info = _lib.ClhInfo()
        _lib.check(L, L.clh_get_solution(self._h, _ptr(x), _ptr(s), _ptr(z), C.byref(info)))
        self.solution = DefaultSolution(x=x, z=z, s=s, status=STATUS[info.status], obj_val=info.obj_val,
                                        obj_val_dual=info.obj_val_dual, setup_time=info.setup_time,
                                        solve_time=info.solve_time, iterations=info.iterations, r_prim=info.r_prim,
                                        r_dual=info.r_dual, info=_info_dict(info))
        return self.solution

    def trace(self) -> np.ndarray:
        """per-iteration rows: pcost dcost gap pres dres k/t mu step (src/info_print.jl:122-137)"""
        rows = C.c_int32()
        buf = np.zeros(8 * (self.settings.max_iter + 2))
        _lib.check(self._L, self._L.clh_get_trace(self._h, _ptr(buf), self.settings.max_iter + 2, C.byref(rows)))
        return buf[: 8 * rows.value].reshape(rows.value, 8)

    # data updates (src/data_updating.jl:26-202)
    def update_P(self, P):
        if sp.issparse(P) or (hasattr(P, "ndim") and np.ndim(P) == 2):
            Pp, Pi, Px = _csc(P, (self._p.n, self._p.n))
            if not (np.array_equal(Pp, self._p.Pp) and np.array_equal(Pi, self._p.Pi)):
                raise ValueError("update_P: sparsity pattern differs from P")
        else:
            Px = np.ascontiguousarray(P, np.float64)
            if Px.size != self._p.Px.size:
                raise ValueError("update_P: wrong number of values")
        self._p.Px = np.ascontiguousarray(Px, np.float64)
        _lib.check(self._L, self._L.clh_update_P(self._h, _ptr(self._p.Px)))

    def update_q(self, q):
        self._p.q = np.ascontiguousarray(q, np.float64).ravel()
        _lib.check(self._L, self._L.clh_update_q(self._h, _ptr(self._p.q)))

    def update_A(self, A):
        if sp.issparse(A) or (hasattr(A, "ndim") and np.ndim(A) == 2):
            Ap, Ai, Ax = _csc(A, (self._p.m, self._p.n))
            if not (np.array_equal(Ap, self._p.Ap) and np.array_equal(Ai, self._p.Ai)):
                raise ValueError("update_A: sparsity pattern differs from A")
        else:
            Ax = np.ascontiguousarray(A, np.float64)
        self._p.Ax = np.ascontiguousarray(Ax, np.float64)
        _lib.check(self._L, self._L.clh_update_A(self._h, _ptr(self._p.Ax)))

    def update_b(self, b):
        self._p.b = np.ascontiguousarray(b, np.float64).ravel()
        _lib.check(self._L, self._L.clh_update_b(self._h, _ptr(self._p.b)))

    def autotune(self, max_seconds: float = 120.0) -> dict:
        """Choose the grid engine's schedule on this problem (before the first
        solve): the default's and the chosen schedule's IPM iterations/s."""
        r = _lib.ClhTuneReport()
        _lib.check(self._L, self._L.clh_solver_autotune(self._h, float(max_seconds), C.byref(r)))
        return _report(r)

    def schedule(self) -> Schedule:
        c = _lib.ClhSchedule()
        _lib.check(self._L, self._L.clh_solver_get_schedule(self._h, C.byref(c)))
        return Schedule.from_c(c)

    def close(self):
        """Release the solver's device memory now."""
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.clh_solver_destroy(h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()


class Server:
    """Continuous-batching solve server (no counterpart in the reference).

    Requests are admitted as slots and the HBM budget free up; all active
    requests advance together in graph-replayed iteration launches.  Small
    requests run on the lane engine (up to 64 requests of one shape per 64-lane
    workgroup, one per lane), larger ones on the workgroup engine
    (Settings.serve_engine).  A request's result does not depend on the
    requests served beside it (bitwise)."""

    def __init__(self, hbm_budget_bytes: int = 8 << 30, max_slots: int = 1024, settings: Optional[Settings] = None,
                 _library=None):
        self._L = _library if _library is not None else _lib.lib()
        self.settings = settings or Settings()
        self._h = C.c_void_p()
        cs = self.settings.to_c()
        _lib.check(self._L, self._L.clh_server_create(C.byref(self._h), int(hbm_budget_bytes), int(max_slots),
                                                      C.byref(cs), None))
        self._dims = {}

    def autotune(self, max_seconds: float = 120.0) -> dict:
        """Choose the schedule of each size class on the queued requests (they
        stay queued): the default's and the chosen schedules' solves/s on them."""
        r = _lib.ClhTuneReport()
        _lib.check(self._L, self._L.clh_server_autotune(self._h, float(max_seconds), C.byref(r)))
        return _report(r)

    def schedule(self, n_plus_m: int) -> Schedule:
        """The schedule a problem with n + m rows and columns is served with."""
        c = _lib.ClhSchedule()
        _lib.check(self._L, self._L.clh_server_get_schedule(self._h, int(n_plus_m), C.byref(c)))
        return Schedule.from_c(c)

    def submit(self, P, q=None, A=None, b=None, cones=None) -> int:
        p = P if isinstance(P, _Problem) else _Problem(P, q, A, b, cones)
        t = C.c_int64()
        _lib.check(self._L, self._L.clh_server_submit(self._h, C.byref(t), *p.args()))
        self._dims[t.value] = (p.n, p.m)
        return t.value

    def progress(self, passes: int = 4, wait: bool = True) -> int:
        """One round (admit, launch `passes` passes, harvest); with wait=False
        the launch runs on while the caller continues and the next call
        finishes it.  Returns the requests still queued or resident."""
        fn = self._L.clh_server_progress if wait else self._L.clh_server_progress_async
        rc = fn(self._h, int(passes))
        if rc < 0:
            raise RuntimeError(self._L.clh_last_error().decode())
        return rc

    def poll(self, ticket: int) -> bool:
        return self._L.clh_server_poll(self._h, int(ticket)) == 1

    def wait(self, ticket: int, passes: int = 4, timeout: float = float("inf")) -> bool:
        """Drive the device until the request has finished (True) or
        `timeout` seconds have passed (False)."""
        rc = self._L.clh_server_wait(self._h, int(ticket), int(passes), float(timeout))
        if rc < 0:
            raise ValueError(self._L.clh_last_error().decode())
        return rc == 1

    def cancel(self, ticket: int) -> bool:
        """End a request early (status MAX_TIME, or what its last iterate
        earns under the reduced tolerances).  False if it had finished."""
        rc = self._L.clh_server_cancel(self._h, int(ticket))
        if rc < 0:
            raise ValueError(self._L.clh_last_error().decode())
        return rc == 0

    def result(self, ticket: int):
        n, m = self._dims.pop(ticket)
        x, s, z = np.zeros(n), np.zeros(m), np.zeros(m)
        info = _lib.ClhInfo()
        lat = C.c_double()
        _lib.check(self._L, self._L.clh_server_result(self._h, int(ticket), _ptr(x), _ptr(s), _ptr(z),
                                                      C.byref(info), C.byref(lat)))
        d = _info_dict(info)
        return DefaultSolution(x=x, z=z, s=s, status=STATUS[info.status], obj_val=info.obj_val,
                               obj_val_dual=info.obj_val_dual, setup_time=0.0, solve_time=info.solve_time,
                               iterations=info.iterations, r_prim=info.r_prim, r_dual=info.r_dual,
                               info=dict(d, latency=lat.value))

    def run(self, passes: int = 4):
        while self.progress(passes) > 0:
            pass

    def stats(self) -> dict:
        st = _lib.ClhServerStats()
        _lib.check(self._L, self._L.clh_server_stats_get(self._h, C.byref(st)))
        return {f[0]: getattr(st, f[0]) for f in st._fields_}

    def close(self):
        """Release the server's device memory and stream now."""
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.clh_server_destroy(h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()
