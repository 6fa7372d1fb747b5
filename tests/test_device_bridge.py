"""Device arrays in and out of the solver (the counterpart of the reference's
Python bridge, src/python/py2jl.jl and jl2py.py): a problem handed over as
device arrays solves to the same result, bitwise, as the same problem from
host arrays; the solution copied device to device equals the host copy; the
data updates from device memory equal the host ones."""
import numpy as np
import pytest
import scipy.sparse as sp

import cuclarabel_amd as ca
from cuclarabel_amd import problems
from cuclarabel_amd.device import DeviceArray, DeviceCSC, DeviceSolver, DeviceVector, device_pointer

from helpers import assert_parity


class _Fake:
    """a host stand-in exposing the device-array interface (no GPU access)"""

    def __init__(self, typestr, shape, strides=None):
        self.__cuda_array_interface__ = {"typestr": typestr, "shape": shape, "strides": strides,
                                         "data": (4096, False), "version": 3}


def test_device_pointer_checks():
    assert device_pointer(_Fake("<f8", (5,)), np.float64, 5) == 4096
    assert device_pointer(DeviceVector(8192, 3), np.float64, 3) == 8192
    with pytest.raises(ValueError, match="type"):
        device_pointer(_Fake("<f4", (5,)), np.float64, 5)
    with pytest.raises(ValueError, match="elements"):
        device_pointer(_Fake("<i4", (4,)), np.int32, 5)
    with pytest.raises(ValueError, match="contiguous"):
        device_pointer(_Fake("<f8", (5,), (16,)), np.float64, 5)
    with pytest.raises(TypeError):
        device_pointer(np.zeros(5), np.float64, 5)


def _to_device(P, q, A, b):
    def csc(M):
        M = sp.csc_matrix(M)
        M.sort_indices()
        return DeviceCSC(DeviceArray.from_host(M.indptr, np.int32), DeviceArray.from_host(M.indices, np.int32),
                         DeviceArray.from_host(M.data, np.float64), M.shape)
    return csc(P), DeviceArray.from_host(q, np.float64), csc(A), DeviceArray.from_host(b, np.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["hip", "hipmixed"])
def test_gpu_device_arrays_match_host_arrays(oracle, method):
    P, q, A, b, cones = problems.portfolio(3)
    st = ca.Settings(direct_solve_method=method)
    host = ca.Solver(P, q, A, b, cones, st).solve()
    Pd, qd, Ad, bd = _to_device(P, q, A, b)
    print("device solver", method, flush=True)
    ds = DeviceSolver(Pd, qd, Ad, bd, cones, st)
    got = ds.solve()
    assert got.status == host.status and got.iterations == host.iterations
    assert np.array_equal(got.x, host.x) and np.array_equal(got.z, host.z) and np.array_equal(got.s, host.s)
    assert_parity(got, oracle.solve(P, q, A, b, cones), name="portfolio")
    x, s, z = DeviceArray(q.size), DeviceArray(b.size), DeviceArray(b.size)
    ds.solution_to(x, s, z)
    assert np.array_equal(x.to_host(), host.x) and np.array_equal(s.to_host(), host.s)
    assert np.array_equal(z.to_host(), host.z)
    # update_q! / update_b! from device memory, against the host path
    q2, b2 = q * 1.1, b.copy()
    b2[-1] *= 0.9
    hs = ca.Solver(P, q, A, b, cones, st)
    hs.solve()
    hs.update_q(q2)
    hs.update_b(b2)
    href = hs.solve()
    ds.update_q(DeviceArray.from_host(q2))
    ds.update_b(DeviceArray.from_host(b2))
    got2 = ds.solve()
    assert got2.iterations == href.iterations and np.array_equal(got2.x, href.x)
    ds.close()
    hs.close()


@pytest.mark.gpu
def test_gpu_device_arrays_rejects_wrong_type():
    P, q, A, b, cones = problems.portfolio(3)
    Pd, qd, Ad, bd = _to_device(P, q, A, b)
    with pytest.raises(ValueError):
        DeviceSolver(Pd, DeviceArray.from_host(q, np.int32), Ad, bd, cones)
