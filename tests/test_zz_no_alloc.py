"""clh_cones_step_length_noalloc (the allocation-free variant of
clh_cones_step_length, its reduction scratch in the library's static device
array) against clh_cones_step_length: the same kernels and reduction order,
so the step lengths are equal bit for bit.  Two calls in a row, the second on
a smaller composite (the scratch keeps the first call's partials), and the
rejection of a composite larger than the scratch."""
import numpy as np
import pytest

import cuclarabel_amd as ca
from cuclarabel_amd.cones import CompositeCone
from test_cones_abi import interior

pytestmark = pytest.mark.gpu


def _large():
    rng = np.random.RandomState(9)
    return ([ca.NonnegativeConeT(20000)] + [ca.SecondOrderConeT(5) for _ in range(2000)]
            + [ca.ExponentialConeT() for _ in range(1500)]
            + [ca.PowerConeT(0.2 + 0.6 * rng.uniform()) for _ in range(500)]
            + [ca.PSDTriangleConeT(10) for _ in range(100)] + [ca.SecondOrderConeT(40) for _ in range(50)])


def _small():
    return [ca.ZeroConeT(2), ca.NonnegativeConeT(40), ca.SecondOrderConeT(3), ca.SecondOrderConeT(9),
            ca.PSDTriangleConeT(3)]


def _directions(K, cones, seed):
    rng = np.random.RandomState(seed)
    s = interior(cones, rng)
    z = interior(cones, rng, dual=True)
    assert K.update_scaling(s, z, float(s @ z) / K.info["degree"])
    return 3 * rng.standard_normal(K.m), 3 * rng.standard_normal(K.m), z, s


def test_noalloc_step_length_equals_step_length():
    # the large composite first (asymmetric: symmetric bound, then backtracking
    # from it), then the smaller symmetric one while the scratch still holds
    # the first call's partials and scalars
    for cones, seed in ((_large(), 1), (_small(), 2)):
        print("composite of", len(cones), "cones", flush=True)
        K = CompositeCone(cones)
        try:
            v = _directions(K, cones, seed)
            for amax in (1.0, 0.3):
                ref = K.step_length(*v, alpha_max=amax)
                got = K.step_length_noalloc(*v, alpha_max=amax)
                assert 0 < ref <= amax
                assert got == ref, (len(cones), amax, got, ref)
        finally:
            K.close()


def test_noalloc_rejects_a_composite_beyond_the_scratch():
    cones = [ca.NonnegativeConeT(600000)]  # 2344 workgroups, the scratch holds 2048
    K = CompositeCone(cones)
    try:
        v = _directions(K, cones, 3)
        assert 0 < K.step_length(*v) <= 1
        with pytest.raises(RuntimeError, match="static scratch"):
            K.step_length_noalloc(*v)
    finally:
        K.close()
