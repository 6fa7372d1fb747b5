"""The composite-cone operations of the C ABI (clh_cones_*: the reference's
src/gpucones/coneops_compositecone_gpu.jl operations) against what the
mathematics fixes and against the oracle's per-cone functions:

  * unit_initialization: the literal initial points of each family;
  * update_scaling: the scaling maps z to s (Hs z = s: NT scaling of the
    symmetric cones, the secant condition of the asymmetric primal-dual
    scaling); the second-order and PSD Hs blocks equal the oracle's
    (oracle.soc_update_scaling / psd_update_scaling);
  * get_Hs agrees with mul_Hs on every dense block;
  * affine_ds = lambda o lambda: its cone-wise trace is <s, z>;
  * ds_from_dz_offset(lambda o lambda, z) = W^T W z = s;
  * combined_ds_shift: shift + sigma_mu e is the Jordan product of the
    returned steps, and <W dz, W^-T ds> = <dz, ds> cone by cone;
  * step_length: the exact boundary (bisection) for the symmetric cones,
    a feasible backtracked step for the asymmetric ones;
  * compute_barrier, margins, scaled_unit_shift, rectify_equilibration:
    closed forms.
The emulated build runs the engine source serially; the gpu-marked tests run
the HIP grid kernels, on the small composite and on a large one."""
import numpy as np
import pytest

import cuclarabel_amd as ca
from cuclarabel_amd.cones import CompositeCone

from helpers import build_emu

S2 = np.sqrt(2.0)


def tri(n):
    return n * (n + 1) // 2


def svec(M):
    n = M.shape[0]
    return np.array([M[r, c] if r == c else S2 * M[r, c] for c in range(n) for r in range(c + 1)])


def smat(v):
    n = int(round((np.sqrt(8 * v.size + 1) - 1) / 2))
    M = np.zeros((n, n))
    k = 0
    for c in range(n):
        for r in range(c + 1):
            M[r, c] = M[c, r] = v[k] if r == c else v[k] / S2
            k += 1
    return M


def blocks(cones):
    """(kind, offset, dim, param) per cone"""
    out, o = [], 0
    for c in cones:
        d = c.dim * (c.dim + 1) // 2 if isinstance(c, ca.PSDTriangleConeT) else c.dim
        out.append((type(c).__name__, o, d, getattr(c, "param", 0.0)))
        o += d
    return out


def small_cones():
    return [ca.ZeroConeT(2), ca.NonnegativeConeT(4), ca.SecondOrderConeT(3), ca.SecondOrderConeT(7),
            ca.ExponentialConeT(), ca.PowerConeT(0.3), ca.PSDTriangleConeT(3), ca.NonnegativeConeT(1),
            ca.SecondOrderConeT(5), ca.PSDTriangleConeT(2)]


def symmetric_cones():
    return [ca.ZeroConeT(1), ca.NonnegativeConeT(5), ca.SecondOrderConeT(4), ca.SecondOrderConeT(9),
            ca.PSDTriangleConeT(3), ca.NonnegativeConeT(2), ca.PSDTriangleConeT(4)]


def interior(cones, rng, dual=False):
    """a point strictly inside each cone (the dual cone for dual=True)"""
    v = []
    for kind, o, d, a in blocks(cones):
        if kind == "ZeroConeT":
            v.append(np.zeros(d))
        elif kind == "NonnegativeConeT":
            v.append(rng.uniform(0.2, 2.0, d))
        elif kind == "SecondOrderConeT":
            x = rng.standard_normal(d - 1)
            v.append(np.r_[np.linalg.norm(x) + rng.uniform(0.2, 1.0), x])
        elif kind == "ExponentialConeT":
            base = np.array([-1.051383945322714, 0.556409619469370, 1.258967884768947])
            v.append(base + 0.05 * rng.standard_normal(3))
        elif kind == "PowerConeT":
            v.append(np.array([np.sqrt(1 + a), np.sqrt(2 - a), 0.0]) + 0.05 * rng.standard_normal(3))
        else:
            n = int(round((np.sqrt(8 * d + 1) - 1) / 2))
            W = rng.standard_normal((n, n))
            v.append(svec(np.eye(n) + 0.2 * W @ W.T))
    return np.concatenate(v)


def in_soc(x, tol=0.0):
    return x[0] - np.linalg.norm(x[1:]) >= -tol


def jordan(kind, x, y):
    if kind == "NonnegativeConeT":
        return x * y
    if kind == "SecondOrderConeT":
        return np.r_[x @ y, x[0] * y[1:] + y[0] * x[1:]]
    X, Y = smat(x), smat(y)
    return svec((X @ Y + Y @ X) / 2)


def unit(kind, d):
    if kind == "NonnegativeConeT":
        return np.ones(d)
    if kind == "SecondOrderConeT":
        return np.r_[1.0, np.zeros(d - 1)]
    n = int(round((np.sqrt(8 * d + 1) - 1) / 2))
    return svec(np.eye(n))


def check_small(lib, oracle):
    rng = np.random.RandomState(3)
    cones = small_cones()
    K = CompositeCone(cones, _library=lib)
    try:
        bl = blocks(cones)
        m = K.m
        assert m == sum(b[2] for b in bl)
        assert K.info["n_soc"] == 3 and K.info["n_exp"] == 1 and K.info["n_pow"] == 1 and K.info["n_psd"] == 2
        assert K.info["symmetric"] == 0

        # unit_initialization (coneops_*_gpu.jl unit_initialization!)
        z0, s0 = K.unit_initialization()
        for kind, o, d, a in bl:
            if kind == "ZeroConeT":
                exp_ = np.zeros(d)
            elif kind == "ExponentialConeT":
                exp_ = np.array([-1.051383945322714, 0.556409619469370, 1.258967884768947])
            elif kind == "PowerConeT":
                exp_ = np.array([np.sqrt(1 + a), np.sqrt(1 + (1 - a)), 0.0])
            else:
                exp_ = unit(kind, d)
            np.testing.assert_allclose(z0[o:o + d], exp_, rtol=0, atol=1e-15)
            np.testing.assert_allclose(s0[o:o + d], exp_, rtol=0, atol=1e-15)

        s = interior(cones, rng)
        z = interior(cones, rng, dual=True)
        nu = K.info["degree"]
        mu = float(s @ z) / nu
        assert K.update_scaling(s, z, mu)

        # Hs z = s (NT scaling / the secant condition of the primal-dual scaling)
        Hz = K.mul_Hs(z)
        np.testing.assert_allclose(Hz, s, rtol=1e-9, atol=1e-11)

        # get_Hs blocks against mul_Hs; second-order and PSD blocks against the oracle
        Hs = K.get_Hs()
        assert Hs.size == K.info["hs_len"]
        x = rng.standard_normal(m)
        y = K.mul_Hs(x)
        h = 0
        for kind, o, d, a in bl:
            if kind in ("ZeroConeT", "NonnegativeConeT"):
                np.testing.assert_allclose(y[o:o + d], Hs[h:h + d] * x[o:o + d], rtol=1e-13, atol=1e-14)
                h += d
                continue
            if kind == "SecondOrderConeT" and d > 5:  # diagonal part only: the expansion carries the rest
                ref = oracle.soc_update_scaling(s[o:o + d], z[o:o + d])
                eta2 = ref["eta"] ** 2
                np.testing.assert_allclose(Hs[h:h + d], eta2 * np.r_[ref["d"], np.ones(d - 1)], rtol=1e-10)
                w = ref["w"]
                full = eta2 * (2 * np.outer(w, w) - np.diag(np.r_[1.0, -np.ones(d - 1)]))
                np.testing.assert_allclose(y[o:o + d], full @ x[o:o + d], rtol=1e-10, atol=1e-12)
                h += d
                continue
            B = Hs[h:h + d * d].reshape(d, d, order="F")
            np.testing.assert_allclose(B, B.T, rtol=0, atol=1e-14)
            np.testing.assert_allclose(y[o:o + d], B @ x[o:o + d], rtol=1e-12, atol=1e-13)
            if kind == "SecondOrderConeT":
                ref = oracle.soc_update_scaling(s[o:o + d], z[o:o + d])
                w = ref["w"]
                full = ref["eta"] ** 2 * (2 * np.outer(w, w) - np.diag(np.r_[1.0, -np.ones(d - 1)]))
                np.testing.assert_allclose(B, full, rtol=1e-10, atol=1e-12)
            if kind == "PSDTriangleConeT":
                n = int(round((np.sqrt(8 * d + 1) - 1) / 2))
                ref = oracle.psd_update_scaling(n, s[o:o + d], z[o:o + d])
                np.testing.assert_allclose(B, ref["Hs"], rtol=1e-9, atol=1e-11)
            if kind in ("ExponentialConeT", "PowerConeT"):
                # the primal-dual scaling (oracle pinned by the secant conditions)
                code = 3 if kind == "ExponentialConeT" else 4
                ref = oracle.nonsym_update_Hs(code, s[o:o + d], z[o:o + d], mu, 0, a)
                np.testing.assert_allclose(B, ref, rtol=1e-11, atol=1e-13)
            h += d * d
        assert h == Hs.size

        # affine_ds = lambda o lambda: cone-wise <e, lambda o lambda> = <s, z>
        ds = K.affine_ds(s)
        for kind, o, d, a in bl:
            if kind == "ZeroConeT":
                assert np.all(ds[o:o + d] == 0)
            elif kind in ("ExponentialConeT", "PowerConeT"):
                np.testing.assert_array_equal(ds[o:o + d], s[o:o + d])
            else:
                np.testing.assert_allclose(unit(kind, d) @ ds[o:o + d], s[o:o + d] @ z[o:o + d], rtol=1e-11)
        # ds_from_dz_offset(lambda o lambda, z) = W'W z = s on the symmetric cones
        out = K.ds_from_dz_offset(ds, z)
        for kind, o, d, a in bl:
            if kind in ("NonnegativeConeT", "SecondOrderConeT", "PSDTriangleConeT"):
                np.testing.assert_allclose(out[o:o + d], s[o:o + d], rtol=1e-9, atol=1e-11)
            elif kind == "ZeroConeT":
                assert np.all(out[o:o + d] == 0)
            else:
                np.testing.assert_array_equal(out[o:o + d], ds[o:o + d])

        # combined_ds_shift: Jordan product of the scaled steps, inner products kept
        dz = rng.standard_normal(m)
        dsv = rng.standard_normal(m)
        sigmu = 0.37
        shift, zz, ss = K.combined_ds_shift(dz, dsv, sigmu)
        for kind, o, d, a in bl:
            if kind in ("NonnegativeConeT", "SecondOrderConeT", "PSDTriangleConeT"):
                np.testing.assert_allclose(shift[o:o + d] + sigmu * unit(kind, d),
                                           jordan(kind, ss[o:o + d], zz[o:o + d]), rtol=1e-11, atol=1e-12)
                np.testing.assert_allclose(zz[o:o + d] @ ss[o:o + d], dz[o:o + d] @ dsv[o:o + d], rtol=1e-9,
                                           atol=1e-11)
            elif kind in ("ExponentialConeT", "PowerConeT"):
                # sigma mu grad f*(z) - eta(step_s, step_z), the third-order
                # correction at the z of the last update_scaling (the oracle's
                # correction is pinned by finite differences of the dual Hessian)
                code = 3 if kind == "ExponentialConeT" else 4
                zc = z[o:o + 3]
                g = (oracle.exp_dual_grad_H(zc) if code == 3 else oracle.pow_dual_grad_H(zc, a))[0]
                eta = oracle.nonsym_higher_correction(code, zc, dsv[o:o + 3], dz[o:o + 3], a)
                np.testing.assert_allclose(shift[o:o + 3], sigmu * g - eta, rtol=1e-10, atol=1e-12)
                np.testing.assert_array_equal(zz[o:o + 3], dz[o:o + 3])
                np.testing.assert_array_equal(ss[o:o + 3], dsv[o:o + 3])

        # margins and scaled_unit_shift (closed forms)
        amin, bsum = K.margins(z)
        ms = []
        for kind, o, d, a in bl:
            if kind == "NonnegativeConeT":
                ms += list(z[o:o + d])
            elif kind == "SecondOrderConeT":
                ms.append(z[o] - np.linalg.norm(z[o + 1:o + d]))
            elif kind == "PSDTriangleConeT":
                ms += list(np.linalg.eigvalsh(smat(z[o:o + d])))
        ms = np.array(ms)
        assert amin == pytest.approx(ms.min(), rel=1e-10)
        assert bsum == pytest.approx(ms[ms > 0].sum(), rel=1e-10)
        zs = K.scaled_unit_shift(z, 0.25, primal=True)
        zd = K.scaled_unit_shift(z + 1.0, 0.25, primal=False)
        for kind, o, d, a in bl:
            if kind == "ZeroConeT":
                assert np.all(zs[o:o + d] == 0)
                np.testing.assert_array_equal(zd[o:o + d], z[o:o + d] + 1.0)
            elif kind in ("ExponentialConeT", "PowerConeT"):
                np.testing.assert_array_equal(zs[o:o + d], z[o:o + d])
            else:
                np.testing.assert_allclose(zs[o:o + d], z[o:o + d] + 0.25 * unit(kind, d), rtol=0, atol=1e-15)

        # step length: feasible for every cone, the asymmetric ones backtracked
        dz = 3 * rng.standard_normal(m)
        dsv = 3 * rng.standard_normal(m)
        al = K.step_length(dz, dsv, z, s, 1.0)
        assert 0 < al <= 1.0
        zn, sn = z + al * dz, s + al * dsv
        for kind, o, d, a in bl:
            if kind == "NonnegativeConeT":
                assert np.all(zn[o:o + d] >= -1e-12) and np.all(sn[o:o + d] >= -1e-12)
            elif kind == "SecondOrderConeT":
                assert in_soc(zn[o:o + d], 1e-9) and in_soc(sn[o:o + d], 1e-9)
            elif kind == "PSDTriangleConeT":
                assert np.linalg.eigvalsh(smat(zn[o:o + d])).min() > -1e-9
                assert np.linalg.eigvalsh(smat(sn[o:o + d])).min() > -1e-9

        # rectify_equilibration: delta = 1 on elementwise cones, mean(e)/e elsewhere
        e = rng.uniform(0.5, 2.0, m)
        delta, changed = K.rectify_equilibration(e)
        assert changed
        for kind, o, d, a in bl:
            if kind in ("ZeroConeT", "NonnegativeConeT"):
                assert np.all(delta[o:o + d] == 1.0)
            else:
                np.testing.assert_allclose(delta[o:o + d], e[o:o + d].mean() / e[o:o + d], rtol=1e-15)
    finally:
        K.close()


def exact_step(cones, x, dx, amax):
    """the largest a <= amax keeping x + a dx in the symmetric cones (bisection)"""
    def ok(a):
        y = x + a * dx
        for kind, o, d, _ in blocks(cones):
            if kind == "NonnegativeConeT" and np.any(y[o:o + d] < 0):
                return False
            if kind == "SecondOrderConeT" and not in_soc(y[o:o + d]):
                return False
            if kind == "PSDTriangleConeT" and np.linalg.eigvalsh(smat(y[o:o + d])).min() < 0:
                return False
        return True
    if ok(amax):
        return amax
    lo, hi = 0.0, amax
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if ok(mid) else (lo, mid)
    return lo


def check_symmetric(lib):
    rng = np.random.RandomState(5)
    cones = symmetric_cones()
    K = CompositeCone(cones, _library=lib)
    try:
        bl = blocks(cones)
        m = K.m
        assert K.info["symmetric"] == 1
        for trial in range(4):
            s = interior(cones, rng)
            z = interior(cones, rng, dual=True)
            dz = 2 * rng.standard_normal(m)
            dsv = 2 * rng.standard_normal(m)
            al = K.step_length(dz, dsv, z, s, 1.0)
            ref = min(exact_step(cones, z, dz, 1.0), exact_step(cones, s, dsv, 1.0))
            assert al == pytest.approx(ref, rel=1e-8, abs=1e-12), trial
        # barrier in closed form at a feasible step
        a = 0.5 * al
        bar = K.compute_barrier(z, s, dz, dsv, a)
        zn, sn = z + a * dz, s + a * dsv
        ref = 0.0
        for kind, o, d, _ in bl:
            if kind == "NonnegativeConeT":
                ref -= np.sum(np.log(sn[o:o + d] * zn[o:o + d]))
            elif kind == "SecondOrderConeT":
                r = lambda v: v[0] ** 2 - v[1:] @ v[1:]
                ref -= 0.5 * np.log(r(sn[o:o + d]) * r(zn[o:o + d]))
            elif kind == "PSDTriangleConeT":
                ref -= np.linalg.slogdet(smat(zn[o:o + d]))[1] + np.linalg.slogdet(smat(sn[o:o + d]))[1]
        assert bar == pytest.approx(ref, rel=1e-10)
        # identity scaling: Hs = I on the symmetric cones
        K.set_identity_scaling()
        x = rng.standard_normal(m)
        y = K.mul_Hs(x)
        for kind, o, d, _ in bl:
            if kind == "ZeroConeT":
                assert np.all(y[o:o + d] == 0)
            else:
                np.testing.assert_allclose(y[o:o + d], x[o:o + d], rtol=1e-15, atol=1e-15)
    finally:
        K.close()


def check_large(lib):
    """many cones per family, spread over many workgroups on the GPU"""
    rng = np.random.RandomState(9)
    cones = ([ca.NonnegativeConeT(20000)] + [ca.SecondOrderConeT(5) for _ in range(2000)]
             + [ca.ExponentialConeT() for _ in range(1500)] + [ca.PowerConeT(0.2 + 0.6 * rng.uniform())
                                                               for _ in range(500)]
             + [ca.PSDTriangleConeT(10) for _ in range(100)] + [ca.SecondOrderConeT(40) for _ in range(50)])
    K = CompositeCone(cones, _library=lib)
    try:
        s = interior(cones, rng)
        z = interior(cones, rng, dual=True)
        mu = float(s @ z) / K.info["degree"]
        assert K.update_scaling(s, z, mu)
        np.testing.assert_allclose(K.mul_Hs(z), s, rtol=1e-8, atol=1e-10)
        ds = K.affine_ds(s)
        out = K.ds_from_dz_offset(ds, z)
        bl = blocks(cones)
        sym = np.zeros(K.m, bool)
        for kind, o, d, a in bl:
            if kind in ("NonnegativeConeT", "SecondOrderConeT", "PSDTriangleConeT"):
                sym[o:o + d] = True
        np.testing.assert_allclose(out[sym], s[sym], rtol=1e-8, atol=1e-10)
        dz = rng.standard_normal(K.m)
        dsv = rng.standard_normal(K.m)
        al = K.step_length(dz, dsv, z, s, 1.0)
        assert 0 < al <= 1.0
        amin, bsum = K.margins(z)
        assert amin > 0 and bsum > 0
        return K.get_Hs()
    finally:
        K.close()


@pytest.fixture(scope="module")
def emu():
    return build_emu()


def test_emu_small(emu, oracle):
    check_small(emu, oracle)


def test_emu_symmetric(emu):
    check_symmetric(emu)


def test_emu_invalid(emu):
    with pytest.raises(ValueError):
        CompositeCone([ca.PowerConeT(1.5)], _library=emu)
    with pytest.raises(ValueError):
        CompositeCone([], _library=emu)


@pytest.mark.gpu
def test_gpu_small(oracle):
    print("cones small", flush=True)
    check_small(None, oracle)


@pytest.mark.gpu
def test_gpu_symmetric():
    print("cones symmetric", flush=True)
    check_symmetric(None)


@pytest.mark.gpu
def test_gpu_large_matches_serial(emu):
    """the grid kernels on 24 150 cones; the Hs blocks match the serial build
    to rounding (contracted multiply-adds differ between the two compilers;
    entries formed by cancellation differ relatively more, so the bound is
    on the largest entry)"""
    print("cones large", flush=True)
    hg = check_large(None)
    he = check_large(emu)
    assert np.abs(hg - he).max() <= 1e-11 * np.abs(he).max()
