"""Pins the CPU oracle before anything is compared against it.

* end to end: every deterministic known answer of the reference test suite
  and the two logged runs of the reference GPU path (tests/golden);
* cone arithmetic: closed-form identities, finite differences of the barrier
  definitions and brute force, chosen so a dropped term, a sign or index
  mistake fails a check;
* the LDL: backward error against a dense solve, inertia.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "reference_cases.json")))


@pytest.fixture(scope="module")
def oracle(reference_oracle):
    """this module pins the oracle as the reference states it (the
    reference's dynamic-regularization rule)"""
    return reference_oracle


def _solve_case(oracle, c):
    n, m = len(c["q"]), len(c["b"])
    P = sp.csc_matrix(np.array(c["P"], float).reshape(n, n))
    A = sp.csc_matrix(np.array(c["A"], float).reshape(m, n))
    kw = {}
    if "update_b" in c:
        kw = dict(update_b=np.array(c["update_b"]), update_P=np.array(c["update_P"]),
                  update_P_without_cscale=c["update_P_without_cscale"])
    return oracle.solve(P, c["q"], A, c["b"], [tuple(x) for x in c["cones"]], **kw)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_reference_known_answers(oracle, case):
    r = _solve_case(oracle, case)
    assert r["status"] == case["status"]
    if "x" in case:
        assert np.linalg.norm(r["x"] - np.array(case["x"])) < case["tol"]
    if "obj" in case:
        assert abs(r["obj_val"] - case["obj"]) < case["tol"]
    if "INFEASIBLE" in case["status"]:
        assert np.isnan(r["obj_val"]) and np.isnan(r["obj_val_dual"])


@pytest.mark.parametrize("name", ["notebook_A", "notebook_B"])
def test_notebook_gpu_traces(oracle, name):
    """Per-iteration agreement with the logged :cudss runs at the printed
    precision (4 significant digits for costs, 3 for the rest).  The primal
    residual column is at round-off level (1e-16..1e-14) and only bounded."""
    c = next(x for x in CASES if x["name"] == name)
    r = _solve_case(oracle, c)
    assert r["status"] == "SOLVED" and r["iterations"] == c["iterations"]
    tr = r["trace"]
    ref = c["trace"]
    assert tr.shape[0] == len(ref)
    for k, row in enumerate(ref):
        got = tr[k]
        for col, digits in [(1, 4), (2, 4), (3, 2), (5, 2), (6, 2), (7, 2)]:
            want = row[col]
            assert float("%.*e" % (digits, got[col - 1])) == pytest.approx(want, rel=10 ** (-digits) * 1.01), (k, col)
        assert got[3] < 1e-13
        if row[8] is not None:
            assert float("%.2e" % got[7]) == pytest.approx(row[8], rel=1.01e-2)
    np.testing.assert_allclose(r["x"], c["x"], atol=1e-9)


# ------------------------------------------------------------------ cones
def test_wright_omega_defining_equation(oracle):
    for z in np.concatenate([np.linspace(0.01, 1 + np.pi, 40), np.linspace(1 + np.pi + 1e-3, 60, 40), [1e3, 1e6]]):
        w = oracle.wright_omega(z)
        assert abs(w + np.log(w) - z) <= 1e-12 * max(1, z)


def _exp_dual_barrier(z):
    return -np.log(-z[2] * z[0]) - np.log(z[1] - z[0] - z[0] * np.log(-z[2] / z[0]))


def _pow_dual_barrier(z, a):
    return (-np.log((z[0] / a) ** (2 * a) * (z[1] / (1 - a)) ** (2 - 2 * a) - z[2] ** 2)
            - (1 - a) * np.log(z[0]) - a * np.log(z[1]))


def _fd_grad(f, x, h=1e-6):
    g = np.zeros(3)
    for i in range(3):
        e = np.zeros(3); e[i] = h * max(1, abs(x[i]))
        g[i] = (f(x + e) - f(x - e)) / (2 * e[i])
    return g


def _fd_hess(gf, x, h=1e-6):
    H = np.zeros((3, 3))
    for i in range(3):
        e = np.zeros(3); e[i] = h * max(1, abs(x[i]))
        H[:, i] = (gf(x + e) - gf(x - e)) / (2 * e[i])
    return H


def test_exp_dual_gradient_hessian_fd(oracle):
    rng = np.random.RandomState(7)
    for _ in range(20):
        z0 = -rng.uniform(0.2, 2); z2 = rng.uniform(0.2, 2)
        z1 = z0 * np.log(-z2 / z0) - z0 + rng.uniform(0.1, 2)  # strictly dual-feasible
        z = np.array([z0, z1, z2])
        g, H = oracle.exp_dual_grad_H(z)
        np.testing.assert_allclose(g, _fd_grad(_exp_dual_barrier, z), rtol=1e-6, atol=1e-8)
        Hfd = _fd_hess(lambda v: oracle.exp_dual_grad_H(v)[0], z)
        np.testing.assert_allclose(H, Hfd, rtol=1e-5, atol=1e-7)


def test_exp_primal_gradient_conjugacy(oracle):
    """The primal barrier is the conjugate of the dual barrier f*, so its
    gradient g(s) = -z with grad f*(z) = -s (checked through the dual gradient,
    itself pinned by finite differences above)."""
    rng = np.random.RandomState(8)
    for _ in range(20):
        s1 = rng.uniform(0.2, 2); s2 = rng.uniform(0.2, 2)
        s0 = s1 * np.log(s2 / s1) - rng.uniform(0.1, 2)
        s = np.array([s0, s1, s2])
        g = oracle.exp_gradient_primal(s)
        gd, _ = oracle.exp_dual_grad_H(-g)
        np.testing.assert_allclose(gd, -s, rtol=1e-9, atol=1e-10)


def test_pow_gradients(oracle):
    """Dual gradient/Hessian: finite differences of the dual barrier.
    Primal gradient: logarithmic homogeneity <s, g> = -3 always; conjugacy
    grad f*(-g) = -s whenever the reference's one-sided Newton iterates.  The
    Newton of coneops_nonsymmetric_common.jl:221-244 stops on the first
    negative step (`dx < eps`), so when its initial guess (coneops_powcone.jl
    :455) lies past the root it returns that guess unchanged; the oracle keeps
    this behaviour and the test checks both branches (on interior points the
    guess branch is the one taken, so the primal gradient is the guess)."""
    rng = np.random.RandomState(9)
    n_iter_branch = n_guess_branch = 0
    for _ in range(60):
        a = rng.uniform(0.15, 0.85)
        s0, s1 = rng.uniform(0.3, 2, 2)
        s = np.array([s0, s1, rng.uniform(-0.9, 0.9) * s0 ** a * s1 ** (1 - a)])
        g = oracle.pow_gradient_primal(s, a)
        assert s @ g == pytest.approx(-3.0, rel=1e-12)
        s3 = abs(s[2])
        phi = s0 ** (2 * a) * s1 ** (2 - 2 * a)
        x0 = -1 / s3 + (2 * s3 + np.sqrt(phi * phi / s3 / s3 + 3 * phi)) / (phi - s3 * s3)
        t1, t2 = x0 * x0, 2 * x0 / s3
        f0 = (2 * a * np.log(2 * a * t1 + (1 + a) * t2) + 2 * (1 - a) * np.log(2 * (1 - a) * t1 + (2 - a) * t2)
              - np.log(phi) - np.log(t1 + t2) - 2 * np.log(t2) - 2 * a * np.log(a) - 2 * (1 - a) * np.log(1 - a))
        if f0 > 0:
            n_iter_branch += 1
            gd, _ = oracle.pow_dual_grad_H(-g, a)
            np.testing.assert_allclose(gd, -s, rtol=1e-7, atol=1e-8)
        else:
            n_guess_branch += 1
            assert abs(g[2]) == pytest.approx(x0, rel=1e-14)
        z0, z1 = rng.uniform(0.3, 2, 2)
        z = np.array([z0, z1, rng.uniform(-0.9, 0.9) * (z0 / a) ** a * (z1 / (1 - a)) ** (1 - a)])
        g, H = oracle.pow_dual_grad_H(z, a)
        np.testing.assert_allclose(g, _fd_grad(lambda v: _pow_dual_barrier(v, a), z), rtol=1e-6, atol=1e-8)
        Hfd = _fd_hess(lambda v: oracle.pow_dual_grad_H(v, a)[0], z)
        np.testing.assert_allclose(H, Hfd, rtol=1e-5, atol=1e-7)
    # on this sample the initial guess always lies past the root
    assert n_iter_branch + n_guess_branch == 60 and n_guess_branch > 0


def _exp_dual_point(rng):
    z0 = -rng.uniform(0.2, 2); z2 = rng.uniform(0.2, 2)
    return np.array([z0, z0 * np.log(-z2 / z0) - z0 + rng.uniform(0.1, 2), z2])


def _exp_primal_point(rng):
    s1 = rng.uniform(0.2, 2); s2 = rng.uniform(0.2, 2)
    return np.array([s1 * np.log(s2 / s1) - rng.uniform(0.1, 2), s1, s2])


def _pow_dual_point(rng, a):
    z0, z1 = rng.uniform(0.3, 2, 2)
    return np.array([z0, z1, rng.uniform(-0.9, 0.9) * (z0 / a) ** a * (z1 / (1 - a)) ** (1 - a)])


def _pow_primal_point(rng, a):
    s0, s1 = rng.uniform(0.3, 2, 2)
    return np.array([s0, s1, rng.uniform(-0.9, 0.9) * s0 ** a * s1 ** (1 - a)])


def _dual_grad_H(oracle, code, z, a):
    return oracle.exp_dual_grad_H(z) if code == 3 else oracle.pow_dual_grad_H(z, a)


@pytest.mark.parametrize("code", [3, 4], ids=["exp", "pow"])
def test_nonsym_third_order_correction_fd(oracle, code):
    """higher_correction! (coneops_expcone.jl:338-385, coneops_powcone.jl:
    329-404) returns eta = 1/2 D^3 f*(z)[u, v] with u = H(z)^-1 ds (the
    reference subtracts it from sigma mu grad in combined_ds_shift!).  The
    third derivative is taken here by central differences of the dual Hessian
    along u, that Hessian itself being pinned by finite differences of the
    dual barrier above; D^3 f* is symmetric in its arguments, which the
    swapped evaluation checks as well."""
    rng = np.random.RandomState(11 + code)
    for _ in range(30):
        a = rng.uniform(0.2, 0.8)
        z = _exp_dual_point(rng) if code == 3 else _pow_dual_point(rng, a)
        H = _dual_grad_H(oracle, code, z, a)[1]
        ds, v = rng.randn(3), rng.randn(3)
        u = np.linalg.solve(H, ds)
        h = 1e-5 / max(1.0, np.linalg.norm(u) / np.linalg.norm(z))
        dH = (_dual_grad_H(oracle, code, z + h * u, a)[1] - _dual_grad_H(oracle, code, z - h * u, a)[1]) / (2 * h)
        eta = oracle.nonsym_higher_correction(code, z, ds, v, a)
        np.testing.assert_allclose(eta, 0.5 * dH @ v, rtol=1e-5, atol=1e-7 * np.abs(dH @ v).max())
        # symmetry: D^3 f*[H^-1 ds, v] . w = D^3 f*[H^-1 ds, w] . v, i.e. with
        # ds' = H v and v' = H^-1 ds the correction is unchanged
        eta_sw = oracle.nonsym_higher_correction(code, z, H @ v, u, a)
        np.testing.assert_allclose(eta_sw, eta, rtol=1e-9, atol=1e-12 * np.abs(eta).max())


@pytest.mark.parametrize("code", [3, 4], ids=["exp", "pow"])
def test_nonsym_primal_dual_scaling_secant(oracle, code):
    """use_primal_dual_scaling (coneops_nonsymmetric_common.jl:82-164) is the
    rank-3 BFGS-type update of Dahl and Andersen: Hs is symmetric positive
    definite and satisfies both secant conditions Hs z = s and Hs zt = st
    with st = grad f*(z) and zt = grad f(s) (the reference's sign convention),
    which follow from the logarithmic homogeneity <s, zt> = <z, st> = -3 alone;
    its third term has weight t = mu ||Hd - st st'/3 - r r'/(<zt, Hd zt> -
    3 mu_t^2)||_F on the unit axis z x zt, with r = Hd zt - mu_t st.  On the
    central path (s = -mu st) and under the Dual strategy it is mu Hd."""
    rng = np.random.RandomState(21 + code)
    n_pd = 0
    for _ in range(40):
        a = rng.uniform(0.2, 0.8)
        z = _exp_dual_point(rng) if code == 3 else _pow_dual_point(rng, a)
        s = _exp_primal_point(rng) if code == 3 else _pow_primal_point(rng, a)
        st, Hd = _dual_grad_H(oracle, code, z, a)
        zt = oracle.exp_gradient_primal(s) if code == 3 else oracle.pow_gradient_primal(s, a)
        mu = s @ z / 3
        Hs = oracle.nonsym_update_Hs(code, s, z, 0.7, 0, a)
        if np.allclose(Hs, mu * Hd, rtol=1e-14, atol=0):
            continue  # the reference's own guard chose the dual scaling
        n_pd += 1
        np.testing.assert_allclose(Hs, Hs.T, rtol=0, atol=0)
        assert np.linalg.eigvalsh(Hs).min() > 0
        np.testing.assert_allclose(Hs @ z, s, rtol=1e-10, atol=1e-12 * np.abs(s).max())
        np.testing.assert_allclose(Hs @ zt, st, rtol=1e-9, atol=1e-11 * np.abs(st).max())
        mut = zt @ st / 3
        r = Hd @ zt - mut * st
        t = mu * np.linalg.norm(Hd - np.outer(st, st) / 3 - np.outer(r, r) / (zt @ Hd @ zt - 3 * mut * mut))
        ax = np.cross(z, zt)
        ax /= np.linalg.norm(ax)
        ax_s = np.outer(s, s) / (s @ z) + np.outer(s + mu * st, s + mu * st) / ((s + mu * st) @ (z + mu * zt))
        assert ax @ (Hs - ax_s) @ ax == pytest.approx(t, rel=1e-9)
        # the Dual strategy: mu Hd with the mu passed in
        np.testing.assert_allclose(oracle.nonsym_update_Hs(code, s, z, 0.7, 1, a), 0.7 * Hd, rtol=1e-15)
        # on the central path the guard falls back to the dual scaling with the local mu
        s_c = -mu * st
        np.testing.assert_allclose(oracle.nonsym_update_Hs(code, s_c, z, 0.7, 0, a), (s_c @ z / 3) * Hd,
                                   rtol=1e-14)
    assert n_pd >= 30


def _soc_point(rng, d):
    v = rng.randn(d - 1)
    return np.concatenate([[np.linalg.norm(v) + rng.uniform(0.1, 1)], v])


def test_soc_nt_scaling_identities(oracle):
    """W z = W^{-1} s = lambda; W^2 from the sparse (d,u,v) form equals
    eta^2 (2 w w' - J) (test/UnitTests/test_coneops_secondordercone.jl:31-91)."""
    rng = np.random.RandomState(10)
    for d in [2, 3, 5, 8]:
        s, z = _soc_point(rng, d), _soc_point(rng, d)
        r = oracle.soc_update_scaling(s, z)
        assert r["ok"]
        w, eta = r["w"], r["eta"]
        w1 = w[1:]
        W = np.zeros((d, d))
        W[0, 0] = w[0]; W[0, 1:] = w1; W[1:, 0] = w1
        W[1:, 1:] = np.eye(d - 1) + np.outer(w1, w1) / (1 + w[0])
        W *= eta
        np.testing.assert_allclose(W @ z, r["lam"], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(np.linalg.solve(W, s), r["lam"], rtol=1e-12, atol=1e-12)
        J = -np.eye(d); J[0, 0] = 1
        W2 = eta ** 2 * (2 * np.outer(w, w) - J)
        np.testing.assert_allclose(W @ W, W2, rtol=1e-12, atol=1e-12)
        D = np.eye(d); D[0, 0] = r["d"]
        W2s = eta ** 2 * (D + np.outer(r["u"], r["u"]) - np.outer(r["v"], r["v"]))
        np.testing.assert_allclose(W2s, W2, rtol=1e-12, atol=1e-12)


def test_soc_step_length_bruteforce(oracle):
    rng = np.random.RandomState(11)
    for d in [2, 3, 4, 6]:
        for _ in range(30):
            x = _soc_point(rng, d)
            y = rng.randn(d) * rng.uniform(0.1, 5)
            a = oracle.soc_step_length_component(x, y, 1e9)
            inside = lambda t: (x[0] + t * y[0]) >= np.linalg.norm(x[1:] + t * y[1:])
            if a < 1e9:
                assert inside(a * (1 - 1e-9)) and not inside(a * (1 + 1e-6) + 1e-12)
            else:
                assert inside(1e6)


def _svec(M):
    n = M.shape[0]
    return np.array([M[r, c] * (1 if r == c else np.sqrt(2)) for c in range(n) for r in range(c + 1)])


def test_psd_nt_identities(oracle):
    """R' Z R = Lambda = R^{-1} S R^{-T}, Hs v = W' W v with W x = svec(R' X R)
    (test/UnitTests/test_coneops_psdtrianglecone.jl:112-139, :213-253)."""
    rng = np.random.RandomState(12)
    for n in [2, 3, 5]:
        Ws, Wz = rng.randn(n, n), rng.randn(n, n)
        S = Ws @ Ws.T + 0.5 * np.eye(n)
        Z = Wz @ Wz.T + 0.5 * np.eye(n)
        r = oracle.psd_update_scaling(n, _svec(S), _svec(Z))
        assert r["ok"]
        R, Rinv, lam = r["R"], r["Rinv"], r["lam"]
        np.testing.assert_allclose(R.T @ Z @ R, np.diag(lam), atol=1e-10)
        np.testing.assert_allclose(Rinv @ S @ Rinv.T, np.diag(lam), atol=1e-10)
        np.testing.assert_allclose(R @ Rinv, np.eye(n), atol=1e-10)
        # Hs = skron(R R') acts as v -> svec(R R' V R R')
        V = rng.randn(n, n); V = V + V.T
        RRt = R @ R.T
        np.testing.assert_allclose(r["Hs"] @ _svec(V), _svec(RRt @ V @ RRt), atol=1e-9)


def test_sym_eigvals_vs_lapack(oracle):
    rng = np.random.RandomState(13)
    for n in [1, 2, 4, 10]:
        M = rng.randn(n, n); M = M + M.T
        np.testing.assert_allclose(oracle.sym_eigvals(M), np.linalg.eigvalsh(M), atol=1e-12)


# ------------------------------------------------------------------ LDL
def test_ldl_backward_error_and_inertia(oracle):
    rng = np.random.RandomState(14)
    n, m = 30, 40
    P = sp.random(n, n, density=0.2, random_state=rng); P = P @ P.T + sp.eye(n)
    A = sp.random(m, n, density=0.15, random_state=rng) + sp.eye(m, n)
    H = sp.diags(rng.uniform(0.1, 2, m))
    K = sp.bmat([[P, A.T], [A, -H]]).tocsc()
    dsigns = np.concatenate([np.ones(n), -np.ones(m)])
    b = rng.randn(n + m)
    x, nnzL = oracle.ldl_solve(sp.triu(K).tocsc(), dsigns, b)
    Kd = K.toarray()
    berr = np.linalg.norm(Kd @ x - b, np.inf) / (np.linalg.norm(Kd, np.inf) * np.linalg.norm(x, np.inf) + np.linalg.norm(b, np.inf))
    assert berr < 1e-13
    np.testing.assert_allclose(x, sla.solve(Kd, b), rtol=1e-8, atol=1e-10)
    # quasi-definite: inertia (n positive, m negative)
    ev = np.linalg.eigvalsh(Kd)
    assert (ev > 0).sum() == n and (ev < 0).sum() == m
    assert nnzL >= sp.tril(K, -1).nnz


@pytest.mark.parametrize("keep", [False, True])
def test_ldl_dynamic_regularization_rules(oracle, keep):
    """A pivot that elimination leaves on the wrong side of zero: K = [[1, 1],
    [1, 1.5]] with signs (+1, -1) gives L = [[1, 0], [1, 1]] and a second
    pivot 1.5 - 1 = +0.5.  The reference's rule (QDLDL) replaces it by
    -delta, the engine's default by -max(delta, 0.5); the solve is then
    exactly that of L diag(1, d2) L^T, [[1, 1], [1, 1 + d2]]."""
    K = sp.csc_matrix(np.array([[1.0, 1.0], [0.0, 1.5]]))
    b = np.array([0.3, -0.7])
    delta = 2e-7
    x, _ = oracle.ldl_solve(K, np.array([1, -1]), b, delta=delta, keep_magnitude=keep)
    d2 = -0.5 if keep else -delta
    np.testing.assert_allclose(x, np.linalg.solve(np.array([[1.0, 1.0], [1.0, 1.0 + d2]]), b), rtol=1e-9)
    # a wrongly signed pivot below delta takes -delta under both rules
    K2 = sp.csc_matrix(np.array([[1.0, 1.0], [0.0, 1.0 + 1e-9]]))
    x2, _ = oracle.ldl_solve(K2, np.array([1, -1]), b, delta=delta, keep_magnitude=keep)
    np.testing.assert_allclose(x2, np.linalg.solve(np.array([[1.0, 1.0], [1.0, 1.0 - delta]]), b), rtol=1e-6)
