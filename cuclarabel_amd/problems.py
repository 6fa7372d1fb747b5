"""Deterministic synthetic workloads (numpy RandomState, a fixed MT19937
stream) in Clarabel's standard form  min 1/2 x'Px + q'x, Ax + s = b, s in K.

Each builder returns (P, q, A, b, cones) with P, A as scipy CSC matrices and
cones as cuclarabel_amd cone objects.  Rows are ordered zero, nonnegative,
second-order (the order the reference's GPU path requires,
src/gpucones/compositecone_type_gpu.jl:75-80).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import NonnegativeConeT, SecondOrderConeT, ZeroConeT


def portfolio(seed: int = 1, n_assets: int = 50, k: int = 5):
    """Factor-model portfolio with transaction costs (README.md:62-97):
    variables [x (n); y (k); s (n); t (n)],
    min x'Dx + y'y - mu'x  s.t. y = F'x, sum(x) + 1e-3 sum(s) = 1, t = -x, x >= 0,
    (s_i, t_i) in SOC(2)."""
    rng = np.random.RandomState(seed)
    n = n_assets
    D = rng.uniform(0, 1, n) * np.sqrt(k)
    F = sp.random(n, k, density=0.5, random_state=rng, data_rvs=rng.standard_normal).tocsc()
    mu = (3 + 9 * rng.uniform(0, 1, n)) / 100
    nv = n + k + n + n
    ix, iy, is_, it = 0, n, n + k, 2 * n + k
    P = sp.block_diag([sp.diags(2 * D), 2 * sp.eye(k), sp.csc_matrix((2 * n, 2 * n))]).tocsc()
    q = np.concatenate([-mu, np.zeros(k + 2 * n)])
    rows, b = [], []
    # zero cone: F'x - y = 0 ; sum(x) + a sum(s) = 1 ; x + t = 0
    Z1 = sp.hstack([F.T, -sp.eye(k), sp.csc_matrix((k, 2 * n))])
    Z2 = sp.csc_matrix(np.concatenate([np.ones(n), np.zeros(k), 1e-3 * np.ones(n), np.zeros(n)])[None, :])
    Z3 = sp.hstack([sp.eye(n), sp.csc_matrix((n, k + n)), sp.eye(n)])
    rows += [Z1, Z2, Z3]
    b += [np.zeros(k), [1.0], np.zeros(n)]
    # nonnegative: -x + s = 0
    rows.append(sp.hstack([-sp.eye(n), sp.csc_matrix((n, nv - n))]))
    b.append(np.zeros(n))
    # SOC(2) on (s_i, t_i): slack = (s_i, t_i)
    S = sp.lil_matrix((2 * n, nv))
    for i in range(n):
        S[2 * i, is_ + i] = -1.0
        S[2 * i + 1, it + i] = -1.0
    rows.append(S.tocsc())
    b.append(np.zeros(2 * n))
    A = sp.vstack(rows).tocsc()
    cones = [ZeroConeT(k + 1 + n), NonnegativeConeT(n)] + [SecondOrderConeT(2) for _ in range(n)]
    return P, q, A, np.concatenate(b), cones


def mpc(seed: int, T: int | None = None, nx: int = 12, nu: int = 4, rng=None):
    """Model-predictive control QP/SOCP over horizon T (T ~ U{10..50} when not
    given): x_{t+1} = Ad x_t + Bd u_t, x_0 = x_init, |x_t| <= xmax (t >= 1),
    |u_t| <= umax, (u3; u1, u2) in SOC(3) per step, cost 1/2 sum |x|^2 + |u|^2.
    Ad = I + 0.05 N(0,1), Bd ~ N(0, 0.1), bounds ~ U(1,5), x_init ~ N(0,1)."""
    if rng is None:
        rng = np.random.RandomState(seed)
    if T is None:
        T = int(rng.randint(10, 51))
    Ad = np.eye(nx) + 0.05 * rng.standard_normal((nx, nx))
    Bd = 0.1 * rng.standard_normal((nx, nu))
    xmax = rng.uniform(1, 5, nx)
    umax = rng.uniform(1, 5, nu)
    x0 = rng.standard_normal(nx)
    nX, nU = nx * (T + 1), nu * T
    n = nX + nU
    P = sp.eye(n, format="csc")
    q = np.zeros(n)
    # zero cone: x_0 = x_init ; x_{t+1} - Ad x_t - Bd u_t = 0
    t = np.arange(T)
    ii = np.arange(nx)
    r_init, c_init, v_init = ii, ii, np.ones(nx)
    rows_dyn = nx + nx * t[:, None] + ii[None, :]                      # (T, nx)
    r1 = rows_dyn.ravel()
    c1 = (nx * (t[:, None] + 1) + ii[None, :]).ravel()
    v1 = np.ones(T * nx)
    r2 = np.repeat(rows_dyn[:, :, None], nx, axis=2).ravel()
    c2 = (nx * t[:, None, None] + np.arange(nx)[None, None, :] + 0 * ii[None, :, None]).ravel()
    v2 = np.broadcast_to(-Ad[None, :, :], (T, nx, nx)).ravel()
    r3 = np.repeat(rows_dyn[:, :, None], nu, axis=2).ravel()
    c3 = (nX + nu * t[:, None, None] + np.arange(nu)[None, None, :] + 0 * ii[None, :, None]).ravel()
    v3 = np.broadcast_to(-Bd[None, :, :], (T, nx, nu)).ravel()
    nzero = nx * (T + 1)
    # nonnegative boxes: +-x_t <= xmax (t = 1..T), +-u_t <= umax
    xcols = (nx * (np.arange(1, T + 1)[:, None]) + ii[None, :]).ravel()
    ucols = (nX + nu * t[:, None] + np.arange(nu)[None, :]).ravel()
    bcols = np.concatenate([xcols, ucols])
    nb = bcols.size
    r4 = nzero + np.arange(2 * nb)
    c4 = np.repeat(bcols, 2)
    v4 = np.tile([1.0, -1.0], nb)
    bb = np.concatenate([np.repeat(np.tile(xmax, T), 2), np.repeat(np.tile(umax, T), 2)])
    nnn = 2 * nb
    # SOC(3) per step: slack = (u3, u1, u2)
    r5 = nzero + nnn + np.arange(3 * T)
    c5 = (nX + nu * t[:, None] + np.array([nu - 1, 0, 1])[None, :]).ravel()
    v5 = -np.ones(3 * T)
    rows = np.concatenate([r_init, r1, r2, r3, r4, r5])
    cols = np.concatenate([c_init, c1, c2, c3, c4, c5])
    vals = np.concatenate([v_init, v1, v2, v3, v4, v5])
    m = nzero + nnn + 3 * T
    A = sp.csc_matrix((vals, (rows, cols)), shape=(m, n))
    b = np.concatenate([x0, np.zeros(T * nx), bb, np.zeros(3 * T)])
    cones = [ZeroConeT(nzero), NonnegativeConeT(nnn)] + [SecondOrderConeT(3) for _ in range(T)]
    return P, q, A, b, cones


def mpc_stream(seed: int = 5, count: int = 10000):
    """The request stream of the serving workload: `count` MPC problems from
    one seeded generator."""
    rng = np.random.RandomState(seed)
    for _ in range(count):
        yield mpc(0, rng=rng)


def large_socp(seed: int = 2, n: int = 20000, m_zero: int = 5000, m_nn: int = 15000, n_soc: int = 2000,
               soc_dim: int = 5, nnz_row: int = 8):
    """Random sparse SOCP: P = diag(U(0.1,1)), q ~ N(0,1), A with nnz_row
    distinct N(0,1) entries per row, b = A x0 + s0 with s0 strictly inside K.
    soc_dim is the dimension of every second-order cone, or a sequence of
    dimensions that the cones take in turn (cone k has soc_dim[k % len])."""
    rng = np.random.RandomState(seed)
    dims = [soc_dim] * n_soc if np.isscalar(soc_dim) else [int(soc_dim[k % len(soc_dim)]) for k in range(n_soc)]
    m = m_zero + m_nn + sum(dims)
    P = sp.diags(rng.uniform(0.1, 1, n)).tocsc()
    q = rng.standard_normal(n)
    cols = np.empty((m, nnz_row), np.int64)
    for i in range(m):
        cols[i] = rng.choice(n, nnz_row, replace=False)
    vals = rng.standard_normal((m, nnz_row))
    A = sp.csc_matrix((vals.ravel(), (np.repeat(np.arange(m), nnz_row), cols.ravel())), shape=(m, n))
    x0 = rng.standard_normal(n)
    s0 = np.zeros(m)
    s0[m_zero:m_zero + m_nn] = rng.uniform(0.1, 1, m_nn)
    o = m_zero + m_nn
    for d in dims:
        v = rng.standard_normal(d - 1)
        s0[o] = np.linalg.norm(v) + rng.uniform(0.1, 1)
        s0[o + 1:o + d] = v
        o += d
    b = A @ x0 + s0
    cones = [ZeroConeT(m_zero), NonnegativeConeT(m_nn)] + [SecondOrderConeT(d) for d in dims]
    return P, q, A, b, cones


def psd_relaxation(seed: int = 4, n: int = 30, n_psd: int = 6, side: int = 4, m_nn: int = 10,
                   density: float = 0.2):
    """Random SDP relaxation (the shape of the batched equal-dimension PSD
    workload, at any size): x in R^n, `n_psd` PSD cones of side `side` and
    `m_nn` nonnegative rows.  Each PSD block's rows are the svec of symmetric
    matrices with N(0,1) entries at the given density; nonnegative rows are
    sparse N(0,1).  b = A x0 + s0 with s0 interior (I + 0.1 W W' per cone,
    U(0.1, 1) for the nonnegative rows) and q = -A' z0 with z0 interior, so the
    problem is feasible and bounded; P = 0.01 I."""
    from . import NonnegativeConeT, PSDTriangleConeT
    rng = np.random.RandomState(seed)
    t = side * (side + 1) // 2
    iu = [(r, c) for c in range(side) for r in range(c + 1)]  # svec order: by columns, upper triangle
    s2 = np.sqrt(2.0)

    def svec(M):
        return np.array([M[r, c] if r == c else s2 * M[r, c] for (r, c) in iu])

    blocks, s0, z0 = [], [], []
    for _ in range(n_psd):
        cols = []
        for _j in range(n):
            M = np.where(rng.uniform(size=(side, side)) < density, rng.standard_normal((side, side)), 0.0)
            cols.append(svec((M + M.T) / 2))
        blocks.append(np.array(cols).T)  # t x n
        W = rng.standard_normal((side, side))
        s0.append(svec(np.eye(side) + 0.1 * W @ W.T))
        V = rng.standard_normal((side, side))
        z0.append(svec(np.eye(side) + 0.1 * V @ V.T))
    Ann = np.where(rng.uniform(size=(m_nn, n)) < density, rng.standard_normal((m_nn, n)), 0.0)
    A = sp.csc_matrix(np.vstack([Ann] + blocks))
    s0 = np.concatenate([rng.uniform(0.1, 1, m_nn)] + s0)
    z0 = np.concatenate([rng.uniform(0.1, 1, m_nn)] + z0)
    x0 = rng.standard_normal(n)
    b = A @ x0 + s0
    q = -(A.T @ z0)
    P = 0.01 * sp.eye(n, format="csc")
    cones = [NonnegativeConeT(m_nn)] + [PSDTriangleConeT(side) for _ in range(n_psd)]
    return P, q, A, b, cones


def logistic_exp_pow(seed: int = 3, n_samples: int = 5000, n_features: int = 200, n_pow: int = 2000,
                     l1: float = 0.1, entropy: float = 0.05, pow_weight: float = 0.01):
    """L1-regularized logistic regression with an entropy term and power-cone
    penalties (the exp/pow-cone workload, at any size).  Features F ~ N(0,1),
    labels y = +-1 Bernoulli(0.5), a_i = -y_i F_i' theta.

    variables [theta (f); r (f); t, u, v, h (one each per sample); p, q (per power cone)]
    min  sum t + l1 sum r - entropy sum h + pow_weight sum (p + q)
    s.t. u_i + v_i <= 1, -r <= theta <= r                      (nonnegative rows)
         (-t_i, 1, u_i), (a_i - t_i, 1, v_i) in K_exp          (log(1 + e^{a_i}) <= t_i)
         (h_i, u_i, 1) in K_exp                                (h_i <= -u_i log u_i)
         (p_k, q_k, theta_{k mod f}) in K_pow(alpha_k), alpha_k ~ U(0.2, 0.8)
    Rows: 2f + n_samples nonnegative, 9 n_samples exponential, 3 n_pow power."""
    from . import ExponentialConeT, NonnegativeConeT, PowerConeT
    rng = np.random.RandomState(seed)
    S, f, K = n_samples, n_features, n_pow
    F = rng.standard_normal((S, f))
    y = np.where(rng.uniform(size=S) < 0.5, -1.0, 1.0)
    alpha = rng.uniform(0.2, 0.8, K)
    o_th, o_r = 0, f
    o_t, o_u, o_v, o_h = 2 * f, 2 * f + S, 2 * f + 2 * S, 2 * f + 3 * S
    o_p, o_q = 2 * f + 4 * S, 2 * f + 4 * S + K
    n = 2 * f + 4 * S + 2 * K
    rows, cols, vals, b = [], [], [], []
    row = [0]

    def put(entries, rhs):
        for c, v in entries:
            rows.append(row[0]); cols.append(c); vals.append(v)
        b.append(rhs)
        row[0] += 1

    smp = np.arange(S)
    for i in smp:  # s = 1 - u - v >= 0
        put([(o_u + i, 1.0), (o_v + i, 1.0)], 1.0)
    for j in range(f):  # s = r - theta >= 0, s = r + theta >= 0
        put([(o_th + j, 1.0), (o_r + j, -1.0)], 0.0)
        put([(o_th + j, -1.0), (o_r + j, -1.0)], 0.0)
    for i in smp:
        put([(o_t + i, 1.0)], 0.0)                       # s = -t
        put([], 1.0)                                     # s = 1
        put([(o_u + i, -1.0)], 0.0)                      # s = u
        put([(o_th + j, y[i] * F[i, j]) for j in range(f)] + [(o_t + i, 1.0)], 0.0)  # s = a - t
        put([], 1.0)
        put([(o_v + i, -1.0)], 0.0)                      # s = v
        put([(o_h + i, -1.0)], 0.0)                      # s = h
        put([(o_u + i, -1.0)], 0.0)                      # s = u
        put([], 1.0)
    for k in range(K):
        put([(o_p + k, -1.0)], 0.0)
        put([(o_q + k, -1.0)], 0.0)
        put([(o_th + k % f, -1.0)], 0.0)
    m = row[0]
    A = sp.csc_matrix((vals, (rows, cols)), shape=(m, n))
    q = np.zeros(n)
    q[o_t:o_t + S] = 1.0
    q[o_r:o_r + f] = l1
    q[o_h:o_h + S] = -entropy
    q[o_p:o_p + 2 * K] = pow_weight
    P = sp.csc_matrix((n, n))
    cones = ([NonnegativeConeT(S + 2 * f)] + [ExponentialConeT() for _ in range(3 * S)]
             + [PowerConeT(float(a)) for a in alpha])
    return P, q, A, np.array(b), cones
