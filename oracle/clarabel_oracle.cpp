// CPU oracle: a from-scratch C++ restatement of the Clarabel.jl v0.10.0 CPU
// path (QDLDL direct solve, presolve and chordal decomposition off).
//
// TEST INFRASTRUCTURE ONLY -- the cuclarabel_amd package never links, loads
// or calls this code.  Used by tests/ (as the checker), by
// __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.
//
// Every function cites the reference file:line range it restates (paths are
// relative to the reference repository root).
#include "clarabel_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iterator>
#include <limits>
#include <set>
#include <vector>

namespace {

using Vec = std::vector<double>;
using IVec = std::vector<int>;
constexpr double kFloatMax = std::numeric_limits<double>::max();
constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kEps = std::numeric_limits<double>::epsilon();

// src/utils/mathutils.jl:4-10
inline double clip(double s, double lo, double hi) { return s < lo ? lo : (s > hi ? hi : s); }
// src/utils/mathutils.jl:12-18
inline double logsafe(double v) { return v < 0 ? -kInf : std::log(v); }

inline double dotv(const double* a, const double* b, int n) {
  double t = 0;
  for (int i = 0; i < n; ++i) t += a[i] * b[i];
  return t;
}
inline double norm2(const double* a, int n) { return std::sqrt(dotv(a, a, n)); }
inline double norm_inf(const double* a, int n) {
  double t = 0;
  for (int i = 0; i < n; ++i) t = std::max(t, std::fabs(a[i]));
  return t;
}
// src/utils/mathutils.jl:58-65
inline double norm_scaled(const Vec& m, const Vec& v) {
  double t = 0;
  for (size_t i = 0; i < v.size(); ++i) { double p = m[i] * v[i]; t += p * p; }
  return std::sqrt(t);
}
// src/utils/mathutils.jl:68-74
inline double norm_inf_scaled(const Vec& m, const Vec& v) {
  double t = 0;
  for (size_t i = 0; i < v.size(); ++i) t = std::max(t, std::fabs(m[i] * v[i]));
  return t;
}

struct Csc {
  int m = 0, n = 0;
  IVec p, i;
  Vec x;
  int nnz() const { return p.empty() ? 0 : p[n]; }
};

// y = alpha * Sym(P) x + beta*y, with only the triu part of P stored
void symv_triu(const Csc& P, const double* x, double* y, double alpha, double beta) {
  for (int r = 0; r < P.n; ++r) y[r] *= beta;
  for (int j = 0; j < P.n; ++j) {
    for (int k = P.p[j]; k < P.p[j + 1]; ++k) {
      int r = P.i[k];
      double v = P.x[k];
      if (r == j) {
        y[r] += alpha * v * x[j];
      } else if (r < j) {
        y[r] += alpha * v * x[j];
        y[j] += alpha * v * x[r];
      }
    }
  }
}
// y = alpha*A x + beta*y
void gemv_csc(const Csc& A, const double* x, double* y, double alpha, double beta) {
  for (int r = 0; r < A.m; ++r) y[r] *= beta;
  for (int j = 0; j < A.n; ++j)
    for (int k = A.p[j]; k < A.p[j + 1]; ++k) y[A.i[k]] += alpha * A.x[k] * x[j];
}
// y = alpha*A' x + beta*y
void gemv_csc_t(const Csc& A, const double* x, double* y, double alpha, double beta) {
  for (int j = 0; j < A.n; ++j) {
    double t = 0;
    for (int k = A.p[j]; k < A.p[j + 1]; ++k) t += A.x[k] * x[A.i[k]];
    y[j] = alpha * t + beta * y[j];
  }
}
// x' Sym(P) y with triu P (src/utils/mathutils.jl:241-279)
double quad_form(const Vec& x, const Csc& P, const Vec& y) {
  double out = 0;
  for (int j = 0; j < P.n; ++j) {
    double t1 = 0, t2 = 0;
    for (int k = P.p[j]; k < P.p[j + 1]; ++k) {
      int i = P.i[k];
      if (i < j) { t1 += P.x[k] * x[i]; t2 += P.x[k] * y[i]; }
      else if (i == j) out += P.x[k] * x[i] * y[i];
    }
    out += t1 * y[j] + t2 * x[j];
  }
  return out;
}

inline int tri(int k) { return k * (k + 1) / 2; }

// ---------------------------------------------------------------------------
// small dense linear algebra (column-major) for the PSD cone
// ---------------------------------------------------------------------------
// Cholesky A = L L' (lower), returns false on a non-positive pivot
bool chol(int n, const double* A, double* L) {
  std::fill(L, L + n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    double s = A[j + j * n];
    for (int k = 0; k < j; ++k) s -= L[j + k * n] * L[j + k * n];
    if (!(s > 0)) return false;
    double d = std::sqrt(s);
    L[j + j * n] = d;
    for (int i = j + 1; i < n; ++i) {
      double t = A[i + j * n];
      for (int k = 0; k < j; ++k) t -= L[i + k * n] * L[j + k * n];
      L[i + j * n] = t / d;
    }
  }
  return true;
}

// cyclic Jacobi eigenvalues of a symmetric matrix, ascending order
void sym_eig(int n, const double* M, double* ev, double* V /*may be null*/) {
  std::vector<double> A(M, M + n * n);
  if (V) { std::fill(V, V + n * n, 0.0); for (int i = 0; i < n; ++i) V[i + i * n] = 1.0; }
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double a = A[i + j * n] * A[i + j * n];
        tot += a;
        if (i != j) off += a;
      }
    if (off <= 1e-32 * tot || off == 0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p + q * n];
        if (apq == 0) continue;
        double app = A[p + p * n], aqq = A[q + q * n];
        double theta = (aqq - app) / (2 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p,q
          double akp = A[k + p * n], akq = A[k + q * n];
          A[k + p * n] = c * akp - s * akq;
          A[k + q * n] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p,q
          double apk = A[p + k * n], aqk = A[q + k * n];
          A[p + k * n] = c * apk - s * aqk;
          A[q + k * n] = s * apk + c * aqk;
        }
        if (V)
          for (int k = 0; k < n; ++k) {
            double vkp = V[k + p * n], vkq = V[k + q * n];
            V[k + p * n] = c * vkp - s * vkq;
            V[k + q * n] = s * vkp + c * vkq;
          }
      }
  }
  for (int i = 0; i < n; ++i) ev[i] = A[i + i * n];
  std::sort(ev, ev + n);
}

// one-sided (Hestenes) Jacobi SVD: A = U diag(S) V', S descending
void svd_jacobi(int n, const double* A, double* U, double* S, double* V) {
  std::copy(A, A + n * n, U);
  std::fill(V, V + n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i + i * n] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double al = 0, be = 0, ga = 0;
        for (int k = 0; k < n; ++k) {
          al += U[k + p * n] * U[k + p * n];
          be += U[k + q * n] * U[k + q * n];
          ga += U[k + p * n] * U[k + q * n];
        }
        if (std::fabs(ga) <= 1e-15 * std::sqrt(al * be) || ga == 0) continue;
        rotated = true;
        double zeta = (be - al) / (2 * ga);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
        double c = 1 / std::sqrt(1 + t * t), s = c * t;
        for (int k = 0; k < n; ++k) {
          double up = U[k + p * n], uq = U[k + q * n];
          U[k + p * n] = c * up - s * uq;
          U[k + q * n] = s * up + c * uq;
          double vp = V[k + p * n], vq = V[k + q * n];
          V[k + p * n] = c * vp - s * vq;
          V[k + q * n] = s * vp + c * vq;
        }
      }
    if (!rotated) break;
  }
  for (int j = 0; j < n; ++j) {
    double s = 0;
    for (int k = 0; k < n; ++k) s += U[k + j * n] * U[k + j * n];
    s = std::sqrt(s);
    S[j] = s;
    if (s > 0)
      for (int k = 0; k < n; ++k) U[k + j * n] /= s;
  }
  // sort descending
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return S[a] > S[b]; });
  std::vector<double> U2(n * n), V2(n * n), S2(n);
  for (int j = 0; j < n; ++j) {
    S2[j] = S[idx[j]];
    for (int k = 0; k < n; ++k) { U2[k + j * n] = U[k + idx[j] * n]; V2[k + j * n] = V[k + idx[j] * n]; }
  }
  std::copy(U2.begin(), U2.end(), U);
  std::copy(V2.begin(), V2.end(), V);
  std::copy(S2.begin(), S2.end(), S);
}

// C = alpha * op(A) op(B) + beta*C, all n x n column-major
void gemm(int n, bool ta, bool tb, double alpha, const double* A, const double* B, double beta, double* C) {
  std::vector<double> T(n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double s = 0;
      for (int k = 0; k < n; ++k) {
        double a = ta ? A[k + i * n] : A[i + k * n];
        double b = tb ? B[j + k * n] : B[k + j * n];
        s += a * b;
      }
      T[i + j * n] = s;
    }
  for (int i = 0; i < n * n; ++i) C[i] = alpha * T[i] + (beta == 0 ? 0.0 : beta * C[i]);
}

// src/cones/coneops_psdtrianglecone.jl:476-490
void svec_to_mat(int n, const double* x, double* M) {
  const double is2 = 1 / std::sqrt(2.0);
  int idx = 0;
  for (int col = 0; col < n; ++col)
    for (int row = 0; row <= col; ++row) {
      if (row == col) M[row + col * n] = x[idx];
      else { M[row + col * n] = x[idx] * is2; M[col + row * n] = x[idx] * is2; }
      ++idx;
    }
}
// src/cones/coneops_psdtrianglecone.jl:493-504
void mat_to_svec(int n, const double* M, double* x) {
  const double is2 = 1 / std::sqrt(2.0);
  int idx = 0;
  for (int col = 0; col < n; ++col)
    for (int row = 0; row <= col; ++row) {
      x[idx] = (row == col) ? M[row + col * n] : (M[row + col * n] + M[col + row * n]) * is2;
      ++idx;
    }
}
// src/cones/coneops_psdtrianglecone.jl:509-547 (full t x t output)
void skron(int n, const double* A, double* out) {
  const double sqrt2 = std::sqrt(2.0);
  int t = tri(n);
  int col = 0;
  for (int l = 0; l < n; ++l)
    for (int k = 0; k <= l; ++k) {
      int row = 0;
      bool kl_eq = (k == l);
      for (int j = 0; j < n; ++j) {
        double Ajl = A[j + l * n], Ajk = A[j + k * n];
        for (int i = 0; i <= j; ++i) {
          if (row > col) break;
          bool ij_eq = (i == j);
          double v;
          if (!ij_eq && !kl_eq) v = A[i + k * n] * Ajl + A[i + l * n] * Ajk;
          else if (ij_eq && !kl_eq) v = sqrt2 * Ajl * Ajk;
          else if (!ij_eq && kl_eq) v = sqrt2 * A[i + l * n] * Ajk;
          else v = Ajl * Ajl;
          out[row + col * t] = v;
          out[col + row * t] = v;
          ++row;
        }
      }
      ++col;
    }
}

// ---------------------------------------------------------------------------
// cones
// ---------------------------------------------------------------------------
struct Cone {
  int type = 0, dim = 0, n = 0;  // dim = numel; n = PSD side
  double alpha = 0;              // power cone parameter
  // NN / SOC
  Vec w, lam;
  double eta = 0;
  bool sparse = false;
  double sd = 0;
  Vec su, sv;
  // exp / pow
  double Hd[9] = {0}, Hs3[9] = {0}, grad[3] = {0}, zc[3] = {0};
  // PSD
  Vec R, Rinv, Hsp, lamp, lisq;
  bool symmetric() const { return type != ORC_EXP && type != ORC_POW; }
  int degree() const {  // src/cones/coneops_*.jl degree()
    switch (type) {
      case ORC_ZERO: return 0;
      case ORC_NONNEG: return dim;
      case ORC_SOC: return 1;
      case ORC_EXP: case ORC_POW: return 3;
      default: return n;
    }
  }
  bool hs_diagonal() const {  // Hs_is_diagonal (coneops_defaults.jl:119-123, coneops_socone.jl:195-199)
    if (type == ORC_ZERO || type == ORC_NONNEG) return true;
    if (type == ORC_SOC) return sparse;
    return false;
  }
  int hs_len() const { return hs_diagonal() ? dim : tri(dim); }
};

enum Scaling { PrimalDual = 0, Dual = 1 };

// ---- SOC helpers (src/cones/coneops_socone.jl:419-517)
double soc_residual(const double* z, int d) {
  double z1 = norm2(z + 1, d - 1);
  return (z[0] - z1) * (z[0] + z1);
}
double sqrt_soc_residual(const double* z, int d) {
  double r = soc_residual(z, d);
  return r > 0 ? std::sqrt(r) : 0.0;
}
double dot_shifted(const double* z, const double* s, const double* dz, const double* ds, double a, int n) {
  double out = 0;
  for (int i = 0; i < n; ++i) out += (z[i] + a * dz[i]) * (s[i] + a * ds[i]);
  return out;
}
double soc_residual_shifted(const double* z, const double* dz, double a, int d) {
  double x0 = z[0] + a * dz[0];
  double x1sq = dot_shifted(z + 1, z + 1, dz + 1, dz + 1, a, d - 1);
  double x1n = std::sqrt(x1sq);
  return (x0 - x1n) * (x0 + x1n);
}
double soc_step_component(const double* x, const double* y, double amax, int d) {
  if (x[0] >= 0 && y[0] < 0) amax = std::min(amax, -x[0] / y[0]);
  double a = soc_residual(y, d);
  double b = 2 * (x[0] * y[0] - dotv(x + 1, y + 1, d - 1));
  double c = std::max(0.0, soc_residual(x, d));
  double dd = b * b - 4 * a * c;
  if ((a > 0 && b > 0) || dd < 0) return amax;
  if (a == 0) return amax;
  if (c == 0) return a >= 0 ? amax : 0.0;
  double t = (b >= 0) ? (-b - std::sqrt(dd)) : (-b + std::sqrt(dd));
  double r1 = (2 * c) / t, r2 = t / (2 * a);
  if (r1 < 0) r1 = kFloatMax;
  if (r2 < 0) r2 = kFloatMax;
  return std::min(amax, std::min(r1, r2));
}

// SOC NT scaling (src/cones/coneops_socone.jl:75-154)
bool soc_update_scaling(Cone& K, const double* s, const double* z) {
  int d = K.dim;
  double zscale = sqrt_soc_residual(z, d), sscale = sqrt_soc_residual(s, d);
  if (zscale == 0 || sscale == 0) return false;
  K.eta = std::sqrt(sscale / zscale);
  double* w = K.w.data();
  for (int i = 0; i < d; ++i) w[i] = s[i] / sscale;
  w[0] += z[0] / zscale;
  for (int i = 1; i < d; ++i) w[i] -= z[i] / zscale;
  double wscale = sqrt_soc_residual(w, d);
  if (wscale == 0) return false;
  for (int i = 0; i < d; ++i) w[i] /= wscale;
  double w1sq = dotv(w + 1, w + 1, d - 1);
  w[0] = std::sqrt(1 + w1sq);
  double gamma = 0.5 * wscale;
  double* lam = K.lam.data();
  lam[0] = gamma;
  double c1 = (gamma + z[0] / zscale) / sscale, c2 = (gamma + s[0] / sscale) / zscale;
  for (int i = 1; i < d; ++i) lam[i] = c1 * s[i] + c2 * z[i];
  double inv = 1.0 / (s[0] / sscale + z[0] / zscale + 2 * gamma);
  for (int i = 1; i < d; ++i) lam[i] *= inv;
  double sc = std::sqrt(sscale * zscale);
  for (int i = 0; i < d; ++i) lam[i] *= sc;
  if (K.sparse) {
    double al = 2 * w[0];
    double wsq = w[0] * w[0] + w1sq;
    double wsqinv = 1 / wsq;
    K.sd = wsqinv / 2;
    double u0 = std::sqrt(wsq - K.sd);
    double u1 = al / u0;
    double v0 = 0;
    double v1 = std::sqrt(2 * (2 + wsqinv) / (2 * wsq - wsqinv));
    K.su[0] = u0;
    for (int i = 1; i < d; ++i) K.su[i] = u1 * w[i];
    K.sv[0] = v0;
    for (int i = 1; i < d; ++i) K.sv[i] = v1 * w[i];
  }
  return true;
}

// SOC W products (src/cones/coneops_socone.jl:314-361)
void soc_mul_W(const Cone& K, double* y, const double* x) {
  int d = K.dim;
  double zeta = dotv(K.w.data() + 1, x + 1, d - 1);
  double c = x[0] + zeta / (1 + K.w[0]);
  y[0] = K.eta * (K.w[0] * x[0] + zeta);
  for (int i = 1; i < d; ++i) y[i] = K.eta * (x[i] + c * K.w[i]);
}
void soc_mul_Winv(const Cone& K, double* y, const double* x) {
  int d = K.dim;
  double zeta = dotv(K.w.data() + 1, x + 1, d - 1);
  double c = -x[0] + zeta / (1 + K.w[0]);
  y[0] = (1 / K.eta) * (K.w[0] * x[0] - zeta);
  for (int i = 1; i < d; ++i) y[i] = (1 / K.eta) * (x[i] + c * K.w[i]);
}
// circ_op! (src/cones/coneops_socone.jl:380-395)
void soc_circ(double* x, const double* y, const double* z, int d) {
  double x0 = dotv(y, z, d);
  double y0 = y[0], z0 = z[0];
  for (int i = 1; i < d; ++i) x[i] = y0 * z[i] + z0 * y[i];
  x[0] = x0;
}

// ---- exponential cone (src/cones/coneops_expcone.jl)
// _wright_omega (coneops_expcone.jl:463-519); the statement `q * zinv`
// (no assignment) at :502 is reproduced as written: q is not updated.
double wright_omega(double z) {
  double w;
  if (z < 0) return std::numeric_limits<double>::quiet_NaN();
  if (z < 1 + M_PI) {
    double zm1 = z - 1;
    double p = zm1;
    w = 1 + 0.5 * p;
    p *= zm1; w += (1 / 16.0) * p;
    p *= zm1; w -= (1 / 192.0) * p;
    p *= zm1; w -= (1 / 3072.0) * p;
    p *= zm1; w += (13 / 61440.0) * p;
  } else {
    double logz = logsafe(z);
    double zinv = 1 / z;
    w = z - logz;
    double q = logz * zinv;
    w += q;
    q *= zinv;
    w += q * (logz / 2 - 1);
    /* q * zinv : value discarded in the reference */
    w += q * (logz * logz / 3. - (3 / 2.) * logz + 1);
  }
  double r = z - w - logsafe(w);
  for (int i = 0; i < 2; ++i) {
    double wp1 = w + 1;
    double t = wp1 * (wp1 + (2. * r) / 3.0);
    w *= 1 + (r / wp1) * (t - 0.5 * r) / (t - r);
    double wp6 = wp1 * wp1 * wp1 * wp1 * wp1 * wp1;
    r = (2 * w * w - 8 * w - 1) / (72.0 * wp6) * r * r * r * r;
  }
  return w;
}
// coneops_expcone.jl:284-297
void exp_gradient_primal(const double* s, double* g) {
  double om = wright_omega(1 - s[0] / s[1] - logsafe(s[1] / s[2]));
  double g1 = 1 / ((om - 1) * s[1]);
  double g2 = g1 + g1 * logsafe(om * s[1] / s[2]) - 1 / s[1];
  double g3 = om / ((1 - om) * s[2]);
  g[0] = g1; g[1] = g2; g[2] = g3;
}
// coneops_expcone.jl:389-418  (H column-major 3x3)
void exp_dual_grad_H(const double* z, double* grad, double* H) {
  double l = logsafe(-z[2] / z[0]);
  double r = -z[0] * l - z[0] + z[1];
  double c2 = 1 / r;
  grad[0] = c2 * l - 1 / z[0];
  grad[1] = -c2;
  grad[2] = (c2 * z[0] - 1) / z[2];
  H[0] = ((r * r - z[0] * r + l * l * z[0] * z[0]) / (r * z[0] * z[0] * r));
  H[3] = (-l / (r * r)); H[1] = H[3];
  H[4] = (1 / (r * r));
  H[6] = ((z[1] - z[0]) / (r * r * z[2])); H[2] = H[6];
  H[7] = (-z[0] / (r * r * z[2])); H[5] = H[7];
  H[8] = ((r * r - z[0] * r + z[0] * z[0]) / (r * r * z[2] * z[2]));
}
// coneops_expcone.jl:253-281
bool exp_primal_feasible(const double* s) {
  if (s[2] > 0 && s[1] > 0) {
    double res = s[1] * logsafe(s[2] / s[1]) - s[0];
    if (res > 0) return true;
  }
  return false;
}
bool exp_dual_feasible(const double* z) {
  if (z[2] > 0 && z[0] < 0) {
    double res = z[1] - z[0] - z[0] * logsafe(-z[2] / z[0]);
    if (res > 0) return true;
  }
  return false;
}
// coneops_expcone.jl:223-248
double exp_barrier_dual(const double* z) {
  double l = logsafe(-z[2] / z[0]);
  return -logsafe(-z[2] * z[0]) - logsafe(z[1] - z[0] - z[0] * l);
}
double exp_barrier_primal(const double* s) {
  double om = wright_omega(1 - s[0] / s[1] - logsafe(s[1] / s[2]));
  om = (om - 1) * (om - 1) / om;
  return -logsafe(om) - 2 * logsafe(s[1]) - logsafe(s[2]) - 3;
}

// ---- power cone (src/cones/coneops_powcone.jl)
// _newton_raphson_onesided (src/cones/coneops_nonsymmetric_common.jl:170-193)
template <class F0, class F1>
double newton_onesided(double x0, F0 f0, F1 f1) {
  int iter = 0;
  double x = x0;
  while (iter < 100) {
    ++iter;
    double dfdx = f1(x);
    double dx = -f0(x) / dfdx;
    if ((dx < kEps) || (std::fabs(dx / x) < std::sqrt(kEps)) || (std::fabs(dfdx) < kEps)) break;
    x += dx;
  }
  return x;
}
// coneops_powcone.jl:449-478
double newton_powcone(double s3, double phi, double a) {
  double x0 = -1 / s3 + (2 * s3 + std::sqrt(phi * phi / s3 / s3 + 3 * phi)) / (phi - s3 * s3);
  double t0 = -2 * a * logsafe(a) - 2 * (1 - a) * logsafe(1 - a);
  auto f0 = [&](double x) {
    double t1 = x * x, t2 = 2 * x / s3;
    return 2 * a * logsafe(2 * a * t1 + (1 + a) * t2) + 2 * (1 - a) * logsafe(2 * (1 - a) * t1 + (2 - a) * t2) -
           logsafe(phi) - logsafe(t1 + t2) - 2 * logsafe(t2) + t0;
  };
  auto f1 = [&](double x) {
    double t1 = x * x, t2 = x * 2 / s3;
    return 2 * a * a / (a * x + (1 + a) / s3) + 2 * (1 - a) * (1 - a) / ((1 - a) * x + (2 - a) / s3) -
           2 * (x + 1 / s3) / (t1 + t2);
  };
  return newton_onesided(x0, f0, f1);
}
// coneops_powcone.jl:288-316
void pow_gradient_primal(const double* s, double a, double* g) {
  double phi = std::pow(s[0], 2 * a) * std::pow(s[1], 2 - 2 * a);
  double abs_s = std::fabs(s[2]);
  if (abs_s > kEps) {
    g[2] = newton_powcone(abs_s, phi, a);
    if (s[2] < 0) g[2] = -g[2];
    g[0] = -(a * g[2] * s[2] + 1 + a) / s[0];
    g[1] = -((1 - a) * g[2] * s[2] + 2 - a) / s[1];
  } else {
    g[2] = 0;
    g[0] = -(1 + a) / s[0];
    g[1] = -(2 - a) / s[1];
  }
}
// coneops_powcone.jl:408-440
void pow_dual_grad_H(const double* z, double a, double* grad, double* H) {
  double phi = std::pow(z[0] / a, 2 * a) * std::pow(z[1] / (1 - a), 2 - 2 * a);
  double psi = phi - z[2] * z[2];
  double g0 = 2 * a * phi / (z[0] * psi), g1 = 2 * (1 - a) * phi / (z[1] * psi), g2 = -2 * z[2] / psi;
  H[0] = g0 * g0 - 2 * a * (2 * a - 1) * phi / (z[0] * z[0] * psi) + (1 - a) / (z[0] * z[0]);
  H[3] = g0 * g1 - 4 * a * (1 - a) * phi / (z[0] * z[1] * psi); H[1] = H[3];
  H[4] = g1 * g1 - 2 * (1 - a) * (1 - 2 * a) * phi / (z[1] * z[1] * psi) + a / (z[1] * z[1]);
  H[6] = g0 * g2; H[2] = H[6];
  H[7] = g1 * g2; H[5] = H[7];
  H[8] = g2 * g2 + 2 / psi;
  grad[0] = -2 * a * phi / (z[0] * psi) - (1 - a) / z[0];
  grad[1] = -2 * (1 - a) * phi / (z[1] * psi) - a / z[1];
  grad[2] = 2 * z[2] / psi;
}
// coneops_powcone.jl:256-284
bool pow_primal_feasible(const double* s, double a) {
  if (s[0] > 0 && s[1] > 0) {
    double res = std::exp(2 * a * logsafe(s[0]) + 2 * (1 - a) * logsafe(s[1])) - s[2] * s[2];
    if (res > 0) return true;
  }
  return false;
}
bool pow_dual_feasible(const double* z, double a) {
  if (z[0] > 0 && z[1] > 0) {
    double res = std::exp(2 * a * logsafe(z[0] / a) + 2 * (1 - a) * logsafe(z[1] / (1 - a))) - z[2] * z[2];
    if (res > 0) return true;
  }
  return false;
}
// coneops_powcone.jl:228-251
double pow_barrier_dual(const double* z, double a) {
  return -logsafe(std::pow(z[0] / a, 2 * a) * std::pow(z[1] / (1 - a), 2 - 2 * a) - z[2] * z[2]) -
         (1 - a) * logsafe(z[0]) - a * logsafe(z[1]);
}
double pow_barrier_primal(const double* s, double a) {
  double g[3];
  pow_gradient_primal(s, a, g);
  return logsafe(std::pow(-g[0] / a, 2 * a) * std::pow(-g[1] / (1 - a), 2 - 2 * a) - g[2] * g[2]) +
         (1 - a) * logsafe(-g[0]) + a * logsafe(-g[1]) - 3;
}

// 3x3 explicit Cholesky (src/utils/mathutils.jl:381-420), column-major
bool chol3_factor(double* L, const double* A) {
  std::fill(L, L + 9, 0.0);
  double t = A[0];
  if (t <= 0) return false;
  L[0] = std::sqrt(t);
  L[1] = A[1] / L[0];
  t = A[4] - L[1] * L[1];
  if (t <= 0) return false;
  L[4] = std::sqrt(t);
  L[2] = A[2] / L[0];
  L[5] = (A[5] - L[1] * L[2]) / L[4];
  t = A[8] - L[2] * L[2] - L[5] * L[5];
  if (t <= 0) return false;
  L[8] = std::sqrt(t);
  return true;
}
void chol3_solve(const double* L, const double* b, double* x) {
  double L11 = L[0], L21 = L[1], L31 = L[2], L22 = L[4], L32 = L[5], L33 = L[8];
  double c1 = b[0] / L11;
  double c2 = (b[1] * L11 - b[0] * L21) / (L11 * L22);
  double c3 = (b[2] * L11 * L22 - b[1] * L11 * L32 + b[0] * L21 * L32 - b[0] * L22 * L31) / (L11 * L22 * L33);
  x[0] = (c1 * L22 * L33 - c2 * L21 * L33 + c3 * L21 * L32 - c3 * L22 * L31) / (L11 * L22 * L33);
  x[1] = (c2 * L33 - c3 * L32) / (L22 * L33);
  x[2] = c3 / L33;
}

// coneops_expcone.jl:338-385 (ds, v) -> eta
void exp_higher_correction(const Cone& K, double* eta, const double* ds, const double* v) {
  const double* H = K.Hd;
  const double* z = K.zc;
  double L[9], u[3];
  if (!chol3_factor(L, H)) { eta[0] = eta[1] = eta[2] = 0; return; }
  chol3_solve(L, ds, u);
  eta[1] = 1;
  eta[2] = -z[0] / z[2];
  eta[0] = logsafe(eta[2]);
  double psi = z[0] * eta[0] - z[0] + z[1];
  double dotpsiu = dotv(eta, u, 3), dotpsiv = dotv(eta, v, 3);
  double coef = ((u[0] * (v[0] / z[0] - v[2] / z[2]) + u[2] * (z[0] * v[2] / z[2] - v[0]) / z[2]) * psi -
                 2 * dotpsiu * dotpsiv) / (psi * psi * psi);
  for (int i = 0; i < 3; ++i) eta[i] *= coef;
  double inv_psi2 = 1 / psi / psi;
  eta[0] += (1 / psi - 2 / z[0]) * u[0] * v[0] / (z[0] * z[0]) - u[2] * v[2] / (z[2] * z[2]) / psi +
            dotpsiu * inv_psi2 * (v[0] / z[0] - v[2] / z[2]) + dotpsiv * inv_psi2 * (u[0] / z[0] - u[2] / z[2]);
  eta[2] += 2 * (z[0] / psi - 1) * u[2] * v[2] / (z[2] * z[2] * z[2]) - (u[2] * v[0] + u[0] * v[2]) / (z[2] * z[2]) / psi +
            dotpsiu * inv_psi2 * (z[0] * v[2] / (z[2] * z[2]) - v[0] / z[2]) +
            dotpsiv * inv_psi2 * (z[0] * u[2] / (z[2] * z[2]) - u[0] / z[2]);
  for (int i = 0; i < 3; ++i) eta[i] /= 2;
}
// coneops_powcone.jl:329-404
void pow_higher_correction(const Cone& K, double* eta, const double* ds, const double* v) {
  const double* H = K.Hd;
  const double* z = K.zc;
  double a = K.alpha;
  double L[9], u[3];
  if (!chol3_factor(L, H)) { eta[0] = eta[1] = eta[2] = 0; return; }
  chol3_solve(L, ds, u);
  double phi = std::pow(z[0] / a, 2 * a) * std::pow(z[1] / (1 - a), 2 - 2 * a);
  double psi = phi - z[2] * z[2];
  eta[0] = 2 * a * phi / z[0];
  eta[1] = 2 * (1 - a) * phi / z[1];
  eta[2] = -2 * z[2];
  double H11 = 2 * a * (2 * a - 1) * phi / (z[0] * z[0]);
  double H12 = 4 * a * (1 - a) * phi / (z[0] * z[1]);
  double H22 = 2 * (1 - a) * (1 - 2 * a) * phi / (z[1] * z[1]);
  double dotpsiu = dotv(eta, u, 3), dotpsiv = dotv(eta, v, 3);
  double Hpv[3] = {H11 * v[0] + H12 * v[1], H12 * v[0] + H22 * v[1], -2 * v[2]};
  double coef = (dotv(u, Hpv, 3) * psi - 2 * dotpsiu * dotpsiv) / (psi * psi * psi);
  double coef2 = 4 * a * (2 * a - 1) * (1 - a) * phi * (u[0] / z[0] - u[1] / z[1]) * (v[0] / z[0] - v[1] / z[1]) / psi;
  double inv_psi2 = 1 / psi / psi;
  eta[0] = coef * eta[0] - 2 * (1 - a) * u[0] * v[0] / (z[0] * z[0] * z[0]) + coef2 / z[0] + Hpv[0] * dotpsiu * inv_psi2;
  eta[1] = coef * eta[1] - 2 * a * u[1] * v[1] / (z[1] * z[1] * z[1]) - coef2 / z[1] + Hpv[1] * dotpsiu * inv_psi2;
  eta[2] = coef * eta[2] + Hpv[2] * dotpsiu * inv_psi2;
  double Hpu[3] = {H11 * u[0] + H12 * u[1], H12 * u[0] + H22 * u[1], -2 * u[2]};
  for (int i = 0; i < 3; ++i) eta[i] = (eta[i] + Hpu[i] * dotpsiv * inv_psi2) / 2;
}

// update_Hs / use_dual_scaling / use_primal_dual_scaling
// (src/cones/coneops_nonsymmetric_common.jl:50-164)
void nonsym_update_Hs(Cone& K, const double* s, const double* z, double mu, Scaling strat) {
  auto use_dual = [&](double m) { for (int i = 0; i < 9; ++i) K.Hs3[i] = m * K.Hd[i]; };
  if (strat == Dual) { use_dual(mu); return; }
  const double* Hd = K.Hd;
  double st[3] = {K.grad[0], K.grad[1], K.grad[2]};
  double zt[3];
  if (K.type == ORC_EXP) exp_gradient_primal(s, zt);
  else pow_gradient_primal(s, K.alpha, zt);
  double dot_sz = dotv(z, s, 3);
  double m = dot_sz / 3;
  double mut = dotv(zt, st, 3) / 3;
  double ds_[3], dz_[3];
  for (int i = 0; i < 3; ++i) { ds_[i] = s[i] + m * st[i]; dz_[i] = z[i] + m * zt[i]; }
  double dot_dsz = dotv(ds_, dz_, 3);
  double de1 = m * mut - 1;
  // dot(zt, H_dual, zt)
  double Hzt[3];
  for (int i = 0; i < 3; ++i) Hzt[i] = Hd[i] * zt[0] + Hd[i + 3] * zt[1] + Hd[i + 6] * zt[2];
  double de2 = dotv(zt, Hzt, 3) - 3 * mut * mut;
  if (std::fabs(de1) > std::sqrt(kEps) && std::fabs(de2) > kEps && dot_sz > 0 && dot_dsz > 0) {
    double tmp[3];
    for (int i = 0; i < 3; ++i) tmp[i] = mut * st[i] - Hd[i] * zt[0] - Hd[i + 3] * zt[1] - Hd[i + 6] * zt[2];
    double Hs[9];
    for (int i = 0; i < 9; ++i) Hs[i] = Hd[i];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Hs[i + 3 * j] -= st[i] * st[j] / 3 + tmp[i] * tmp[j] / de2;
    double fro = 0;
    for (int i = 0; i < 9; ++i) fro += Hs[i] * Hs[i];
    double t = m * std::sqrt(fro);
    double ax[3] = {z[1] * zt[2] - z[2] * zt[1], z[2] * zt[0] - z[0] * zt[2], z[0] * zt[1] - z[1] * zt[0]};
    double an = norm2(ax, 3);
    for (int i = 0; i < 3; ++i) ax[i] /= an;
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j)
        Hs[i + 3 * j] = s[i] * s[j] / dot_sz + ds_[i] * ds_[j] / dot_dsz + t * ax[i] * ax[j];
    Hs[1] = Hs[3]; Hs[2] = Hs[6]; Hs[5] = Hs[7];
    for (int i = 0; i < 9; ++i) K.Hs3[i] = Hs[i];
  } else {
    use_dual(m);
  }
}

// backtrack_search (coneops_nonsymmetric_common.jl:5-34)
template <class F>
double backtrack_search(const double* dq, const double* q, double a_init, double a_min, double step, F infn) {
  double a = a_init;
  double wq[3];
  while (true) {
    for (int i = 0; i < 3; ++i) wq[i] = q[i] + a * dq[i];
    if (infn(wq)) break;
    if ((a *= step) < a_min) { a = 0; break; }
  }
  return a;
}

// ---------------------------------------------------------------------------
// composite cone
// ---------------------------------------------------------------------------
struct Cones {
  std::vector<Cone> c;
  IVec rng;      // start offset of each cone in (s,z)
  IVec blk;      // start offset of each cone's Hs block
  int numel = 0, degree = 0, hslen = 0;
  bool symmetric = true;

  void init(int nc, const int* type, const int* dim, const double* param) {
    c.resize(nc);
    rng.resize(nc);
    blk.resize(nc);
    for (int k = 0; k < nc; ++k) {
      Cone& K = c[k];
      K.type = type[k];
      if (K.type == ORC_EXP || K.type == ORC_POW) K.dim = 3;
      else if (K.type == ORC_PSD) { K.n = dim[k]; K.dim = tri(K.n); }
      else K.dim = dim[k];
      K.alpha = param ? param[k] : 0;
      if (K.type == ORC_NONNEG) { K.w.assign(K.dim, 0); K.lam.assign(K.dim, 0); }
      if (K.type == ORC_SOC) {
        K.w.assign(K.dim, 0); K.lam.assign(K.dim, 0);
        K.sparse = K.dim > 5;  // SOC_NO_EXPANSION_MAX_SIZE (src/Clarabel.jl:9, cones/cone_types.jl:106)
        if (K.sparse) { K.su.assign(K.dim, 0); K.sv.assign(K.dim, 0); }
      }
      if (K.type == ORC_PSD) {
        int n = K.n;
        K.R.assign(n * n, 0); K.Rinv.assign(n * n, 0); K.Hsp.assign(K.dim * K.dim, 0);
        K.lamp.assign(n, 0); K.lisq.assign(n, 0);
      }
      rng[k] = numel;
      blk[k] = hslen;
      numel += K.dim;
      hslen += K.hs_len();
      degree += K.degree();
      symmetric = symmetric && K.symmetric();
    }
  }
};

// ---- composite operations (src/cones/coneops_compositecone.jl)

// margins (:49-63) and per-family margins
void margins(Cones& C, const Vec& z, double& amin, double& bsum) {
  amin = kInf;  // typemax(T)
  bsum = 0;
  for (size_t k = 0; k < C.c.size(); ++k) {
    Cone& K = C.c[k];
    const double* zi = z.data() + C.rng[k];
    double a = 0, b = 0;
    switch (K.type) {
      case ORC_ZERO: a = kFloatMax; b = 0; break;
      case ORC_NONNEG:
        if (K.dim > 0) { a = zi[0]; for (int i = 1; i < K.dim; ++i) a = std::min(a, zi[i]); }
        else a = kFloatMax;
        for (int i = 0; i < K.dim; ++i) b += zi[i] > 0 ? zi[i] : 0;
        break;
      case ORC_SOC: a = zi[0] - norm2(zi + 1, K.dim - 1); b = std::max(0.0, a); break;
      case ORC_PSD: {
        int n = K.n;
        if (K.dim == 0) { a = kFloatMax; b = 0; break; }
        std::vector<double> M(n * n), e(n);
        svec_to_mat(n, zi, M.data());
        sym_eig(n, M.data(), e.data(), nullptr);
        a = e[0];
        for (int i = 0; i < n; ++i) if (e[i] > 0) b += e[i];
        break;
      }
      default: break;  // asymmetric cones never reach margins
    }
    amin = std::min(amin, a);
    bsum += b;
  }
}

enum PD { Primal = 0, DualC = 1 };
// scaled_unit_shift! (:65-77 and per cone)
void scaled_unit_shift(Cones& C, Vec& z, double a, PD pd) {
  for (size_t k = 0; k < C.c.size(); ++k) {
    Cone& K = C.c[k];
    double* zi = z.data() + C.rng[k];
    switch (K.type) {
      case ORC_ZERO: if (pd == Primal) for (int i = 0; i < K.dim; ++i) zi[i] = 0; break;
      case ORC_NONNEG: for (int i = 0; i < K.dim; ++i) zi[i] += a; break;
      case ORC_SOC: zi[0] += a; break;
      case ORC_PSD: for (int j = 0; j < K.n; ++j) zi[tri(j + 1) - 1] += a; break;
      default: break;
    }
  }
}

// unit_initialization! (:80-90)
void unit_initialization(Cones& C, Vec& z, Vec& s) {
  for (size_t k = 0; k < C.c.size(); ++k) {
    Cone& K = C.c[k];
    double* zi = z.data() + C.rng[k];
    double* si = s.data() + C.rng[k];
    switch (K.type) {
      case ORC_ZERO: for (int i = 0; i < K.dim; ++i) zi[i] = si[i] = 0; break;
      case ORC_NONNEG: for (int i = 0; i < K.dim; ++i) zi[i] = si[i] = 1; break;
      case ORC_SOC:
        for (int i = 0; i < K.dim; ++i) zi[i] = si[i] = 0;
        si[0] = 1; zi[0] = 1; break;
      case ORC_EXP:  // coneops_expcone.jl:36-52
        si[0] = -1.051383945322714; si[1] = 0.556409619469370; si[2] = 1.258967884768947;
        for (int i = 0; i < 3; ++i) zi[i] = si[i];
        break;
      case ORC_POW:  // coneops_powcone.jl:36-63
        si[0] = std::sqrt(1 + K.alpha); si[1] = std::sqrt(1 + (1 - K.alpha)); si[2] = 0;
        for (int i = 0; i < 3; ++i) zi[i] = si[i];
        break;
      case ORC_PSD:
        for (int i = 0; i < K.dim; ++i) zi[i] = si[i] = 0;
        for (int j = 0; j < K.n; ++j) { si[tri(j + 1) - 1] += 1; zi[tri(j + 1) - 1] += 1; }
        break;
    }
  }
}

// set_identity_scaling! (:92-101)
void set_identity_scaling(Cones& C) {
  for (auto& K : C.c) {
    switch (K.type) {
      case ORC_NONNEG: std::fill(K.w.begin(), K.w.end(), 1.0); break;
      case ORC_SOC:
        std::fill(K.w.begin(), K.w.end(), 0.0);
        K.w[0] = 1; K.eta = 1;
        if (K.sparse) {
          K.sd = 0.5;
          std::fill(K.su.begin(), K.su.end(), 0.0); K.su[0] = std::sqrt(0.5);
          std::fill(K.sv.begin(), K.sv.end(), 0.0);
        }
        break;
      case ORC_PSD: {
        int n = K.n, t = K.dim;
        std::fill(K.R.begin(), K.R.end(), 0.0); std::fill(K.Rinv.begin(), K.Rinv.end(), 0.0);
        for (int i = 0; i < n; ++i) { K.R[i + i * n] = 1; K.Rinv[i + i * n] = 1; }
        std::fill(K.Hsp.begin(), K.Hsp.end(), 0.0);
        for (int i = 0; i < t; ++i) K.Hsp[i + i * t] = 1;
        break;
      }
      default: break;
    }
  }
}

// PSD update_scaling (src/cones/coneops_psdtrianglecone.jl:79-142)
bool psd_update_scaling(Cone& K, const double* s, const double* z) {
  int n = K.n;
  if (K.dim == 0) return true;
  std::vector<double> S(n * n), Z(n * n), L1(n * n), L2(n * n), tmp(n * n), U(n * n), V(n * n), sv(n);
  svec_to_mat(n, s, S.data());
  svec_to_mat(n, z, Z.data());
  if (!chol(n, S.data(), L1.data()) || !chol(n, Z.data(), L2.data())) return false;
  gemm(n, true, false, 1.0, L2.data(), L1.data(), 0.0, tmp.data());  // L2' L1
  svd_jacobi(n, tmp.data(), U.data(), sv.data(), V.data());
  for (int i = 0; i < n; ++i) { K.lamp[i] = sv[i]; K.lisq[i] = 1 / std::sqrt(sv[i]); }
  // R = L1 V Λ^{-1/2}
  gemm(n, false, false, 1.0, L1.data(), V.data(), 0.0, K.R.data());
  for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) K.R[i + j * n] *= K.lisq[j];
  // Rinv = Λ^{-1/2} U' L2'
  gemm(n, true, true, 1.0, U.data(), L2.data(), 0.0, K.Rinv.data());
  for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) K.Rinv[i + j * n] *= K.lisq[i];
  // Hs = skron(R R')
  gemm(n, false, true, 1.0, K.R.data(), K.R.data(), 0.0, tmp.data());
  // symmetrize exactly as syrk + Symmetric(upper) would
  for (int j = 0; j < n; ++j) for (int i = j + 1; i < n; ++i) tmp[i + j * n] = tmp[j + i * n];
  skron(n, tmp.data(), K.Hsp.data());
  return true;
}

// PSD W products (coneops_psdtrianglecone.jl:413-443)
void psd_mul_W_inner(const Cone& K, bool transpose, double* y, const double* x, const Vec& Rx) {
  int n = K.n;
  std::vector<double> X(n * n), Y(n * n), tmp(n * n);
  svec_to_mat(n, x, X.data());
  if (transpose) {
    gemm(n, false, true, 1.0, X.data(), Rx.data(), 0.0, tmp.data());  // X R'
    gemm(n, false, false, 1.0, Rx.data(), tmp.data(), 0.0, Y.data());  // R X R'
  } else {
    gemm(n, true, false, 1.0, Rx.data(), X.data(), 0.0, tmp.data());   // R' X
    gemm(n, false, false, 1.0, tmp.data(), Rx.data(), 0.0, Y.data());  // R' X R
  }
  mat_to_svec(n, Y.data(), y);
}

bool update_scaling(Cones& C, const Vec& s, const Vec& z, double mu, Scaling strat) {
  for (size_t k = 0; k < C.c.size(); ++k) {
    Cone& K = C.c[k];
    const double* si = s.data() + C.rng[k];
    const double* zi = z.data() + C.rng[k];
    bool ok = true;
    switch (K.type) {
      case ORC_ZERO: break;
      case ORC_NONNEG:  // coneops_nncone.jl:77-89
        for (int i = 0; i < K.dim; ++i) { K.lam[i] = std::sqrt(si[i] * zi[i]); K.w[i] = std::sqrt(si[i] / zi[i]); }
        break;
      case ORC_SOC: ok = soc_update_scaling(K, si, zi); break;
      case ORC_EXP: case ORC_POW:  // coneops_expcone.jl:63-83, coneops_powcone.jl:65-85
        if (K.type == ORC_EXP) exp_dual_grad_H(zi, K.grad, K.Hd);
        else pow_dual_grad_H(zi, K.alpha, K.grad, K.Hd);
        nonsym_update_Hs(K, si, zi, mu, strat);
        for (int i = 0; i < 3; ++i) K.zc[i] = zi[i];
        break;
      case ORC_PSD: ok = psd_update_scaling(K, si, zi); break;
    }
    if (!ok) return false;
  }
  return true;
}

// get_Hs! (:123-132 and per cone) into the packed block vector
void get_Hs(Cones& C, Vec& Hs) {
  for (size_t k = 0; k < C.c.size(); ++k) {
    Cone& K = C.c[k];
    double* h = Hs.data() + C.blk[k];
    switch (K.type) {
      case ORC_ZERO: for (int i = 0; i < K.dim; ++i) h[i] = 0; break;
      case ORC_NONNEG: for (int i = 0; i < K.dim; ++i) h[i] = K.w[i] * K.w[i]; break;
      case ORC_SOC: {  // coneops_socone.jl:156-193
        double eta2 = K.eta * K.eta;
        if (K.sparse) {
          for (int i = 0; i < K.dim; ++i) h[i] = eta2;
          h[0] *= K.sd;
        } else {
          const double s2 = std::sqrt(2.0);
          h[0] = (s2 * K.w[0] - 1) * (s2 * K.w[0] + 1);
          int hidx = 1;
          for (int col = 1; col < K.dim; ++col) {
            double wc = K.w[col];
            for (int row = 0; row <= col; ++row) { h[hidx] = 2 * K.w[row] * wc; ++hidx; }
            h[hidx - 1] += 1;
          }
          for (int i = 0; i < tri(K.dim); ++i) h[i] *= eta2;
        }
        break;
      }
      case ORC_EXP: case ORC_POW: {  // pack_triu of the 3x3 Hs
        int idx = 0;
        for (int col = 0; col < 3; ++col) for (int row = 0; row <= col; ++row) h[idx++] = K.Hs3[row + 3 * col];
        break;
      }
      case ORC_PSD: {
        int t = K.dim, idx = 0;
        for (int col = 0; col < t; ++col) for (int row = 0; row <= col; ++row) h[idx++] = K.Hsp[row + t * col];
        break;
      }
    }
  }
}

// mul_Hs! (:138-150)
void mul_Hs(Cones& C, Vec& y, const Vec& x) {
  for (size_t k = 0; k < C.c.size(); ++k) {
    Cone& K = C.c[k];
    double* yi = y.data() + C.rng[k];
    const double* xi = x.data() + C.rng[k];
    switch (K.type) {
      case ORC_ZERO: for (int i = 0; i < K.dim; ++i) yi[i] = 0; break;
      case ORC_NONNEG: for (int i = 0; i < K.dim; ++i) yi[i] = K.w[i] * (K.w[i] * xi[i]); break;
      case ORC_SOC: {  // coneops_socone.jl:202-217
        double c = 2 * dotv(K.w.data(), xi, K.dim);
        double e2 = K.eta * K.eta;
        for (int i = 0; i < K.dim; ++i) yi[i] = xi[i];
        yi[0] = -xi[0];
        for (int i = 0; i < K.dim; ++i) yi[i] += c * K.w[i];
        for (int i = 0; i < K.dim; ++i) yi[i] *= e2;
        break;
      }
      case ORC_EXP: case ORC_POW: {
        double t[3];
        for (int i = 0; i < 3; ++i) t[i] = K.Hs3[i] * xi[0] + K.Hs3[i + 3] * xi[1] + K.Hs3[i + 6] * xi[2];
        for (int i = 0; i < 3; ++i) yi[i] = t[i];
        break;
      }
      case ORC_PSD: {
        std::vector<double> w(K.dim);
        psd_mul_W_inner(K, false, w.data(), xi, K.R);
        psd_mul_W_inner(K, true, yi, w.data(), K.R);
        break;
      }
    }
  }
}

// affine_ds! (:153-165)
void affine_ds(Cones& C, Vec& ds, const Vec& s) {
  for (size_t k = 0; k < C.c.size(); ++k) {
    Cone& K = C.c[k];
    double* d = ds.data() + C.rng[k];
    const double* si = s.data() + C.rng[k];
    switch (K.type) {
      case ORC_ZERO: for (int i = 0; i < K.dim; ++i) d[i] = 0; break;
      case ORC_NONNEG: for (int i = 0; i < K.dim; ++i) d[i] = K.lam[i] * K.lam[i]; break;
      case ORC_SOC: soc_circ(d, K.lam.data(), K.lam.data(), K.dim); break;
      case ORC_EXP: case ORC_POW: for (int i = 0; i < 3; ++i) d[i] = si[i]; break;
      case ORC_PSD:
        for (int i = 0; i < K.dim; ++i) d[i] = 0;
        for (int j = 0; j < K.n; ++j) d[tri(j + 1) - 1] = K.lamp[j] * K.lamp[j];
        break;
    }
  }
}

// combined_ds_shift! (:167-184); modifies step_z, step_s of symmetric cones in place
// (src/cones/coneops_symmetric_common.jl:2-35)
void combined_ds_shift(Cones& C, Vec& shift, Vec& step_z, Vec& step_s, double sigmu) {
  for (size_t k = 0; k < C.c.size(); ++k) {
    Cone& K = C.c[k];
    int o = C.rng[k], d = K.dim;
    double* sh = shift.data() + o;
    double* sz = step_z.data() + o;
    double* ss = step_s.data() + o;
    switch (K.type) {
      case ORC_ZERO: for (int i = 0; i < d; ++i) sh[i] = 0; break;
      case ORC_NONNEG:
        for (int i = 0; i < d; ++i) { sz[i] = sz[i] * K.w[i]; ss[i] = ss[i] / K.w[i]; }
        for (int i = 0; i < d; ++i) sh[i] = ss[i] * sz[i];
        for (int i = 0; i < d; ++i) sh[i] += -sigmu;
        break;
      case ORC_SOC: {
        std::vector<double> tmp(sz, sz + d);
        soc_mul_W(K, sz, tmp.data());
        tmp.assign(ss, ss + d);
        soc_mul_Winv(K, ss, tmp.data());
        soc_circ(sh, ss, sz, d);
        sh[0] += -sigmu;
        break;
      }
      case ORC_EXP: case ORC_POW: {  // coneops_expcone.jl:130-148, coneops_powcone.jl:132-150
        double eta[3];
        if (K.type == ORC_EXP) exp_higher_correction(K, eta, ss, sz);
        else pow_higher_correction(K, eta, ss, sz);
        for (int i = 0; i < 3; ++i) sh[i] = K.grad[i] * sigmu - eta[i];
        break;
      }
      case ORC_PSD: {
        int n = K.n;
        std::vector<double> tmp(sz, sz + d);
        psd_mul_W_inner(K, false, sz, tmp.data(), K.R);
        tmp.assign(ss, ss + d);
        psd_mul_W_inner(K, true, ss, tmp.data(), K.Rinv);
        std::vector<double> Y(n * n), Z(n * n), X(n * n);
        svec_to_mat(n, ss, Y.data());
        svec_to_mat(n, sz, Z.data());
        // X = (YZ + ZY)/2
        gemm(n, false, false, 0.5, Y.data(), Z.data(), 0.0, X.data());
        gemm(n, false, false, 0.5, Z.data(), Y.data(), 1.0, X.data());
        // Symmetric(X) uses the upper triangle
        for (int j = 0; j < n; ++j) for (int i = j + 1; i < n; ++i) X[i + j * n] = X[j + i * n];
        mat_to_svec(n, X.data(), sh);
        for (int j = 0; j < n; ++j) sh[tri(j + 1) - 1] += -sigmu;
        break;
      }
    }
  }
}

// Δs_from_Δz_offset! (:186-203)
void ds_from_dz_offset(Cones& C, Vec& out, const Vec& ds, Vec& work, const Vec& z) {
  for (size_t k = 0; k < C.c.size(); ++k) {
    Cone& K = C.c[k];
    int o = C.rng[k], d = K.dim;
    double* oi = out.data() + o;
    const double* dsi = ds.data() + o;
    const double* zi = z.data() + o;
    double* wi = work.data() + o;
    switch (K.type) {
      case ORC_ZERO: for (int i = 0; i < d; ++i) oi[i] = 0; break;
      case ORC_NONNEG: for (int i = 0; i < d; ++i) oi[i] = dsi[i] / zi[i]; break;
      case ORC_SOC: {  // coneops_socone.jl:242-269
        double resz = soc_residual(zi, d);
        double l1ds1 = dotv(K.lam.data() + 1, dsi + 1, d - 1);
        double w1ds1 = dotv(K.w.data() + 1, dsi + 1, d - 1);
        for (int i = 0; i < d; ++i) oi[i] = -zi[i];
        oi[0] = zi[0];
        double c = K.lam[0] * dsi[0] - l1ds1;
        for (int i = 0; i < d; ++i) oi[i] *= c / resz;
        oi[0] += K.eta * w1ds1;
        for (int i = 1; i < d; ++i) oi[i] += K.eta * (dsi[i] + w1ds1 / (1 + K.w[0]) * K.w[i]);
        for (int i = 0; i < d; ++i) oi[i] *= (1 / K.lam[0]);
        break;
      }
      case ORC_EXP: case ORC_POW: for (int i = 0; i < 3; ++i) oi[i] = dsi[i]; break;
      case ORC_PSD: {
        int n = K.n;
        std::vector<double> X(n * n), Z(n * n);
        svec_to_mat(n, dsi, Z.data());
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j) X[i + j * n] = 2 * Z[i + j * n] / (K.lamp[i] + K.lamp[j]);
        mat_to_svec(n, X.data(), wi);
        psd_mul_W_inner(K, true, oi, wi, K.R);
        break;
      }
    }
  }
}

// PSD step length component (coneops_psdtrianglecone.jl:445-473)
double psd_step_component(const Cone& K, const double* dsvec, double amax) {
  int n = K.n;
  if (K.dim == 0) return amax;
  std::vector<double> M(n * n), e(n);
  svec_to_mat(n, dsvec, M.data());
  for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) M[i + j * n] *= K.lisq[i] * K.lisq[j];
  sym_eig(n, M.data(), e.data(), nullptr);
  double g = e[0];
  if (g < 0) return std::min(1 / (-g), amax);
  return amax;
}

// step_length (:206-243)
double step_length(Cones& C, const Vec& dz, const Vec& ds, const Vec& z, const Vec& s, const orc_settings& st,
                   double amax) {
  double a = amax;
  auto inner = [&](double al, bool symcond) {
    for (size_t k = 0; k < C.c.size(); ++k) {
      Cone& K = C.c[k];
      if (K.symmetric() == symcond) continue;
      int o = C.rng[k], d = K.dim;
      const double *dzi = dz.data() + o, *dsi = ds.data() + o, *zi = z.data() + o, *si = s.data() + o;
      double az = al, as = al;
      switch (K.type) {
        case ORC_ZERO: break;
        case ORC_NONNEG:  // coneops_nncone.jl:151-170
          for (int i = 0; i < d; ++i) {
            if (dzi[i] < 0) az = std::min(az, -zi[i] / dzi[i]);
            if (dsi[i] < 0) as = std::min(as, -si[i] / dsi[i]);
          }
          break;
        case ORC_SOC:
          az = soc_step_component(zi, dzi, al, d);
          as = soc_step_component(si, dsi, al, d);
          break;
        case ORC_EXP: {
          double step = st.linesearch_backtrack_step, amin = st.min_terminate_step_length;
          az = backtrack_search(dzi, zi, al, amin, step, exp_dual_feasible);
          as = backtrack_search(dsi, si, al, amin, step, exp_primal_feasible);
          break;
        }
        case ORC_POW: {
          double step = st.linesearch_backtrack_step, amin = st.min_terminate_step_length;
          double pa = K.alpha;
          az = backtrack_search(dzi, zi, al, amin, step, [&](const double* v) { return pow_dual_feasible(v, pa); });
          as = backtrack_search(dsi, si, al, amin, step, [&](const double* v) { return pow_primal_feasible(v, pa); });
          break;
        }
        case ORC_PSD: {
          std::vector<double> dv(d);
          psd_mul_W_inner(K, false, dv.data(), dzi, K.R);
          az = psd_step_component(K, dv.data(), al);
          psd_mul_W_inner(K, true, dv.data(), dsi, K.Rinv);
          as = psd_step_component(K, dv.data(), al);
          break;
        }
      }
      al = std::min(al, std::min(az, as));
    }
    return al;
  };
  a = inner(a, false);
  if (!C.symmetric) a = std::min(a, st.max_step_fraction);
  a = inner(a, true);
  return a;
}

// compute_barrier (:246-265)
double compute_barrier(Cones& C, const Vec& z, const Vec& s, const Vec& dz, const Vec& ds, double a) {
  double bar = 0;
  for (size_t k = 0; k < C.c.size(); ++k) {
    Cone& K = C.c[k];
    int o = C.rng[k], d = K.dim;
    const double *zi = z.data() + o, *si = s.data() + o, *dzi = dz.data() + o, *dsi = ds.data() + o;
    switch (K.type) {
      case ORC_ZERO: break;
      case ORC_NONNEG:
        for (int i = 0; i < d; ++i) bar -= logsafe((si[i] + a * dsi[i]) * (zi[i] + a * dzi[i]));
        break;
      case ORC_SOC: {
        double rs = soc_residual_shifted(si, dsi, a, d), rz = soc_residual_shifted(zi, dzi, a, d);
        if (rs > 0 && rz > 0) bar += -logsafe(rs * rz) / 2;
        else bar += kInf;
        break;
      }
      case ORC_EXP: case ORC_POW: {
        double cz[3], cs[3];
        for (int i = 0; i < 3; ++i) { cz[i] = zi[i] + a * dzi[i]; cs[i] = si[i] + a * dsi[i]; }
        if (K.type == ORC_EXP) bar += exp_barrier_dual(cz) + exp_barrier_primal(cs);
        else bar += pow_barrier_dual(cz, K.alpha) + pow_barrier_primal(cs, K.alpha);
        break;
      }
      case ORC_PSD: {
        int n = K.n;
        auto logdet = [&](const double* x, const double* dx) {
          std::vector<double> q(d), Q(n * n), L(n * n);
          for (int i = 0; i < d; ++i) q[i] = x[i] + a * dx[i];
          svec_to_mat(n, q.data(), Q.data());
          if (!chol(n, Q.data(), L.data())) return kInf;  // typemax(T)
          double ld = 0;
          for (int i = 0; i < n; ++i) ld += 2 * std::log(L[i + i * n]);
          return ld;
        };
        bar -= logdet(zi, dzi);
        bar -= logdet(si, dsi);
        break;
      }
    }
  }
  return bar;
}

// rectify_equilibration! (:28-47; coneops_defaults.jl:32-44; nn/zero overrides)
bool rectify_equilibration(Cones& C, Vec& delta, const Vec& e) {
  bool any = false;
  std::fill(delta.begin(), delta.end(), 1.0);
  for (size_t k = 0; k < C.c.size(); ++k) {
    Cone& K = C.c[k];
    int o = C.rng[k], d = K.dim;
    if (K.type == ORC_ZERO || K.type == ORC_NONNEG) continue;
    double mean = 0;
    for (int i = 0; i < d; ++i) mean += e[o + i];
    mean /= d;
    for (int i = 0; i < d; ++i) delta[o + i] = mean / e[o + i];
    any = true;
  }
  return any;
}

// ---------------------------------------------------------------------------
// minimum-degree ordering on an explicit elimination graph.  The reference
// uses AMD.jl through QDLDL (src/kktsolvers/direct-ldl/directldl_qdldl.jl:18-25);
// the permutation only changes rounding, not the factorized operator.
// ---------------------------------------------------------------------------
IVec min_degree_order(int n, const IVec& Kp, const IVec& Ki) {
  std::vector<IVec> adj(n);
  for (int j = 0; j < n; ++j)
    for (int k = Kp[j]; k < Kp[j + 1]; ++k) {
      int i = Ki[k];
      if (i != j) { adj[i].push_back(j); adj[j].push_back(i); }
    }
  for (auto& a : adj) { std::sort(a.begin(), a.end()); a.erase(std::unique(a.begin(), a.end()), a.end()); }
  std::vector<char> done(n, 0);
  IVec order;
  order.reserve(n);
  std::set<std::pair<int, int>> pq;  // (degree, index): smallest degree, then smallest index
  for (int i = 0; i < n; ++i) pq.insert({(int)adj[i].size(), i});
  IVec merged, nb, cleaned;
  while (!pq.empty()) {
    int v = pq.begin()->second;
    pq.erase(pq.begin());
    done[v] = 1;
    order.push_back(v);
    nb.clear();
    for (int u : adj[v]) if (!done[u]) nb.push_back(u);
    for (int u : nb) {
      merged.clear();
      std::set_union(adj[u].begin(), adj[u].end(), nb.begin(), nb.end(), std::back_inserter(merged));
      cleaned.clear();
      for (int x : merged) if (x != u && !done[x]) cleaned.push_back(x);
      pq.erase({(int)adj[u].size(), u});
      adj[u].swap(cleaned);
      pq.insert({(int)adj[u].size(), u});
    }
    IVec().swap(adj[v]);
  }
  return order;
}

// ---------------------------------------------------------------------------
// QDLDL-style up-looking LDL' with sign-based dynamic regularization
// (restates the QDLDL.jl 0.4 algorithm reached from
//  src/kktsolvers/direct-ldl/directldl_qdldl.jl:18-25, :62-86)
// ---------------------------------------------------------------------------
struct Ldl {
  int n = 0;
  IVec perm, iperm;        // perm[k] = original index of pivot k
  IVec Ap, Ai;             // permuted triu pattern
  IVec Amap;               // permuted nz -> original KKT nz
  Vec Ax;
  IVec etree, Lnz, Lp, Li;
  Vec Lx, D, Dinv;
  IVec dsigns;             // permuted signs
  double eps = 1e-13, delta = 2e-7;
  bool reg = true;
  bool keep = false;  // sign * max(delta, |d|) instead of sign * delta (not the reference's rule)
  double flops = 0;

  void setup(int n_, const IVec& Kp, const IVec& Ki, const IVec& dsigns_orig) {
    n = n_;
    perm = min_degree_order(n, Kp, Ki);
    iperm.assign(n, 0);
    for (int k = 0; k < n; ++k) iperm[perm[k]] = k;
    // permuted triu: entry (i,j) -> (pi,pj), store in column max(pi,pj)
    IVec cnt(n + 1, 0);
    for (int j = 0; j < n; ++j)
      for (int k = Kp[j]; k < Kp[j + 1]; ++k) {
        int pi = iperm[Ki[k]], pj = iperm[j];
        cnt[std::max(pi, pj) + 1]++;
      }
    Ap.assign(n + 1, 0);
    for (int j = 0; j < n; ++j) Ap[j + 1] = Ap[j] + cnt[j + 1];
    int nz = Ap[n];
    Ai.assign(nz, 0);
    Amap.assign(nz, 0);
    Ax.assign(nz, 0);
    IVec next(Ap.begin(), Ap.end() - 1);
    for (int j = 0; j < n; ++j)
      for (int k = Kp[j]; k < Kp[j + 1]; ++k) {
        int pi = iperm[Ki[k]], pj = iperm[j];
        int c = std::max(pi, pj), r = std::min(pi, pj);
        Ai[next[c]] = r;
        Amap[next[c]] = k;
        next[c]++;
      }
    dsigns.assign(n, 1);
    for (int k = 0; k < n; ++k) dsigns[k] = dsigns_orig[perm[k]];
    // elimination tree and column counts (QDLDL_etree)
    etree.assign(n, -1);
    Lnz.assign(n, 0);
    IVec work(n, -1);
    for (int j = 0; j < n; ++j) {
      work[j] = j;
      for (int p = Ap[j]; p < Ap[j + 1]; ++p) {
        int i = Ai[p];
        while (work[i] != j) {
          if (etree[i] == -1) etree[i] = j;
          Lnz[i]++;
          work[i] = j;
          i = etree[i];
        }
      }
    }
    Lp.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) Lp[i + 1] = Lp[i] + Lnz[i];
    Li.assign(Lp[n], 0);
    Lx.assign(Lp[n], 0);
    D.assign(n, 0);
    Dinv.assign(n, 0);
    flops = 0;
    for (int i = 0; i < n; ++i) flops += (double)Lnz[i] * (Lnz[i] + 1);
  }

  // numeric factorization from KKT values (already diag-shifted)
  bool factor(const Vec& Kvals) {
    for (size_t p = 0; p < Ax.size(); ++p) Ax[p] = Kvals[Amap[p]];
    IVec yMarker(n, 0), yIdx(n), elimBuffer(n), LNext(n);
    Vec yVals(n, 0.0);
    for (int i = 0; i < n; ++i) LNext[i] = Lp[i];
    for (int k = 0; k < n; ++k) {
      int nnzY = 0;
      D[k] = 0;
      for (int p = Ap[k]; p < Ap[k + 1]; ++p) {
        int bidx = Ai[p];
        if (bidx == k) { D[k] = Ax[p]; continue; }
        yVals[bidx] = Ax[p];
        int nextIdx = bidx;
        if (yMarker[nextIdx] == 0) {
          yMarker[nextIdx] = 1;
          elimBuffer[0] = nextIdx;
          int nnzE = 1;
          nextIdx = etree[bidx];
          while (nextIdx != -1 && nextIdx < k) {
            if (yMarker[nextIdx] == 1) break;
            yMarker[nextIdx] = 1;
            elimBuffer[nnzE++] = nextIdx;
            nextIdx = etree[nextIdx];
          }
          while (nnzE) yIdx[nnzY++] = elimBuffer[--nnzE];
        }
      }
      for (int i = nnzY - 1; i >= 0; --i) {
        int cidx = yIdx[i];
        int tmpIdx = LNext[cidx];
        double yc = yVals[cidx];
        for (int j = Lp[cidx]; j < tmpIdx; ++j) yVals[Li[j]] -= Lx[j] * yc;
        Li[tmpIdx] = k;
        Lx[tmpIdx] = yc * Dinv[cidx];
        D[k] -= yc * Lx[tmpIdx];
        LNext[cidx]++;
        yVals[cidx] = 0;
        yMarker[cidx] = 0;
      }
      if (reg) {
        int sg = dsigns[k];
        if (D[k] * sg <= eps) D[k] = (keep && std::fabs(D[k]) > delta ? std::fabs(D[k]) : delta) * sg;
      }
      if (D[k] == 0) return false;
      Dinv[k] = 1 / D[k];
    }
    for (int k = 0; k < n; ++k) if (!std::isfinite(Dinv[k])) return false;
    return true;
  }

  void solve(const double* b, double* x) const {
    Vec y(n);
    for (int k = 0; k < n; ++k) y[k] = b[perm[k]];
    for (int i = 0; i < n; ++i)
      for (int j = Lp[i]; j < Lp[i + 1]; ++j) y[Li[j]] -= Lx[j] * y[i];
    for (int i = 0; i < n; ++i) y[i] *= Dinv[i];
    for (int i = n - 1; i >= 0; --i)
      for (int j = Lp[i]; j < Lp[i + 1]; ++j) y[i] -= Lx[j] * y[Li[j]];
    for (int k = 0; k < n; ++k) x[perm[k]] = y[k];
  }
};

// ---------------------------------------------------------------------------
// KKT assembly (src/kktsolvers/direct-ldl/directldl_kkt_assembly.jl:15-175)
// triu CSC of [P A'; A -Hs] plus sparse SOC expansion columns
// ---------------------------------------------------------------------------
struct KktMap {
  IVec P, A, Hsblocks, diag_full, diagP;
  struct SocMap { IVec u, v; int D[2]; };
  std::vector<SocMap> soc;  // one per sparse SOC, in cone order
};

void assemble_kkt(const Csc& P, const Csc& A, Cones& C, Csc& K, KktMap& map, int& p) {
  int n = P.n, m = A.m;
  p = 0;
  for (auto& c : C.c) if (c.type == ORC_SOC && c.sparse) p += 2;
  int N = n + m + p;
  // per-column entry lists of (row, source-tag, source-index)
  struct Ent { int row; int tag; int idx; };
  enum { T_P, T_PDIAG, T_A, T_HS, T_SV, T_SU, T_SD };
  std::vector<std::vector<Ent>> cols(N);
  for (int j = 0; j < n; ++j) {
    bool hasdiag = false;
    for (int k = P.p[j]; k < P.p[j + 1]; ++k) {
      cols[j].push_back({P.i[k], T_P, k});
      if (P.i[k] == j) hasdiag = true;
    }
    if (!hasdiag) cols[j].push_back({j, T_PDIAG, j});
  }
  // A' : row i of A goes to column n+i with rows = A's columns
  for (int j = 0; j < A.n; ++j)
    for (int k = A.p[j]; k < A.p[j + 1]; ++k) cols[n + A.i[k]].push_back({j, T_A, k});
  int pcol = n + m;
  map.soc.clear();
  for (size_t ci = 0; ci < C.c.size(); ++ci) {
    Cone& K = C.c[ci];
    int row0 = n + C.rng[ci];
    int b0 = C.blk[ci];
    if (K.hs_diagonal()) {
      for (int i = 0; i < K.dim; ++i) cols[row0 + i].push_back({row0 + i, T_HS, b0 + i});
    } else {
      int kidx = 0;
      for (int col = 0; col < K.dim; ++col)
        for (int row = 0; row <= col; ++row) cols[row0 + col].push_back({row0 + row, T_HS, b0 + kidx++});
    }
    if (K.type == ORC_SOC && K.sparse) {
      int si = (int)map.soc.size();
      map.soc.push_back({});
      map.soc[si].u.assign(K.dim, 0);
      map.soc[si].v.assign(K.dim, 0);
      for (int i = 0; i < K.dim; ++i) cols[pcol].push_back({row0 + i, T_SV, i + (si << 20)});
      for (int i = 0; i < K.dim; ++i) cols[pcol + 1].push_back({row0 + i, T_SU, i + (si << 20)});
      cols[pcol].push_back({pcol, T_SD, 0 + (si << 20)});
      cols[pcol + 1].push_back({pcol + 1, T_SD, 1 + (si << 20)});
      pcol += 2;
    }
  }
  K.m = K.n = N;
  K.p.assign(N + 1, 0);
  for (int j = 0; j < N; ++j) K.p[j + 1] = K.p[j] + (int)cols[j].size();
  K.i.assign(K.p[N], 0);
  K.x.assign(K.p[N], 0.0);
  map.P.assign(P.nnz(), -1);
  map.A.assign(A.nnz(), -1);
  map.Hsblocks.assign(C.hslen, -1);
  map.diag_full.assign(N, -1);
  map.diagP.assign(n, -1);
  for (int j = 0; j < N; ++j) {
    auto& c = cols[j];
    std::stable_sort(c.begin(), c.end(), [](const Ent& a, const Ent& b) { return a.row < b.row; });
    for (size_t t = 0; t < c.size(); ++t) {
      int dst = K.p[j] + (int)t;
      const Ent& e = c[t];
      K.i[dst] = e.row;
      switch (e.tag) {
        case T_P: map.P[e.idx] = dst; K.x[dst] = P.x[e.idx]; break;
        case T_PDIAG: break;
        case T_A: map.A[e.idx] = dst; K.x[dst] = A.x[e.idx]; break;
        case T_HS: map.Hsblocks[e.idx] = dst; break;
        case T_SV: map.soc[e.idx >> 20].v[e.idx & 0xFFFFF] = dst; break;
        case T_SU: map.soc[e.idx >> 20].u[e.idx & 0xFFFFF] = dst; break;
        case T_SD: map.soc[e.idx >> 20].D[e.idx & 0xFFFFF] = dst; break;
      }
      if (e.row == j) map.diag_full[j] = dst;
    }
  }
  for (int j = 0; j < n; ++j) map.diagP[j] = map.diag_full[j];
}

// ---------------------------------------------------------------------------
// solver state
// ---------------------------------------------------------------------------
struct Variables {
  Vec x, s, z;
  double tau = 1, kappa = 1;
  void resize(int n, int m) { x.assign(n, 0); s.assign(m, 0); z.assign(m, 0); }
};
struct Residuals {
  Vec rx, rz, rx_inf, rz_inf, Px;
  double rtau = 0, dot_qx = 0, dot_bz = 0, dot_sz = 0, dot_xPx = 0;
};
struct Info {
  double mu = 0, sigma = 0, step_length = 0;
  int iterations = 0;
  double cost_primal = 0, cost_dual = 0, res_primal = 0, res_dual = 0, res_primal_inf = 0, res_dual_inf = 0;
  double gap_abs = 0, gap_rel = 0, ktratio = 0;
  double prev_cost_primal = kInf, prev_cost_dual = kInf, prev_res_primal = kInf, prev_res_dual = kInf;
  double prev_gap_abs = kInf, prev_gap_rel = kInf;
  double solve_time = 0, setup_time = 0;
  int status = ORC_UNSOLVED;
};

struct Solver {
  orc_settings st;
  int n = 0, m = 0, p = 0, N = 0;
  Csc P, A;  // scaled data, P triu
  Vec q, b;
  // equilibration
  Vec d, dinv, e, einv;
  double c = 1;
  double normq = 0, normb = 0;
  bool normb_valid = true, normq_valid = true;
  Cones cones;
  // KKT
  Csc K;
  KktMap map;
  IVec Dsigns;
  Vec Hsblocks;
  Ldl ldl;
  Vec kx, kb, work1, work2;
  // KKT system
  Vec x1, z1, x2, z2, workx, workz, work_conic;
  Variables vars, step_rhs, step_lhs, prev_vars;
  Residuals res;
  Info info;
  std::vector<double> trace;  // 8 per iteration record

  // --------------------------------------------------------------- setup
  void setup(int n_, int m_, const int* Pp, const int* Pi, const double* Px, const double* q_, const int* Ap,
             const int* Ai, const double* Ax, const double* b_, int nc, const int* ct, const int* cd,
             const double* cp, const orc_settings& s) {
    st = s;
    n = n_;
    m = m_;
    // triu(P) copy (src/problemdata.jl:19-23)
    P.m = P.n = n;
    P.p.assign(n + 1, 0);
    for (int j = 0; j < n; ++j) {
      for (int k = Pp[j]; k < Pp[j + 1]; ++k)
        if (Pi[k] <= j) { P.i.push_back(Pi[k]); P.x.push_back(Px[k]); }
      P.p[j + 1] = (int)P.i.size();
    }
    A.m = m; A.n = n;
    A.p.assign(Ap, Ap + n + 1);
    A.i.assign(Ai, Ai + Ap[n]);
    A.x.assign(Ax, Ax + Ap[n]);
    q.assign(q_, q_ + n);
    b.assign(b_, b_ + m);
    for (auto& v : b) v = std::min(v, 1e20);  // src/problemdata.jl:47-51
    normq = norm_inf(q.data(), n);
    normb = norm_inf(b.data(), m);
    cones.init(nc, ct, cd, cp);
    d.assign(n, 1); dinv.assign(n, 1); e.assign(m, 1); einv.assign(m, 1); c = 1;
    equilibrate();
    // KKT system (src/kktsystem.jl:23-55, kktsolver_directldl.jl:47-100)
    assemble_kkt(P, A, cones, K, map, p);
    N = n + m + p;
    Dsigns.assign(N, 1);
    for (int i = n; i < n + m; ++i) Dsigns[i] = -1;
    {
      int pc = n + m;
      for (auto& cc : cones.c)
        if (cc.type == ORC_SOC && cc.sparse) { Dsigns[pc] = -1; Dsigns[pc + 1] = 1; pc += 2; }
    }
    Hsblocks.assign(cones.hslen, 0);
    ldl.reg = st.dynamic_regularization_enable != 0;
    ldl.eps = st.dynamic_regularization_eps;
    ldl.delta = st.dynamic_regularization_delta;
    ldl.keep = st.dynamic_regularization_keep_magnitude != 0;
    ldl.setup(N, K.p, K.i, Dsigns);
    kx.assign(N, 0); kb.assign(N, 0); work1.assign(N, 0); work2.assign(N, 0);
    x1.assign(n, 0); z1.assign(m, 0); x2.assign(n, 0); z2.assign(m, 0);
    workx.assign(n, 0); workz.assign(m, 0); work_conic.assign(m, 0);
    vars.resize(n, m); step_rhs.resize(n, m); step_lhs.resize(n, m); prev_vars.resize(n, m);
    res.rx.assign(n, 0); res.rz.assign(m, 0); res.rx_inf.assign(n, 0); res.rz_inf.assign(m, 0); res.Px.assign(n, 0);
  }

  // ------------------------------------------------ equilibration
  // data_equilibrate! (src/problemdata.jl:196-284)
  void scale_data(const Vec* dw, const Vec& ew) {
    if (dw) {
      for (int j = 0; j < n; ++j)
        for (int k = P.p[j]; k < P.p[j + 1]; ++k) P.x[k] *= (*dw)[P.i[k]] * (*dw)[j];
      for (int j = 0; j < n; ++j)
        for (int k = A.p[j]; k < A.p[j + 1]; ++k) A.x[k] *= ew[A.i[k]] * (*dw)[j];
      for (int j = 0; j < n; ++j) q[j] *= (*dw)[j];
    } else {
      for (int k = 0; k < A.nnz(); ++k) A.x[k] *= ew[A.i[k]];
    }
    for (int i = 0; i < m; ++i) b[i] *= ew[i];
  }
  void equilibrate() {
    if (!st.equilibrate_enable) return;
    Vec dw(n), ew(m);
    double smin = st.equilibrate_min_scaling, smax = st.equilibrate_max_scaling;
    for (int it = 0; it < st.equilibrate_max_iter; ++it) {
      // kkt_col_norms! (src/utils/mathutils.jl:86-98)
      std::fill(dw.begin(), dw.end(), 0.0);
      for (int j = 0; j < n; ++j)
        for (int k = P.p[j]; k < P.p[j + 1]; ++k) {
          double t = std::fabs(P.x[k]);
          int r = P.i[k];
          dw[j] = dw[j] > t ? dw[j] : t;
          dw[r] = dw[r] > t ? dw[r] : t;
        }
      for (int j = 0; j < n; ++j)
        for (int k = A.p[j]; k < A.p[j + 1]; ++k) { double t = std::fabs(A.x[k]); dw[j] = dw[j] > t ? dw[j] : t; }
      std::fill(ew.begin(), ew.end(), 0.0);
      for (int k = 0; k < A.nnz(); ++k) { int r = A.i[k]; double t = std::fabs(A.x[k]); ew[r] = ew[r] > t ? ew[r] : t; }
      for (auto& v : dw) if (v == 0) v = 1;
      for (auto& v : ew) if (v == 0) v = 1;
      for (auto& v : dw) v = 1 / std::sqrt(v);
      for (auto& v : ew) v = 1 / std::sqrt(v);
      for (int j = 0; j < n; ++j) dw[j] = clip(dw[j], smin / d[j], smax / d[j]);
      for (int i = 0; i < m; ++i) ew[i] = clip(ew[i], smin / e[i], smax / e[i]);
      scale_data(&dw, ew);
      for (int j = 0; j < n; ++j) d[j] *= dw[j];
      for (int i = 0; i < m; ++i) e[i] *= ew[i];
      // col_norms!(dwork, P) on the triu storage (mathutils.jl:100-122)
      std::fill(dw.begin(), dw.end(), 0.0);
      for (int j = 0; j < n; ++j)
        for (int k = P.p[j]; k < P.p[j + 1]; ++k) { double t = std::fabs(P.x[k]); dw[j] = dw[j] > t ? dw[j] : t; }
      double mean = 0;
      for (double v : dw) mean += v;
      mean /= n;
      double inf_q = norm_inf(q.data(), n);
      if (mean != 0 && inf_q != 0) {
        double scale_cost = std::max(inf_q, mean);
        double ctmp = 1 / scale_cost;
        ctmp = clip(ctmp, smin / c, smax / c);
        for (auto& v : P.x) v *= ctmp;
        for (auto& v : q) v *= ctmp;
        c *= ctmp;
      }
    }
    if (rectify_equilibration(cones, ew, e)) {
      scale_data(nullptr, ew);
      for (int i = 0; i < m; ++i) e[i] *= ew[i];
    }
    for (int j = 0; j < n; ++j) dinv[j] = 1 / d[j];
    for (int i = 0; i < m; ++i) einv[i] = 1 / e[i];
  }

  // ------------------------------------------------ KKT solver
  // _kktsolver_update_inner! (kktsolver_directldl.jl:213-247)
  bool kktsolver_update() {
    get_Hs(cones, Hsblocks);
    for (size_t t = 0; t < Hsblocks.size(); ++t) K.x[map.Hsblocks[t]] = -Hsblocks[t];
    size_t si = 0;
    for (auto& cc : cones.c) {
      if (!(cc.type == ORC_SOC && cc.sparse)) continue;
      // _csc_update_sparsecone (directldl_datamaps.jl:61-79)
      auto& sm = map.soc[si++];
      double eta2 = cc.eta * cc.eta;
      for (int i = 0; i < cc.dim; ++i) { K.x[sm.u[i]] = cc.su[i] * -eta2; K.x[sm.v[i]] = cc.sv[i] * -eta2; }
      K.x[sm.D[0]] = -eta2;
      K.x[sm.D[1]] = +eta2;
    }
    return regularize_and_refactor();
  }
  // kktsolver_directldl.jl:249-296
  bool regularize_and_refactor() {
    Vec diag_kkt(N);
    for (int i = 0; i < N; ++i) diag_kkt[i] = K.x[map.diag_full[i]];
    if (st.static_regularization_enable) {
      double maxdiag = norm_inf(diag_kkt.data(), N);
      double eps = st.static_regularization_constant + st.static_regularization_proportional * maxdiag;
      for (int i = 0; i < N; ++i) K.x[map.diag_full[i]] = diag_kkt[i] + (Dsigns[i] == 1 ? eps : -eps);
    }
    bool ok = ldl.factor(K.x);
    if (st.static_regularization_enable)
      for (int i = 0; i < N; ++i) K.x[map.diag_full[i]] = diag_kkt[i];
    return ok;
  }
  void setrhs(const Vec& rx, const Vec& rz) {
    for (int i = 0; i < n; ++i) kb[i] = rx[i];
    for (int i = 0; i < m; ++i) kb[n + i] = rz[i];
    for (int i = n + m; i < N; ++i) kb[i] = 0;
  }
  // e = b - K xi, returns ||e||_inf (kktsolver_directldl.jl:457-468)
  double refine_error(Vec& ev, const Vec& bv, const Vec& xi) {
    for (int i = 0; i < N; ++i) ev[i] = bv[i];
    symv_triu(K, xi.data(), ev.data(), -1.0, 1.0);
    return norm_inf(ev.data(), N);
  }
  // kktsolver_solve! + _iterative_refinement (kktsolver_directldl.jl:348-451)
  bool kktsolver_solve(Vec* lx, Vec* lz) {
    ldl.solve(kb.data(), kx.data());
    bool ok = true;
    if (st.iterative_refinement_enable) {
      Vec& ev = work1;
      Vec& dx = work2;
      double normb_ = norm_inf(kb.data(), N);
      double norme = refine_error(ev, kb, kx);
      if (!std::isfinite(norme)) ok = false;
      else {
        for (int it = 0; it < st.iterative_refinement_max_iter; ++it) {
          if (norme <= st.iterative_refinement_abstol + st.iterative_refinement_reltol * normb_) break;
          double lastnorme = norme;
          ldl.solve(ev.data(), dx.data());
          for (int i = 0; i < N; ++i) dx[i] += kx[i];
          norme = refine_error(ev, kb, dx);
          if (!std::isfinite(norme)) { ok = false; break; }
          double ratio = lastnorme / norme;
          if (ratio < st.iterative_refinement_stop_ratio) {
            if (ratio > 1) kx.swap(dx);
            break;
          }
          kx.swap(dx);
        }
      }
    } else {
      for (int i = 0; i < N; ++i) if (!std::isfinite(kx[i])) ok = false;
    }
    if (ok) {
      if (lx) for (int i = 0; i < n; ++i) (*lx)[i] = kx[i];
      if (lz) for (int i = 0; i < m; ++i) (*lz)[i] = kx[n + i];
    }
    return ok;
  }

  // ------------------------------------------------ KKT system (src/kktsystem.jl)
  bool kkt_update() {
    if (!kktsolver_update()) return false;
    for (int i = 0; i < n; ++i) workx[i] = -q[i];
    setrhs(workx, b);
    return kktsolver_solve(&x2, &z2);
  }
  bool kkt_solve_initial_point() {
    bool ok;
    if (P.nnz() == 0) {
      std::fill(workx.begin(), workx.end(), 0.0);
      workz = b;
      setrhs(workx, workz);
      ok = kktsolver_solve(&vars.x, &vars.s);
      for (auto& v : vars.s) v = -v;
      if (!ok) return ok;
      for (int i = 0; i < n; ++i) workx[i] = -q[i];
      std::fill(workz.begin(), workz.end(), 0.0);
      setrhs(workx, workz);
      ok = kktsolver_solve(nullptr, &vars.z);
    } else {
      for (int i = 0; i < n; ++i) workx[i] = -q[i];
      workz = b;
      setrhs(workx, workz);
      ok = kktsolver_solve(&vars.x, &vars.z);
      for (int i = 0; i < m; ++i) vars.s[i] = -vars.z[i];
    }
    return ok;
  }
  bool kkt_solve(Variables& lhs, Variables& rhs, bool affine) {
    workx = rhs.x;
    Vec& dsc = work_conic;
    if (affine) dsc = vars.s;
    else ds_from_dz_offset(cones, dsc, rhs.s, lhs.z, vars.z);
    for (int i = 0; i < m; ++i) workz[i] = dsc[i] - rhs.z[i];
    setrhs(workx, workz);
    if (!kktsolver_solve(&x1, &z1)) return false;
    Vec& xi = workx;
    for (int i = 0; i < n; ++i) xi[i] = vars.x[i] / vars.tau;
    double tau_num = rhs.tau - rhs.kappa / vars.tau + dotv(q.data(), x1.data(), n) + dotv(b.data(), z1.data(), m) +
                     2 * quad_form(xi, P, x1);
    for (int i = 0; i < n; ++i) xi[i] -= x2[i];
    double tau_den = vars.kappa / vars.tau - dotv(q.data(), x2.data(), n) - dotv(b.data(), z2.data(), m);
    tau_den += quad_form(xi, P, xi) - quad_form(x2, P, x2);
    lhs.tau = tau_num / tau_den;
    for (int i = 0; i < n; ++i) lhs.x[i] = x1[i] + lhs.tau * x2[i];
    for (int i = 0; i < m; ++i) lhs.z[i] = z1[i] + lhs.tau * z2[i];
    mul_Hs(cones, lhs.s, lhs.z);
    for (int i = 0; i < m; ++i) lhs.s[i] = -(lhs.s[i] + dsc[i]);
    lhs.kappa = -(rhs.kappa + vars.kappa * lhs.tau) / vars.tau;
    return true;
  }

  // ------------------------------------------------ residuals / info
  void residuals_update() {  // src/residuals.jl:1-37
    double qx = dotv(q.data(), vars.x.data(), n);
    double bz = dotv(b.data(), vars.z.data(), m);
    double sz = dotv(vars.s.data(), vars.z.data(), m);
    std::fill(res.Px.begin(), res.Px.end(), 0.0);
    symv_triu(P, vars.x.data(), res.Px.data(), 1.0, 0.0);
    double xPx = dotv(vars.x.data(), res.Px.data(), n);
    gemv_csc_t(A, vars.z.data(), res.rx_inf.data(), -1.0, 0.0);
    for (int i = 0; i < m; ++i) res.rz_inf[i] = vars.s[i];
    gemv_csc(A, vars.x.data(), res.rz_inf.data(), 1.0, 1.0);
    for (int i = 0; i < n; ++i) res.rx[i] = res.rx_inf[i] - res.Px[i] - q[i] * vars.tau;
    for (int i = 0; i < m; ++i) res.rz[i] = res.rz_inf[i] - b[i] * vars.tau;
    res.rtau = qx + bz + vars.kappa + xPx / vars.tau;
    res.dot_qx = qx; res.dot_bz = bz; res.dot_sz = sz; res.dot_xPx = xPx;
  }
  double calc_mu() { return (res.dot_sz + vars.tau * vars.kappa) / (cones.degree + 1); }
  void info_update() {  // src/info.jl:1-64
    double tinv = 1 / vars.tau;
    if (!normb_valid) { normb = norm_inf_scaled(einv, b); normb_valid = true; }
    if (!normq_valid) {
      Vec t(n);
      double cinv = 1 / c;
      normq = norm_inf_scaled(dinv, q) * cinv;
      normq_valid = true;
    }
    double cinv = 1 / c;
    double xPx2 = res.dot_xPx * tinv * tinv / 2;
    info.cost_primal = (+res.dot_qx * tinv + xPx2) * cinv;
    info.cost_dual = (-res.dot_bz * tinv - xPx2) * cinv;
    double normx = norm_scaled(d, vars.x);
    double normz = norm_scaled(e, vars.z) * cinv;
    double norms = norm_scaled(einv, vars.s);
    info.res_primal_inf = (norm_scaled(dinv, res.rx_inf) * cinv) / std::max(1.0, normz);
    info.res_dual_inf = std::max(norm_scaled(dinv, res.Px) / std::max(1.0, normx),
                                 norm_scaled(einv, res.rz_inf) / std::max(1.0, normx + norms));
    normx *= tinv; normz *= tinv; norms *= tinv;
    info.res_primal = norm_scaled(einv, res.rz) * tinv / std::max(1.0, normb + normx + norms);
    info.res_dual = norm_scaled(dinv, res.rx) * tinv * cinv / std::max(1.0, normq + normx + normz);
    info.gap_abs = std::fabs(info.cost_primal - info.cost_dual);
    info.gap_rel = info.gap_abs / std::max(1.0, std::min(std::fabs(info.cost_primal), std::fabs(info.cost_dual)));
    info.ktratio = vars.kappa * tinv;
  }
  // _check_convergence (src/info.jl:307-408)
  void check_convergence(double tga, double tgr, double tf, double tia, double tir, double tkt, int sstat, int pstat,
                         int dstat) {
    auto is_solved = [&]() {
      return ((info.gap_abs < tga) || (info.gap_rel < tgr)) && (info.res_primal < tf) && (info.res_dual < tf);
    };
    if (info.ktratio <= 1 && is_solved()) info.status = sstat;
    else if (info.ktratio > 1000. / tkt) {
      if ((res.dot_bz < -tia) && (info.res_primal_inf < -tir * res.dot_bz)) info.status = pstat;
      else if ((res.dot_qx < -tia) && (info.res_dual_inf < -tir * res.dot_qx)) info.status = dstat;
    }
  }
  // info_check_termination! (src/info.jl:135-190)
  bool check_termination(int iter, double elapsed) {
    info.status = ORC_UNSOLVED;
    check_convergence(st.tol_gap_abs, st.tol_gap_rel, st.tol_feas, st.tol_infeas_abs, st.tol_infeas_rel,
                      st.tol_ktratio, ORC_SOLVED, ORC_PRIMAL_INFEASIBLE, ORC_DUAL_INFEASIBLE);
    if (info.status == ORC_UNSOLVED && iter > 1 &&
        (info.res_dual > info.prev_res_dual || info.res_primal > info.prev_res_primal)) {
      if (info.ktratio < 100 * kEps && (info.prev_gap_abs < st.tol_gap_abs || info.prev_gap_rel < st.tol_gap_rel))
        info.status = ORC_INSUFFICIENT_PROGRESS;
      if (info.ktratio < 1) {
        if ((info.res_dual > 100 * st.tol_feas && info.res_dual > 100 * info.prev_res_dual) ||
            (info.res_primal > 100 * st.tol_feas && info.res_primal > 100 * info.prev_res_primal))
          info.status = ORC_INSUFFICIENT_PROGRESS;
      }
    }
    if (info.status == ORC_UNSOLVED) {
      if (st.max_iter == info.iterations) info.status = ORC_MAX_ITERATIONS;
      else if (elapsed > st.time_limit) info.status = ORC_MAX_TIME;
    }
    return info.status != ORC_UNSOLVED;
  }
  void save_prev_iterate() {
    info.prev_cost_primal = info.cost_primal; info.prev_cost_dual = info.cost_dual;
    info.prev_res_primal = info.res_primal; info.prev_res_dual = info.res_dual;
    info.prev_gap_abs = info.gap_abs; info.prev_gap_rel = info.gap_rel;
    prev_vars = vars;
  }
  void reset_to_prev_iterate() {
    info.cost_primal = info.prev_cost_primal; info.cost_dual = info.prev_cost_dual;
    info.res_primal = info.prev_res_primal; info.res_dual = info.prev_res_dual;
    info.gap_abs = info.prev_gap_abs; info.gap_rel = info.prev_gap_rel;
    vars = prev_vars;
  }
  void record() {
    trace.push_back(info.cost_primal);
    trace.push_back(info.cost_dual);
    trace.push_back(std::min(info.gap_abs, info.gap_rel));
    trace.push_back(info.res_primal);
    trace.push_back(info.res_dual);
    trace.push_back(info.ktratio);
    trace.push_back(info.mu);
    trace.push_back(info.iterations > 0 ? info.step_length : std::nan(""));
  }

  // ------------------------------------------------ step length
  double get_step_length(bool combined, Scaling scaling) {  // src/solver.jl:449-464, variables.jl:14-43
    double at = step_lhs.tau < 0 ? -vars.tau / step_lhs.tau : kFloatMax;
    double ak = step_lhs.kappa < 0 ? -vars.kappa / step_lhs.kappa : kFloatMax;
    double a = std::min(std::min(at, ak), 1.0);
    a = step_length(cones, step_lhs.z, step_lhs.s, vars.z, vars.s, st, a);
    if (combined) a *= st.max_step_fraction;
    if (!cones.symmetric && combined && scaling == Dual) {
      // solver_backtrack_step_to_barrier (src/solver.jl:468-485)
      double stp = st.linesearch_backtrack_step;
      for (int j = 0; j < 50; ++j) {
        double bar = barrier(a);
        if (bar < 1) return a;
        a = stp * a;
      }
    }
    return a;
  }
  double barrier(double a) {  // variables_barrier (src/variables.jl:46-72)
    double cc = cones.degree + 1;
    double ct = vars.tau + a * step_lhs.tau, ck = vars.kappa + a * step_lhs.kappa;
    double sz = dot_shifted(vars.z.data(), vars.s.data(), step_lhs.z.data(), step_lhs.s.data(), a, m);
    double mu = (sz + ct * ck) / cc;
    double bar = cc * logsafe(mu) - logsafe(ct) - logsafe(ck);
    bar += compute_barrier(cones, vars.z, vars.s, step_lhs.z, step_lhs.s, a);
    return bar;
  }

  // ------------------------------------------------ the IPM loop (src/solver.jl:231-422)
  void shift_to_cone_interior(Vec& z, PD pd) {  // src/variables.jl:276-304
    double mm, pm;
    margins(cones, z, mm, pm);
    double target = std::max(1.0, 0.1 * pm / cones.degree);
    if (mm <= 0) {
      scaled_unit_shift(cones, z, -mm, pd);
      scaled_unit_shift(cones, z, target, pd);
    } else if (mm < target) {
      scaled_unit_shift(cones, z, target - mm, pd);
    } else {
      scaled_unit_shift(cones, z, 0.0, pd);
    }
  }
  void default_start() {  // src/solver.jl:425-446
    if (cones.symmetric) {
      set_identity_scaling(cones);
      kkt_update();
      kkt_solve_initial_point();
      shift_to_cone_interior(vars.s, Primal);
      shift_to_cone_interior(vars.z, DualC);
      vars.tau = 1; vars.kappa = 1;
    } else {
      unit_initialization(cones, vars.z, vars.s);
      std::fill(vars.x.begin(), vars.x.end(), 0.0);
      vars.tau = 1; vars.kappa = 1;
    }
  }

  void solve() {
    auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&]() {
      return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() + info.setup_time;
    };
    int iter = 0;
    double sigma = 1, alpha = 0, mu = kInf;
    double setup_time = info.setup_time;
    info = Info();
    info.setup_time = setup_time;
    trace.clear();
    default_start();
    Scaling scaling = PrimalDual;  // allows_primal_dual_scaling is true for every supported cone
    while (true) {
      residuals_update();
      mu = calc_mu();
      info.mu = mu; info.step_length = alpha; info.sigma = sigma; info.iterations = iter;
      info_update();
      bool isdone = check_termination(iter, elapsed());
      record();
      if (isdone) {
        // _strategy_checkpoint_insufficient_progress (src/solver.jl:521-540)
        if (info.status != ORC_INSUFFICIENT_PROGRESS) break;
        reset_to_prev_iterate();
        if (!cones.symmetric && scaling == PrimalDual) {
          info.status = ORC_UNSOLVED;
          scaling = Dual;
          continue;
        }
        break;
      }
      bool sc_ok = update_scaling(cones, vars.s, vars.z, mu, scaling);
      if (!sc_ok) { info.status = ORC_NUMERICAL_ERROR; break; }
      ++iter;
      bool ok = kkt_update();
      // affine step rhs (src/variables.jl:161-175)
      step_rhs.x = res.rx;
      step_rhs.z = res.rz;
      affine_ds(cones, step_rhs.s, vars.s);
      step_rhs.tau = res.rtau;
      step_rhs.kappa = vars.tau * vars.kappa;
      ok = ok && kkt_solve(step_lhs, step_rhs, true);
      if (ok) {
        alpha = get_step_length(false, scaling);
        sigma = (1 - alpha) * (1 - alpha) * (1 - alpha);  // (1-α)^3, a literal power (src/solver.jl:514-517)
        double mm = iter > 1 ? 1.0 : alpha;
        // combined step rhs (src/variables.jl:178-216)
        double sigmu = sigma * mu;
        for (int i = 0; i < n; ++i) step_rhs.x[i] = (1 - sigma) * res.rx[i];
        step_rhs.tau = (1 - sigma) * res.rtau;
        step_rhs.kappa = -sigmu + mm * step_lhs.tau * step_lhs.kappa + vars.tau * vars.kappa;
        if (mm != 1) for (auto& v : step_lhs.z) v *= mm;
        combined_ds_shift(cones, step_rhs.z, step_lhs.z, step_lhs.s, sigmu);
        for (int i = 0; i < m; ++i) step_rhs.s[i] += step_rhs.z[i];
        for (int i = 0; i < m; ++i) step_rhs.z[i] = (1 - sigma) * res.rz[i];
        ok = kkt_solve(step_lhs, step_rhs, false);
      }
      // _strategy_checkpoint_numerical_error (src/solver.jl:543-557)
      if (!ok) {
        if (!cones.symmetric && scaling == PrimalDual) { scaling = Dual; alpha = 0; continue; }
        info.status = ORC_NUMERICAL_ERROR;
        alpha = 0;
        break;
      }
      alpha = get_step_length(true, scaling);
      // _strategy_checkpoint_small_step (src/solver.jl:560-573)
      if (!cones.symmetric && scaling == PrimalDual && alpha < st.min_switch_step_length) {
        scaling = Dual; alpha = 0; continue;
      } else if (alpha <= std::max(0.0, st.min_terminate_step_length)) {
        info.status = ORC_INSUFFICIENT_PROGRESS;
        alpha = 0;
        break;
      }
      save_prev_iterate();
      // variables_add_step! (src/variables.jl:146-158)
      for (int i = 0; i < n; ++i) vars.x[i] += alpha * step_lhs.x[i];
      for (int i = 0; i < m; ++i) vars.s[i] += alpha * step_lhs.s[i];
      for (int i = 0; i < m; ++i) vars.z[i] += alpha * step_lhs.z[i];
      vars.tau += alpha * step_lhs.tau;
      vars.kappa += alpha * step_lhs.kappa;
    }
    if (alpha == 0) {
      info.mu = mu; info.step_length = alpha; info.sigma = sigma; info.iterations = iter;
    }
    // info_post_process! (src/info.jl:275-291)
    if (info.status == ORC_NUMERICAL_ERROR || info.status == ORC_INSUFFICIENT_PROGRESS ||
        info.status == ORC_MAX_ITERATIONS || info.status == ORC_MAX_TIME) {
      check_convergence(st.reduced_tol_gap_abs, st.reduced_tol_gap_rel, st.reduced_tol_feas,
                        st.reduced_tol_infeas_abs, st.reduced_tol_infeas_rel, st.reduced_tol_ktratio,
                        ORC_ALMOST_SOLVED, ORC_ALMOST_PRIMAL_INFEASIBLE, ORC_ALMOST_DUAL_INFEASIBLE);
    }
    info.solve_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // solution_post_process! (src/solution.jl:2-50) + variables_unscale! (src/variables.jl:343-371)
  void get_solution(double* x, double* s, double* z, orc_result* r) {
    int stt = info.status;
    bool infeas = stt == ORC_PRIMAL_INFEASIBLE || stt == ORC_DUAL_INFEASIBLE || stt == ORC_ALMOST_PRIMAL_INFEASIBLE ||
                  stt == ORC_ALMOST_DUAL_INFEASIBLE;
    r->status = stt;
    r->obj_val = infeas ? std::nan("") : info.cost_primal;
    r->obj_val_dual = infeas ? std::nan("") : info.cost_dual;
    r->iterations = info.iterations;
    r->r_prim = info.res_primal;
    r->r_dual = info.res_dual;
    r->solve_time = info.solve_time;
    r->setup_time = info.setup_time;
    double scaleinv = infeas ? 1 / vars.kappa : 1 / vars.tau;
    double cinv = 1 / c;
    for (int i = 0; i < n; ++i) x[i] = vars.x[i] * d[i] * scaleinv;
    for (int i = 0; i < m; ++i) z[i] = vars.z[i] * e[i] * (scaleinv * cinv);
    for (int i = 0; i < m; ++i) s[i] = vars.s[i] * einv[i] * scaleinv;
    r->trace_len = (int)(trace.size() / 8);
    r->nnz_L = ldl.Lp.empty() ? 0 : ldl.Lp[N];
    r->flops_factor = ldl.flops;
  }

  // update_P! / update_b! (src/data_updating.jl:71-104, :188-202)
  void update_P(const double* Px_new_triu) {
    for (int j = 0; j < n; ++j)
      for (int k = P.p[j]; k < P.p[j + 1]; ++k) {
        double v = Px_new_triu[k] * d[P.i[k]] * d[j];
        if (!st.update_P_without_cscale) v *= c;
        P.x[k] = v;
        K.x[map.P[k]] = v;
      }
  }
  void update_b(const double* bnew) {
    for (int i = 0; i < m; ++i) b[i] = bnew[i] * e[i];
    normb_valid = false;
  }
  // update_q! (src/data_updating.jl:164-179): q = d .* q_new * c
  void update_q(const double* qnew) {
    for (int j = 0; j < n; ++j) q[j] = qnew[j] * d[j] * c;
    normq_valid = false;
  }
  // update_A! (src/data_updating.jl:120-134, _update_matrix :234-247 with
  // kktsolver_update_A! kktsolver_directldl.jl:383-388): A = E A_new D,
  // written through to the KKT values
  void update_A(const double* Axnew) {
    for (int j = 0; j < n; ++j)
      for (int k = A.p[j]; k < A.p[j + 1]; ++k) {
        A.x[k] = e[A.i[k]] * Axnew[k] * d[j];
        K.x[map.A[k]] = A.x[k];
      }
  }
};

}  // namespace

extern "C" {

void orc_default_settings(orc_settings* s) {
  s->max_iter = 200;
  s->time_limit = kInf;
  s->verbose = 0;
  s->max_step_fraction = 0.99;
  s->tol_gap_abs = 1e-8; s->tol_gap_rel = 1e-8; s->tol_feas = 1e-8;
  s->tol_infeas_abs = 1e-8; s->tol_infeas_rel = 1e-8; s->tol_ktratio = 1e-6;
  s->reduced_tol_gap_abs = 5e-5; s->reduced_tol_gap_rel = 5e-5; s->reduced_tol_feas = 1e-4;
  s->reduced_tol_infeas_abs = 5e-12; s->reduced_tol_infeas_rel = 5e-5; s->reduced_tol_ktratio = 1e-4;
  s->equilibrate_enable = 1; s->equilibrate_max_iter = 10;
  s->equilibrate_min_scaling = 1e-4; s->equilibrate_max_scaling = 1e4;
  s->linesearch_backtrack_step = 0.8; s->min_switch_step_length = 1e-2; s->min_terminate_step_length = 1e-4;
  s->static_regularization_enable = 1; s->static_regularization_constant = 1e-8;
  s->static_regularization_proportional = kEps * kEps;
  s->dynamic_regularization_enable = 1; s->dynamic_regularization_eps = 1e-13; s->dynamic_regularization_delta = 2e-7;
  s->iterative_refinement_enable = 1; s->iterative_refinement_reltol = 1e-12; s->iterative_refinement_abstol = 1e-12;
  s->iterative_refinement_max_iter = 10; s->iterative_refinement_stop_ratio = 5;
  s->update_P_without_cscale = 0;
  s->dynamic_regularization_keep_magnitude = 0;
}

static void copy_trace(const Solver& S, double* trace) {
  if (!trace) return;
  std::memcpy(trace, S.trace.data(), S.trace.size() * sizeof(double));
}

int orc_solve(int32_t n, int32_t m, const int32_t* Pp, const int32_t* Pi, const double* Px, const double* q,
              const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* b, int32_t ncones,
              const int32_t* cone_type, const int32_t* cone_dim, const double* cone_param, const orc_settings* settings,
              double* x, double* s, double* z, orc_result* res, double* trace) {
  orc_settings st;
  if (settings) st = *settings; else orc_default_settings(&st);
  Solver S;
  auto t0 = std::chrono::steady_clock::now();
  S.setup(n, m, Pp, Pi, Px, q, Ap, Ai, Ax, b, ncones, cone_type, cone_dim, cone_param, st);
  S.info.setup_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  S.solve();
  S.get_solution(x, s, z, res);
  copy_trace(S, trace);
  return 0;
}

int orc_solve_update(int32_t n, int32_t m, const int32_t* Pp, const int32_t* Pi, const double* Px, const double* q,
                     const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* b, int32_t ncones,
                     const int32_t* cone_type, const int32_t* cone_dim, const double* cone_param,
                     const orc_settings* settings, const double* Px_new, const double* b_new, const double* q_new,
                     const double* Ax_new, double* x, double* s, double* z, orc_result* res, double* trace) {
  orc_settings st;
  if (settings) st = *settings; else orc_default_settings(&st);
  Solver S;
  S.setup(n, m, Pp, Pi, Px, q, Ap, Ai, Ax, b, ncones, cone_type, cone_dim, cone_param, st);
  S.solve();
  if (b_new) S.update_b(b_new);
  if (q_new) S.update_q(q_new);
  if (Ax_new) S.update_A(Ax_new);
  if (Px_new) {
    // Px_new holds values for the triu entries of the original pattern, in order
    std::vector<double> vals;
    for (int j = 0; j < n; ++j)
      for (int k = Pp[j]; k < Pp[j + 1]; ++k)
        if (Pi[k] <= j) vals.push_back(Px_new[k]);
    S.update_P(vals.data());
  }
  S.solve();
  S.get_solution(x, s, z, res);
  copy_trace(S, trace);
  return 0;
}

double orc_wright_omega(double zz) { return wright_omega(zz); }
void orc_exp_gradient_primal(const double* s, double* g) { exp_gradient_primal(s, g); }
void orc_exp_dual_grad_H(const double* z, double* grad, double* H) { exp_dual_grad_H(z, grad, H); }
void orc_pow_gradient_primal(const double* s, double alpha, double* g) { pow_gradient_primal(s, alpha, g); }
void orc_pow_dual_grad_H(const double* z, double alpha, double* grad, double* H) { pow_dual_grad_H(z, alpha, grad, H); }

int orc_soc_update_scaling(int32_t dim, const double* s, const double* z, double* w, double* lambda, double* eta,
                           double* d, double* u, double* v) {
  Cone K;
  K.type = ORC_SOC; K.dim = dim;
  K.w.assign(dim, 0); K.lam.assign(dim, 0);
  K.sparse = true; K.su.assign(dim, 0); K.sv.assign(dim, 0);
  bool ok = soc_update_scaling(K, s, z);
  std::copy(K.w.begin(), K.w.end(), w);
  std::copy(K.lam.begin(), K.lam.end(), lambda);
  *eta = K.eta; *d = K.sd;
  std::copy(K.su.begin(), K.su.end(), u);
  std::copy(K.sv.begin(), K.sv.end(), v);
  return ok ? 1 : 0;
}
double orc_soc_step_length_component(int32_t dim, const double* x, const double* y, double amax) {
  return soc_step_component(x, y, amax, dim);
}
int orc_psd_update_scaling(int32_t n, const double* s, const double* z, double* R, double* Rinv, double* lambda,
                           double* Hs) {
  Cone K;
  K.type = ORC_PSD; K.n = n; K.dim = tri(n);
  K.R.assign(n * n, 0); K.Rinv.assign(n * n, 0); K.Hsp.assign(K.dim * K.dim, 0); K.lamp.assign(n, 0);
  K.lisq.assign(n, 0);
  bool ok = psd_update_scaling(K, s, z);
  std::copy(K.R.begin(), K.R.end(), R);
  std::copy(K.Rinv.begin(), K.Rinv.end(), Rinv);
  std::copy(K.lamp.begin(), K.lamp.end(), lambda);
  std::copy(K.Hsp.begin(), K.Hsp.end(), Hs);
  return ok ? 1 : 0;
}
void orc_sym_eigvals(int32_t n, const double* M, double* evals) { sym_eig(n, M, evals, nullptr); }

// one exponential (type 3) or power (type 4) cone with its dual gradient and
// Hessian evaluated at z, as update_scaling! leaves it
static Cone nonsym_cone(int32_t type, double alpha, const double* z) {
  Cone K;
  K.type = type; K.dim = 3; K.alpha = alpha;
  if (type == ORC_EXP) exp_dual_grad_H(z, K.grad, K.Hd);
  else pow_dual_grad_H(z, alpha, K.grad, K.Hd);
  for (int i = 0; i < 3; ++i) K.zc[i] = z[i];
  return K;
}
void orc_nonsym_higher_correction(int32_t type, double alpha, const double* z, const double* ds, const double* v,
                                  double* eta) {
  Cone K = nonsym_cone(type, alpha, z);
  if (type == ORC_EXP) exp_higher_correction(K, eta, ds, v);
  else pow_higher_correction(K, eta, ds, v);
}
void orc_nonsym_update_Hs(int32_t type, double alpha, const double* s, const double* z, double mu, int32_t strategy,
                          double* Hs) {
  Cone K = nonsym_cone(type, alpha, z);
  nonsym_update_Hs(K, s, z, mu, strategy == 1 ? Dual : PrimalDual);
  std::copy(K.Hs3, K.Hs3 + 9, Hs);
}

int orc_ldl_solve(int32_t n, const int32_t* Kp, const int32_t* Ki, const double* Kx, const int32_t* dsigns,
                  double reg_eps, double reg_delta, int32_t keep_magnitude, const double* b, double* x,
                  int64_t* nnz_L) {
  Ldl L;
  IVec kp(Kp, Kp + n + 1), ki(Ki, Ki + Kp[n]);
  IVec ds(dsigns, dsigns + n);
  L.eps = reg_eps; L.delta = reg_delta; L.reg = true; L.keep = keep_magnitude != 0;
  L.setup(n, kp, ki, ds);
  Vec kx(Kx, Kx + Kp[n]);
  if (!L.factor(kx)) return 1;
  L.solve(b, x);
  if (nnz_L) *nnz_L = L.Lp[n];
  return 0;
}

}  // extern "C"
