// Structural analysis for the device engine (host side, once per shape).
#include "clh_symbolic.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <thread>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <functional>
#include <future>
#include <limits>
#include <queue>
#include <stdexcept>
#include <unordered_map>

#include "clh_types.h"

namespace clh {

namespace {

inline int tri(int k) { return k * (k + 1) / 2; }
// CLH_SETUP_TIMING=1 prints the host analysis phases' wall times to stderr
inline void setup_mark(const char* what) {
  static const bool on = std::getenv("CLH_SETUP_TIMING") != nullptr;
  static auto t0 = std::chrono::steady_clock::now();
  if (!on) return;
  const auto t = std::chrono::steady_clock::now();
  std::fprintf(stderr, "[clh setup] %-28s %8.3f s\n", what, std::chrono::duration<double>(t - t0).count());
  t0 = t;
}

// run work(j0, j1) over column ranges of a CSC pattern on up to 16 host
// threads, the ranges balanced by entries (serially below a million entries)
template <class F>
void parallel_columns(int ncols, int64_t nnz, const int32_t* colptr, F&& work) {
  const int nt = nnz < 1000000 ? 1 : (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
  if (nt == 1) { work(0, ncols); return; }
  std::vector<int> cut(nt + 1, ncols);
  cut[0] = 0;
  for (int t = 1; t < nt; ++t)
    cut[t] = (int)(std::lower_bound(colptr, colptr + ncols + 1, (int32_t)(nnz * t / nt)) - colptr);
  for (int t = 1; t <= nt; ++t) cut[t] = std::max(cut[t], cut[t - 1]);
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) th.emplace_back(work, cut[t], cut[t + 1]);
  for (auto& x : th) x.join();
}

// 64-bit key of a sparsity pattern: 8 bytes per step, each word mixed in by
// a multiply and a xor-shift (a request's pattern is hashed on every submit,
// so this runs at memory speed rather than a byte per multiply)
struct Hash {
  uint64_t h = 1469598103934665603ull;
  void word(uint64_t w) {
    h = (h ^ w) * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
  }
  void add(const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
      uint64_t w;
      std::memcpy(&w, c + i, 8);
      word(w);
    }
    uint64_t t = n;  // the length, and the last bytes
    for (; i < n; ++i) t = (t << 8) | c[i];
    word(t);
  }
  template <class T> void v(T x) { add(&x, sizeof(T)); }
};

// Minimum-degree ordering on the explicit elimination graph, ties broken by
// the smaller index; rows whose initial degree exceeds max(16, 10 sqrt(N))
// are ordered last (the dense-row rule of AMD).
//
// Nodes flagged in `wait` (KKT pivots that are the bare regularization: the
// equality rows, whose -Hs block is zero, and variables without a P
// diagonal) become eligible only once a neighbour has been eliminated, so
// that their pivot has received a contribution -a^2/d from an eliminated
// pivot d first: an FP32 factorization that pivots on a bare eps grows like
// 1/eps and loses every digit.  When only waiting nodes remain they are all
// released.
std::vector<int32_t> min_degree(int N, const std::vector<int32_t>& Kp, const std::vector<int32_t>& Ki,
                                const int8_t* wait = nullptr) {
  std::vector<std::vector<int32_t>> adj(N);
  for (int j = 0; j < N; ++j)
    for (int k = Kp[j]; k < Kp[j + 1]; ++k) {
      int i = Ki[k];
      if (i != j) { adj[i].push_back(j); adj[j].push_back(i); }
    }
  for (auto& a : adj) { std::sort(a.begin(), a.end()); a.erase(std::unique(a.begin(), a.end()), a.end()); }
  const int dense_thr = std::max(16, (int)(10 * std::sqrt((double)N)));
  std::vector<char> dense(N, 0), gone(N, 0);
  std::vector<int32_t> order;
  order.reserve(N);
  for (int i = 0; i < N; ++i)
    if ((int)adj[i].size() > dense_thr) dense[i] = 1;
  // remove dense nodes from the graph
  for (int i = 0; i < N; ++i)
    if (!dense[i]) {
      auto& a = adj[i];
      a.erase(std::remove_if(a.begin(), a.end(), [&](int32_t u) { return dense[u] != 0; }), a.end());
    }
  using E = std::pair<int, int>;  // (degree, node)
  std::priority_queue<E, std::vector<E>, std::greater<E>> heap;
  std::vector<int> deg(N);
  std::vector<char> parked(N, 0);
  int nparked = 0;
  for (int i = 0; i < N; ++i)
    if (!dense[i]) {
      deg[i] = (int)adj[i].size();
      if (wait && wait[i] && !adj[i].empty()) { parked[i] = 1; ++nparked; }
      else heap.push({deg[i], i});
    }
  std::vector<int32_t> nb, merged;
  while (!heap.empty() || nparked > 0) {
    if (heap.empty()) {
      for (int i = 0; i < N; ++i)
        if (parked[i]) { parked[i] = 0; heap.push({deg[i], i}); }
      nparked = 0;
      continue;
    }
    E top = heap.top();
    heap.pop();
    int v = top.second;
    if (gone[v] || parked[v] || top.first != deg[v]) continue;
    gone[v] = 1;
    order.push_back(v);
    nb.clear();
    for (int u : adj[v]) if (!gone[u]) nb.push_back(u);
    for (int u : nb) {
      merged.clear();
      merged.reserve(adj[u].size() + nb.size());
      auto a = adj[u].begin(), ae = adj[u].end();
      auto b = nb.begin(), be = nb.end();
      while (a != ae || b != be) {
        int x;
        if (b == be || (a != ae && *a < *b)) x = *a++;
        else if (a == ae || *b < *a) x = *b++;
        else { x = *a++; ++b; }
        if (x != u && !gone[x]) merged.push_back(x);
      }
      adj[u].swap(merged);
      deg[u] = (int)adj[u].size();
      if (parked[u]) { parked[u] = 0; --nparked; }
      heap.push({deg[u], u});
    }
    std::vector<int32_t>().swap(adj[v]);
  }
  for (int i = 0; i < N; ++i)
    if (dense[i]) order.push_back(i);
  return order;
}


// Approximate minimum degree on the quotient graph (Amestoy, Davis & Duff,
// "An approximate minimum degree ordering algorithm", SIAM J. Matrix Anal.
// Appl. 17(4), 1996): an eliminated node becomes an element holding the
// pattern of its column of L, the elements it covers are absorbed, each
// variable's degree is the paper's approximate external degree (bounded by
// the exact one), and variables with the same adjacency are merged into
// supervariables.  Its cost follows the matrix rather than the factor, so it
// serves matrices whose factor is large.  `wait` and the dense-row rule as in
// min_degree.
std::vector<int32_t> approx_min_degree(int N, const std::vector<int32_t>& Kp, const std::vector<int32_t>& Ki,
                                       const int8_t* wait = nullptr) {
  enum : char { VAR = 0, ELEMENT = 1, ABSORBED = 2, MERGED = 3, DENSE = 4 };
  std::vector<std::vector<int32_t>> var(N), elt(N), Le(N);
  for (int j = 0; j < N; ++j)
    for (int k = Kp[j]; k < Kp[j + 1]; ++k) {
      const int i = Ki[k];
      if (i != j) { var[i].push_back(j); var[j].push_back(i); }
    }
  for (auto& a : var) { std::sort(a.begin(), a.end()); a.erase(std::unique(a.begin(), a.end()), a.end()); }
  const int dense_thr = std::max(16, (int)(10 * std::sqrt((double)N)));
  std::vector<char> state(N, VAR);
  for (int i = 0; i < N; ++i)
    if ((int)var[i].size() > dense_thr) state[i] = DENSE;
  for (int i = 0; i < N; ++i)
    if (state[i] != DENSE) {
      auto& a = var[i];
      a.erase(std::remove_if(a.begin(), a.end(), [&](int32_t u) { return state[u] == DENSE; }), a.end());
    }
  std::vector<int> nv(N, 1), deg(N, 0), esize(N, 0), mnext(N, -1), mtail(N);
  for (int i = 0; i < N; ++i) mtail[i] = i;
  // degree lists
  std::vector<int> head(N + 2, -1), nx(N, -1), pv(N, -1);
  std::vector<char> inb(N, 0), parked(N, 0);
  int mindeg = 0, nparked = 0;
  auto ins = [&](int i) {
    const int d = std::min(deg[i], N);
    nx[i] = head[d]; pv[i] = -1;
    if (head[d] >= 0) pv[head[d]] = i;
    head[d] = i; inb[i] = 1;
    if (d < mindeg) mindeg = d;
  };
  auto rem = [&](int i) {
    const int d = std::min(deg[i], N);
    if (pv[i] >= 0) nx[pv[i]] = nx[i];
    else head[d] = nx[i];
    if (nx[i] >= 0) pv[nx[i]] = pv[i];
    inb[i] = 0;
  };
  int nleft = 0;
  for (int i = 0; i < N; ++i) {
    if (state[i] == DENSE) continue;
    ++nleft;
    deg[i] = (int)var[i].size();
    if (wait && wait[i] && !var[i].empty()) { parked[i] = 1; ++nparked; }
    else ins(i);
  }
  std::vector<int64_t> w(N, 0);
  int64_t wflg = 1;
  std::vector<int> mark(N, 0);
  int mk = 0;
  std::vector<int32_t> order, Lp;
  order.reserve(N);
  std::vector<std::pair<uint64_t, int>> hv;
  while (nleft > 0) {
    while (mindeg <= N && head[mindeg] < 0) ++mindeg;
    if (mindeg > N) {  // only waiting nodes are left
      for (int i = 0; i < N; ++i)
        if (parked[i]) { parked[i] = 0; ins(i); }
      nparked = 0;
      continue;
    }
    const int p = head[mindeg];
    rem(p);
    // the new element: the variables of p's elements and of its own edges
    mark[p] = ++mk;
    Lp.clear();
    for (int e : elt[p]) {
      if (state[e] != ELEMENT) continue;
      for (int i : Le[e])
        if (state[i] == VAR && mark[i] != mk) { mark[i] = mk; Lp.push_back(i); }
      state[e] = ABSORBED;
      std::vector<int32_t>().swap(Le[e]);
    }
    for (int i : var[p])
      if (state[i] == VAR && mark[i] != mk) { mark[i] = mk; Lp.push_back(i); }
    state[p] = ELEMENT;
    std::vector<int32_t>().swap(var[p]);
    std::vector<int32_t>().swap(elt[p]);
    for (int q = p; q >= 0; q = mnext[q]) order.push_back(q);
    nleft -= nv[p];
    int lpw = 0;
    for (int i : Lp) lpw += nv[i];
    esize[p] = lpw;
    for (int i : Lp) {
      if (inb[i]) rem(i);
      if (parked[i]) { parked[i] = 0; --nparked; }
    }
    // w(e) - wflg = the weight of e outside the new element
    for (int i : Lp)
      for (int e : elt[i])
        if (state[e] == ELEMENT && e != p) {
          if (w[e] < wflg) w[e] = wflg + esize[e];
          w[e] -= nv[i];
        }
    for (int i : Lp) {
      int64_t d = 0;
      auto& E = elt[i];
      size_t o = 0;
      for (int e : E) {
        if (state[e] != ELEMENT || e == p) continue;
        const int64_t we = w[e] - wflg;
        if (we <= 0) {  // e lies inside the new element: absorbed
          state[e] = ABSORBED;
          std::vector<int32_t>().swap(Le[e]);
          continue;
        }
        d += we;
        E[o++] = e;
      }
      E.resize(o);
      E.push_back(p);
      auto& V = var[i];
      o = 0;
      for (int j : V) {
        if (state[j] != VAR || mark[j] == mk) continue;
        d += nv[j];
        V[o++] = j;
      }
      V.resize(o);
      d += lpw - nv[i];
      d = std::min<int64_t>(d, (int64_t)deg[i] + lpw - nv[i]);
      d = std::min<int64_t>(d, (int64_t)nleft - nv[i]);
      deg[i] = (int)std::max<int64_t>(0, d);
    }
    wflg += (int64_t)N + 2;
    // supervariables: the variables of the new element with the same
    // elements and variables are merged
    hv.clear();
    for (int i : Lp) {
      if (state[i] != VAR) continue;
      uint64_t h = 0;
      for (int e : elt[i]) h += (uint64_t)e * 2654435761u + 1;
      for (int j : var[i]) h += (uint64_t)j * 40503u + 7;
      h ^= (uint64_t)(elt[i].size() * 31 + var[i].size());
      hv.push_back({h, i});
    }
    std::sort(hv.begin(), hv.end());
    for (size_t a = 0; a < hv.size();) {
      size_t b = a + 1;
      while (b < hv.size() && hv[b].first == hv[a].first) ++b;
      if (b - a > 1) {
        for (size_t x = a; x < b; ++x) {
          std::sort(elt[hv[x].second].begin(), elt[hv[x].second].end());
          std::sort(var[hv[x].second].begin(), var[hv[x].second].end());
        }
        for (size_t x = a; x < b; ++x) {
          const int i = hv[x].second;
          if (state[i] != VAR) continue;
          for (size_t y = x + 1; y < b; ++y) {
            const int j = hv[y].second;
            if (state[j] != VAR || elt[i] != elt[j] || var[i] != var[j]) continue;
            nv[i] += nv[j];
            deg[i] = std::max(0, deg[i] - nv[j]);
            nv[j] = 0;
            state[j] = MERGED;
            mnext[mtail[i]] = j;
            mtail[i] = mtail[j];
            std::vector<int32_t>().swap(var[j]);
            std::vector<int32_t>().swap(elt[j]);
          }
        }
      }
      a = b;
    }
    for (int i : Lp)
      if (state[i] == VAR) {
        ins(i);
        Le[p].push_back(i);
      }
  }
  for (int i = 0; i < N; ++i)
    if (state[i] == DENSE) order.push_back(i);
  return order;
}

// Nested dissection by recursive bisection with BFS level sets (George's
// method): degree-1 vertices (slack rows touching one variable) are ordered
// first since eliminating them creates no fill, dense rows last, and every
// remaining connected piece is split at the middle level of a BFS from a
// pseudo-peripheral vertex; the separator is ordered after both halves.
// Pieces of at most 64 vertices are ordered by minimum degree.
std::vector<int32_t> nested_dissection(int N, const std::vector<int32_t>& Kp, const std::vector<int32_t>& Ki,
                                       const int8_t* wait = nullptr) {
  std::vector<std::vector<int32_t>> adj(N);
  for (int j = 0; j < N; ++j)
    for (int k = Kp[j]; k < Kp[j + 1]; ++k) {
      int i = Ki[k];
      if (i != j) { adj[i].push_back(j); adj[j].push_back(i); }
    }
  for (auto& a : adj) { std::sort(a.begin(), a.end()); a.erase(std::unique(a.begin(), a.end()), a.end()); }
  const int dense_thr = std::max(16, (int)(10 * std::sqrt((double)N)));
  std::vector<char> dense(N, 0);
  for (int i = 0; i < N; ++i) if ((int)adj[i].size() > dense_thr) dense[i] = 1;
  for (int i = 0; i < N; ++i) {
    auto& a = adj[i];
    a.erase(std::remove_if(a.begin(), a.end(), [&](int32_t u) { return dense[u] != 0; }), a.end());
  }
  std::vector<int32_t> order;
  order.reserve(N);
  std::vector<int> tag(N, -1);
  int next_tag = 0;
  std::vector<int32_t> rest;
  for (int i = 0; i < N; ++i) {
    if (dense[i]) continue;
    if (adj[i].size() <= 1 && !(wait && wait[i] && adj[i].size() == 1)) order.push_back(i);
    else rest.push_back(i);
  }
  for (int i : order) tag[i] = -2;  // already ordered
  auto induced_deg = [&](int v, int t) {
    int d = 0;
    for (int u : adj[v]) if (tag[u] == t) ++d;
    return d;
  };
  // minimum degree restricted to the vertices tagged t
  auto md_local = [&](const std::vector<int32_t>& S, int t) {
    std::vector<std::vector<int32_t>> la(S.size());
    std::vector<int> loc(0);
    std::unordered_map<int, int> idx;
    for (size_t k = 0; k < S.size(); ++k) idx[S[k]] = (int)k;
    for (size_t k = 0; k < S.size(); ++k)
      for (int u : adj[S[k]]) if (tag[u] == t) la[k].push_back(idx[u]);
    std::vector<int32_t> lp(S.size() + 1, 0), li;
    for (size_t k = 0; k < S.size(); ++k) {
      for (int u : la[k]) if (u < (int)k) li.push_back(u);
      li.push_back((int)k);
      lp[k + 1] = (int)li.size();
    }
    std::vector<int8_t> lw;
    if (wait) {
      lw.resize(S.size());
      for (size_t k = 0; k < S.size(); ++k) lw[k] = wait[S[k]];
    }
    std::vector<int32_t> o = min_degree((int)S.size(), lp, li, wait ? lw.data() : nullptr);
    for (int k : o) order.push_back(S[k]);
  };
  std::function<void(std::vector<int32_t>&)> nd = [&](std::vector<int32_t>& S) {
    if (S.empty()) return;
    int t = next_tag++;
    for (int v : S) tag[v] = t;
    if (S.size() <= 64) { md_local(S, t); for (int v : S) tag[v] = -2; return; }
    // connected components of the induced subgraph
    std::vector<int> comp_of;
    std::vector<std::vector<int32_t>> comps;
    {
      std::vector<int32_t> q;
      int ct = next_tag++;
      for (int v : S) {
        if (tag[v] != t) continue;
        std::vector<int32_t> c;
        q.assign(1, v);
        tag[v] = ct;
        for (size_t h = 0; h < q.size(); ++h) {
          int x = q[h];
          c.push_back(x);
          for (int u : adj[x]) if (tag[u] == t) { tag[u] = ct; q.push_back(u); }
        }
        comps.push_back(std::move(c));
      }
    }
    if (comps.size() > 1) {
      for (auto& c : comps) nd(c);
      return;
    }
    std::vector<int32_t>& C = comps[0];
    int t2 = next_tag++;
    for (int v : C) tag[v] = t2;
    // pseudo-peripheral vertex and BFS levels
    std::vector<int> level(0);
    std::unordered_map<int, int> lev;
    auto bfs = [&](int root, std::vector<std::vector<int32_t>>& L) {
      L.clear();
      lev.clear();
      lev[root] = 0;
      L.push_back({root});
      while (true) {
        std::vector<int32_t> nxt;
        for (int x : L.back())
          for (int u : adj[x])
            if (tag[u] == t2 && !lev.count(u)) { lev[u] = (int)L.size(); nxt.push_back(u); }
        if (nxt.empty()) break;
        L.push_back(std::move(nxt));
      }
    };
    int root = C[0];
    for (int v : C) if (induced_deg(v, t2) < induced_deg(root, t2)) root = v;
    std::vector<std::vector<int32_t>> L;
    bfs(root, L);
    for (int rep = 0; rep < 4; ++rep) {
      int cand = L.back()[0];
      for (int v : L.back()) if (induced_deg(v, t2) < induced_deg(cand, t2)) cand = v;
      std::vector<std::vector<int32_t>> L2;
      bfs(cand, L2);
      if (L2.size() <= L.size()) break;
      L.swap(L2);
      root = cand;
    }
    if (L.size() < 3) { md_local(C, t2); for (int v : C) tag[v] = -2; return; }
    bfs(root, L);
    size_t half = C.size() / 2, acc = 0;
    int mid = 1;
    for (size_t l = 0; l < L.size(); ++l) {
      acc += L[l].size();
      if (acc >= half) { mid = (int)l; break; }
    }
    mid = std::max(1, std::min(mid, (int)L.size() - 2));
    std::vector<int32_t> A, Bv, sep;
    for (int l = 0; l < (int)L.size(); ++l) {
      if (l < mid) A.insert(A.end(), L[l].begin(), L[l].end());
      else if (l > mid) Bv.insert(Bv.end(), L[l].begin(), L[l].end());
    }
    for (int v : L[mid]) {  // keep only separator vertices that touch the far side
      bool touches = false;
      for (int u : adj[v]) if (tag[u] == t2 && lev.count(u) && lev[u] == mid + 1) { touches = true; break; }
      (touches ? sep : A).push_back(v);
    }
    for (int v : C) tag[v] = -3;  // detached until each part is tagged again
    nd(A);
    nd(Bv);
    int ts = next_tag++;
    for (int v : sep) tag[v] = ts;
    md_local(sep, ts);
    for (int v : sep) tag[v] = -2;
  };
  nd(rest);
  for (int i = 0; i < N; ++i) if (dense[i]) order.push_back(i);
  return order;
}
}  // namespace

uint64_t structure_key(int n, int m, const int32_t* Pp, const int32_t* Pi, const int32_t* Ap, const int32_t* Ai,
                       const std::vector<ConeSpec>& cones, bool fp32_order, const SymOpts& opts) {
  Hash h;
  h.v(n); h.v(m); h.v(fp32_order ? 1 : 0);
  h.v(opts.relax_max_cols); h.v(opts.relax_zero_pct); h.v(opts.group_lanes);
  // the diagnostic overrides of the ordering and of supernode relaxation
  // change the structure built for the same pattern
  for (const char* name : {"CLH_ORDERING", "CLH_RELAX"}) {
    const char* e = std::getenv(name);
    h.v(e ? 1 : 0);
    if (e) h.add(e, std::strlen(e));
  }
  // the whole stored pattern of P: the map from the caller's entries to the
  // upper triangle depends on which lower entries are stored too
  h.add(Pp, sizeof(int32_t) * (n + 1));
  h.add(Pi, sizeof(int32_t) * Pp[n]);
  h.add(Ap, sizeof(int32_t) * (n + 1));
  h.add(Ai, sizeof(int32_t) * Ap[n]);
  for (auto& c : cones) { h.v(c.type); h.v(c.dim); h.v(c.param); }
  return h.h;
}

// Entries of L for a fill-reducing order, counted without forming L: the
// elimination tree by Liu's algorithm with path compression, then for each
// row the row subtree walked from its entries up to the first node already
// marked for that row (the nonzeros of that row of L).  O(nnz(L)) time, O(N)
// memory; stops early and returns false once the count passes `cap`.
static bool count_fill(const Structure& T, const std::vector<int32_t>& order, double cap, double& nnzL,
                       double& flops) {
  const int N = T.N;
  std::vector<int32_t> ip(N);
  for (int k = 0; k < N; ++k) ip[order[k]] = k;
  // permuted lower triangle by rows: row r holds columns c < r
  std::vector<int32_t> rp(N + 1, 0);
  for (int j = 0; j < N; ++j)
    for (int k = T.Kp[j]; k < T.Kp[j + 1]; ++k) {
      const int a = ip[T.Ki[k]], b = ip[j];
      if (a != b) rp[std::max(a, b) + 1]++;
    }
  for (int i = 0; i < N; ++i) rp[i + 1] += rp[i];
  std::vector<int32_t> rc(rp[N]), nx(rp.begin(), rp.end() - 1);
  for (int j = 0; j < N; ++j)
    for (int k = T.Kp[j]; k < T.Kp[j + 1]; ++k) {
      const int a = ip[T.Ki[k]], b = ip[j];
      if (a != b) rc[nx[std::max(a, b)]++] = std::min(a, b);
    }
  std::vector<int32_t> parent(N, -1), anc(N, -1), mark(N, -1);
  for (int r = 0; r < N; ++r)
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      int i = rc[k];
      while (anc[i] != -1 && anc[i] != r) { const int t = anc[i]; anc[i] = r; i = t; }
      if (anc[i] == -1) { anc[i] = r; parent[i] = r; }
    }
  double cnt = 0;
  std::vector<int32_t> colcnt(N, 0);  // entries below the diagonal per column of L
  for (int r = 0; r < N; ++r) {
    mark[r] = r;
    for (int k = rp[r]; k < rp[r + 1]; ++k)
      for (int c = rc[k]; c != -1 && mark[c] != r; c = parent[c]) {
        mark[c] = r;
        ++colcnt[c];
        cnt += 1;
      }
    if (cnt > cap) return false;
  }
  nnzL = cnt;
  flops = 0;  // as the symbolic analysis counts them: sum of c (c + 1) over the columns
  for (int c = 0; c < N; ++c) flops += (double)colcnt[c] * (colcnt[c] + 1);
  return true;
}

// elimination tree, postorder, symbolic factor, supernodes, level schedule,
// update and assembly lists for the given fill-reducing order
// Relaxed supernodes: a column joins its child's supernode when the child is
// the previous column even if their structures differ, as long as the panel
// then holds at most relax_zero_pct percent of explicit zeros and
// relax_max_cols columns (SymOpts; by default 25% and 128).
// Chains of single-column supernodes (a cone block eliminated column by
// column before its dense part) otherwise give every later column one update
// record per earlier column, each as tall as the chain; on the serving
// shapes the fewer, wider levels are also faster (MPC stream: 15 levels to 7
// on a T = 12 request, 3793 to 4184 served solves/s).

// cost_only: stop once nnz(L), the flops and the critical-path estimate are
// known (what the choice between two orders needs), before the update records
// What the costing part of the analysis leaves for the rest of it.
struct SymState {
  std::vector<int32_t> iperm;
  std::vector<int> snode_of;
  std::vector<std::vector<std::array<int, 3>>> upd;
};
static void symbolic_finish(Structure& T, SymState& S);

// keep: with cost_only, where the costed order's state goes, so that
// symbolic_finish can complete the analysis without repeating it
static void symbolic_factor(Structure& T, const std::vector<int32_t>& order, bool cost_only = false,
                            SymState* keep = nullptr) {
  const int N = T.N;
  T.sn_col.clear(); T.sn_rows.clear(); T.upd_src.clear(); T.upd_roff.clear(); T.upd_ncol.clear();
  T.asm_dst.clear(); T.asm_src.clear(); T.asm_sgn.clear();
  std::vector<int32_t> iperm(N);
  auto lower_pattern = [&](const std::vector<int32_t>& ip, std::vector<int32_t>& Lp, std::vector<int32_t>& Li) {
    // column c of the permuted lower triangle holds rows r > c
    std::vector<int> cc(N + 1, 0);
    for (int j = 0; j < N; ++j)
      for (int k = T.Kp[j]; k < T.Kp[j + 1]; ++k) {
        int a = ip[T.Ki[k]], b = ip[j];
        if (a != b) cc[std::min(a, b) + 1]++;
      }
    Lp.assign(N + 1, 0);
    for (int i = 0; i < N; ++i) Lp[i + 1] = Lp[i] + cc[i + 1];
    Li.assign(Lp[N], 0);
    std::vector<int> nx(Lp.begin(), Lp.end() - 1);
    for (int j = 0; j < N; ++j)
      for (int k = T.Kp[j]; k < T.Kp[j + 1]; ++k) {
        int a = ip[T.Ki[k]], b = ip[j];
        if (a != b) Li[nx[std::min(a, b)]++] = std::max(a, b);
      }
  };
  auto etree_of = [&](const std::vector<int32_t>& Lp, const std::vector<int32_t>& Li) {
    // parent via the row lists (the transpose of the lower pattern, one CSR):
    // for each row r, its columns c < r with path compression
    std::vector<int32_t> rp(N + 1, 0), rc(Lp[N]);
    for (int k = 0; k < Lp[N]; ++k) rp[Li[k] + 1]++;
    for (int r = 0; r < N; ++r) rp[r + 1] += rp[r];
    {
      std::vector<int32_t> nx(rp.begin(), rp.end() - 1);
      for (int c = 0; c < N; ++c)
        for (int k = Lp[c]; k < Lp[c + 1]; ++k) rc[nx[Li[k]]++] = c;
    }
    std::vector<int> parent(N, -1), anc(N, -1);
    for (int r = 0; r < N; ++r)
      for (int k = rp[r]; k < rp[r + 1]; ++k) {
        int i = rc[k];
        while (anc[i] != -1 && anc[i] != r) { int nx = anc[i]; anc[i] = r; i = nx; }
        if (anc[i] == -1) { anc[i] = r; parent[i] = r; }
      }
    return parent;
  };
  for (int k = 0; k < N; ++k) iperm[order[k]] = k;
  std::vector<int32_t> Lp0, Li0;
  lower_pattern(iperm, Lp0, Li0);
  std::vector<int> par0 = etree_of(Lp0, Li0);
  setup_mark("etree");
  // postorder
  std::vector<int> head(N, -1), nextc(N, -1), post;
  post.reserve(N);
  for (int j = N - 1; j >= 0; --j)
    if (par0[j] >= 0) { nextc[j] = head[par0[j]]; head[par0[j]] = j; }
  std::vector<int> stack;
  for (int r = 0; r < N; ++r) {
    if (par0[r] != -1) continue;
    stack.push_back(r);
    while (!stack.empty()) {
      int v = stack.back();
      if (head[v] != -1) { int c = head[v]; head[v] = nextc[c]; stack.push_back(c); }
      else { stack.pop_back(); post.push_back(v); }
    }
  }
  T.perm.resize(N);
  for (int k = 0; k < N; ++k) T.perm[k] = order[post[k]];
  for (int k = 0; k < N; ++k) iperm[T.perm[k]] = k;
  std::vector<int32_t> Lp, Li;
  lower_pattern(iperm, Lp, Li);
  // the tree of the postordered matrix is the first tree relabelled
  // (a postorder is a topological order of the same tree)
  std::vector<int> parent(N, -1);
  {
    std::vector<int> pos(N);
    for (int k = 0; k < N; ++k) pos[post[k]] = k;
    for (int k = 0; k < N; ++k) parent[k] = par0[post[k]] >= 0 ? pos[par0[post[k]]] : -1;
  }

  setup_mark("postorder");
  // ---- column counts of L (below the diagonal)
  // Default: the row-subtree ("skeleton") method of Gilbert, Ng and Peyton
  // (SIAM J. Matrix Anal. Appl. 15, 1994), in the form of Davis, Direct
  // Methods for Sparse Linear Systems (2006), sec. 4.5: L(i, j) != 0 for
  // i > j exactly when j lies in the row subtree of i, the union of the tree
  // paths from the columns k < i of row i's entries up to i.  Visiting the
  // columns in postorder (the matrix is postordered: post[k] = k), j is a new
  // leaf of row subtree i when no column seen before in that subtree starts
  // at or after first[j] (the first descendant of j); each leaf adds one at
  // j and, after the first, takes one off at the least common ancestor with
  // the previous leaf; the counts are then summed up the tree.  O(nnz(A))
  // with path compression, against O(nnz(L)) for the row structures
  // themselves, which only the supernodes' last columns need (built below).
  // CLH_SYMBOLIC_EXPLICIT=1 (diagnostics) builds every column's structure
  // instead, the way the counts were first taken.
  std::vector<int> ccount(N);
  std::vector<std::vector<int32_t>> strc;
  const bool explicit_structs = [] {
    const char* e = std::getenv("CLH_SYMBOLIC_EXPLICIT");
    return e && e[0] == '1';
  }();
  if (explicit_structs) {
    std::vector<std::vector<int>> children(N);
    for (int j = 0; j < N; ++j) if (parent[j] >= 0) children[parent[j]].push_back(j);
    strc.resize(N);
    std::vector<int> mark(N, -1);
    for (int j = 0; j < N; ++j) {
      auto& st = strc[j];
      mark[j] = j;
      for (int k = Lp[j]; k < Lp[j + 1]; ++k) { int r = Li[k]; if (mark[r] != j) { mark[r] = j; st.push_back(r); } }
      for (int c : children[j])
        for (int r : strc[c]) if (r != j && mark[r] != j) { mark[r] = j; st.push_back(r); }
      std::sort(st.begin(), st.end());
    }
    for (int j = 0; j < N; ++j) ccount[j] = (int)strc[j].size();
  } else {
    std::vector<int> first(N, -1), maxfirst(N, -1), prevleaf(N, -1), anc(N), delta(N, 0);
    for (int k = 0; k < N; ++k) {
      int j = k;
      delta[j] = first[j] == -1 ? 1 : 0;  // a leaf of the elimination tree
      for (; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
    }
    for (int i = 0; i < N; ++i) anc[i] = i;
    for (int j = 0; j < N; ++j) {
      if (parent[j] != -1) delta[parent[j]]--;
      for (int p = Lp[j]; p < Lp[j + 1]; ++p) {
        const int i = Li[p];  // i > j
        if (first[j] <= maxfirst[i]) continue;  // j is not a leaf of row subtree i
        maxfirst[i] = first[j];
        const int jprev = prevleaf[i];
        prevleaf[i] = j;
        delta[j]++;
        if (jprev != -1) {  // a later leaf: one off at lca(jprev, j)
          int q = jprev;
          while (q != anc[q]) q = anc[q];
          for (int v = jprev; v != q;) { const int nv = anc[v]; anc[v] = q; v = nv; }
          delta[q]--;
        }
      }
      if (parent[j] != -1) anc[j] = parent[j];
    }
    for (int j = 0; j < N; ++j) if (parent[j] != -1) delta[parent[j]] += delta[j];
    for (int j = 0; j < N; ++j) ccount[j] = delta[j] - 1;  // without the diagonal
  }
  T.nnzL = 0;
  T.flops = 0;
  for (int j = 0; j < N; ++j) {
    T.nnzL += ccount[j];
    T.flops += (double)ccount[j] * (ccount[j] + 1);
  }

  setup_mark("row structures");
  // ---- supernodes: maximal runs of columns whose structures nest
  // (struct(j-1) = {j} U struct(j)), so every panel is dense, or, relaxed,
  // runs where each column is the parent of the previous one (struct(j-1)
  // minus j lies in struct(j), so the panel rows are the supernode's columns
  // and the structure of its last column; earlier columns hold zeros in the
  // rows they lack); other children of a column reach it through the update
  // lists
  std::vector<int> snode_of(N);
  T.sn_col.clear();
  {
    bool relax = T.opts.relax_max_cols > 0;
    if (const char* e = std::getenv("CLH_RELAX")) relax = std::atoi(e) != 0;  // diagnostics: force on or off
    const int relax_cols = T.opts.relax_max_cols > 0 ? T.opts.relax_max_cols : 128;
    const double relax_zeros = 0.01 * T.opts.relax_zero_pct;
    int first = 0;
    double zeros = 0;
    for (int j = 0; j < N; ++j) {
      bool merge = false;
      double dz = 0;
      if (j > 0 && parent[j - 1] == j) {
        if (ccount[j - 1] == ccount[j] + 1) {
          merge = true;
        } else if (relax) {
          const double ns = j - first + 1, nr = ns + ccount[j];
          dz = (double)(j - first) * (1 + ccount[j] - ccount[j - 1]);
          merge = ns <= relax_cols && zeros + dz <= relax_zeros * (ns * nr - 0.5 * ns * (ns - 1));
        }
      }
      if (merge) {
        zeros += dz;
      } else {
        T.sn_col.push_back(j);
        first = j;
        zeros = 0;
      }
      snode_of[j] = (int)T.sn_col.size() - 1;
    }
  }
  T.nsn = (int)T.sn_col.size();
  T.sn_col.push_back(N);
  const int nsn = T.nsn;
  T.sn_rowptr.assign(nsn + 1, 0);
  T.sn_lptr.assign(nsn + 1, 0);
  for (int s = 0; s < nsn; ++s) {
    int f = T.sn_col[s], ns = T.sn_col[s + 1] - f;
    int nr = ns + ccount[f + ns - 1];
    T.sn_rowptr[s + 1] = T.sn_rowptr[s] + nr;
    T.sn_lptr[s + 1] = T.sn_lptr[s] + (int64_t)nr * ns;
  }
  T.Lsize = T.sn_lptr[nsn];
  T.sn_iptr.assign(nsn + 1, 0);
  T.maxns = 1;
  {
    std::vector<int> heights;
    for (int s = 0; s < nsn; ++s) {
      int ns = T.sn_col[s + 1] - T.sn_col[s];
      T.sn_iptr[s + 1] = T.sn_iptr[s] + (ns > 1 ? (int64_t)ns * ns : 0);
      T.maxns = std::max(T.maxns, ns);
      if (ns > 1) heights.push_back(T.sn_rowptr[s + 1] - T.sn_rowptr[s]);
    }
    int p90 = 0;
    if (!heights.empty()) {
      std::sort(heights.begin(), heights.end());
      p90 = heights[(heights.size() * 9) / 10];
    }
    T.group_hint = T.opts.group_lanes > 0 ? T.opts.group_lanes : (p90 <= 48 ? 16 : (p90 <= 96 ? 32 : 64));
  }
  T.sn_rows.assign(T.sn_rowptr[nsn], 0);
  if (explicit_structs) {
    for (int s = 0; s < nsn; ++s) {
      const int f = T.sn_col[s], l = T.sn_col[s + 1] - 1;
      int p = T.sn_rowptr[s];
      for (int c = f; c <= l; ++c) T.sn_rows[p++] = c;
      for (int r : strc[l]) T.sn_rows[p++] = r;
    }
  } else {
    // the supernodes' row lists, children first (postorder): the rows
    // below the last column l of supernode s are those of its columns'
    // entries and of its child supernodes' lists that lie below l (a
    // column's structure is its entries and its children's structures;
    // within s each column is the parent of the one before it)
    std::vector<std::vector<int>> snch(nsn);
    std::vector<int> mark(N, -1), lst;
    for (int s = 0; s < nsn; ++s) {
      const int f = T.sn_col[s], l = T.sn_col[s + 1] - 1;
      lst.clear();
      for (int c = f; c <= l; ++c)
        for (int k = Lp[c]; k < Lp[c + 1]; ++k) {
          const int r = Li[k];
          if (r > l && mark[r] != s) { mark[r] = s; lst.push_back(r); }
        }
      for (int t : snch[s]) {
        const int nst = T.sn_col[t + 1] - T.sn_col[t];
        for (int q = T.sn_rowptr[t] + nst; q < T.sn_rowptr[t + 1]; ++q) {
          const int r = T.sn_rows[q];
          if (r > l && mark[r] != s) { mark[r] = s; lst.push_back(r); }
        }
      }
      std::sort(lst.begin(), lst.end());
      if ((int)lst.size() != ccount[l])
        throw std::logic_error("symbolic analysis: a supernode's row list disagrees with its column count");
      int p = T.sn_rowptr[s];
      for (int c = f; c <= l; ++c) T.sn_rows[p++] = c;
      for (int r : lst) T.sn_rows[p++] = r;
      if (!lst.empty()) snch[snode_of[lst[0]]].push_back(s);
    }
  }
  setup_mark("supernodes");
  // supernodal tree and levels (height above the leaves)
  std::vector<int> snpar(nsn, -1), lev(nsn, 0);
  for (int s = 0; s < nsn; ++s) {
    int ns = T.sn_col[s + 1] - T.sn_col[s];
    int nr = T.sn_rowptr[s + 1] - T.sn_rowptr[s];
    if (nr > ns) snpar[s] = snode_of[T.sn_rows[T.sn_rowptr[s] + ns]];
  }
  for (int s = 0; s < nsn; ++s)
    if (snpar[s] >= 0) lev[snpar[s]] = std::max(lev[snpar[s]], lev[s] + 1);
  T.nlev = 0;
  for (int s = 0; s < nsn; ++s) T.nlev = std::max(T.nlev, lev[s] + 1);
  T.lev_ptr.assign(T.nlev + 1, 0);
  for (int s = 0; s < nsn; ++s) T.lev_ptr[lev[s] + 1]++;
  for (int l = 0; l < T.nlev; ++l) T.lev_ptr[l + 1] += T.lev_ptr[l];
  T.lev_sn.assign(nsn, 0);
  {
    // level 0 lists its single-column supernodes first: the engine gives
    // each of them one thread instead of one wavefront
    std::vector<int> nx(T.lev_ptr.begin(), T.lev_ptr.end() - 1);
    T.nleaf = 0;
    for (int s = 0; s < nsn; ++s)
      if (lev[s] == 0 && T.sn_col[s + 1] - T.sn_col[s] == 1) { T.lev_sn[nx[0]++] = s; T.nleaf++; }
    for (int s = 0; s < nsn; ++s)
      if (!(lev[s] == 0 && T.sn_col[s + 1] - T.sn_col[s] == 1)) T.lev_sn[nx[lev[s]]++] = s;
  }
  setup_mark("tree and levels");
  if (std::getenv("CLH_SETUP_TIMING")) {  // the top levels' shapes
    for (int l = std::max(0, T.nlev - 4); l < T.nlev; ++l) {
      int mx = 0, cnt = T.lev_ptr[l + 1] - T.lev_ptr[l];
      for (int t = T.lev_ptr[l]; t < T.lev_ptr[l + 1]; ++t) mx = std::max(mx, T.sn_col[T.lev_sn[t] + 1] - T.sn_col[T.lev_sn[t]]);
      std::fprintf(stderr, "[clh setup]   level %d: %d supernodes, widest %d columns\n", l, cnt, mx);
    }
  }
  // update lists grouped by target supernode
  std::vector<std::vector<std::array<int, 3>>> upd(nsn);
  for (int d = 0; d < nsn; ++d) {
    int r0 = T.sn_rowptr[d], nr = T.sn_rowptr[d + 1] - r0;
    int ns = T.sn_col[d + 1] - T.sn_col[d];
    int t = ns;
    while (t < nr) {
      int tgt = snode_of[T.sn_rows[r0 + t]];
      int e = t;
      while (e < nr && snode_of[T.sn_rows[r0 + e]] == tgt) ++e;
      upd[tgt].push_back({d, t, e - t});
      t = e;
    }
  }
  T.upd_ptr.assign(nsn + 1, 0);
  for (int s = 0; s < nsn; ++s) T.upd_ptr[s + 1] = T.upd_ptr[s] + (int)upd[s].size();
  setup_mark("update lists");
  // critical path of the level-synchronous factorization with 4 wavefronts:
  // a supernode costs ~4 steps of setup, one per column and ~2 per update
  {
    const int nw = 4;
    T.path_cost = 0;
    for (int l = 0; l < T.nlev; ++l) {
      double mx = 0, sum = 0, leaves = 0;
      for (int t = T.lev_ptr[l]; t < T.lev_ptr[l + 1]; ++t) {
        int sn = T.lev_sn[t];
        if (l == 0 && t < T.lev_ptr[0] + T.nleaf) { leaves += 1; continue; }
        double c = 4 + (T.sn_col[sn + 1] - T.sn_col[sn]) + 2.0 * upd[sn].size();
        mx = std::max(mx, c);
        sum += c;
      }
      T.path_cost += 1 + std::max(mx, sum / nw) + leaves / 256.0 * 4;
    }
  }
  SymState st;
  st.iperm = std::move(iperm);
  st.snode_of = std::move(snode_of);
  st.upd = std::move(upd);
  if (cost_only) {
    if (keep) *keep = std::move(st);
    return;
  }
  symbolic_finish(T, st);
}

// update records, assembly lists and row-owner lists of a costed order
static void symbolic_finish(Structure& T, SymState& S) {
  const int N = T.N, nsn = T.nsn;
  const std::vector<int32_t>& iperm = S.iperm;
  const std::vector<int>& snode_of = S.snode_of;
  auto& upd = S.upd;
  for (int s = 0; s < nsn; ++s)
    for (auto& u : upd[s]) { T.upd_src.push_back(u[0]); T.upd_roff.push_back(u[1]); T.upd_ncol.push_back(u[2]); }
  setup_mark("critical path");
  // packed update records + relative row positions (no row map at run time)
  T.upd.clear();
  T.upd_rel.clear();
  for (int s = 0; s < nsn; ++s) {
    const int r0s = T.sn_rowptr[s], nrs = T.sn_rowptr[s + 1] - r0s;
    const int32_t* rows_s = T.sn_rows.data() + r0s;
    for (auto& u : upd[s]) {
      const int d = u[0], roff = u[1];
      UpdDesc r{};
      r.src = d; r.roff = roff; r.ncol = u[2];
      r.fd = T.sn_col[d]; r.nsd = T.sn_col[d + 1] - T.sn_col[d];
      r.nrd = T.sn_rowptr[d + 1] - T.sn_rowptr[d];
      r.rel = (int32_t)T.upd_rel.size();
      r.lptr = T.sn_lptr[d];
      for (int a = roff; a < r.nrd; ++a) {
        const int row = T.sn_rows[T.sn_rowptr[d] + a];
        const int pos = (int)(std::lower_bound(rows_s, rows_s + nrs, row) - rows_s);
        if (pos >= nrs || rows_s[pos] != row) throw std::runtime_error("internal: update row outside the target");
        T.upd_rel.push_back(pos);
      }
      T.upd.push_back(r);
    }
  }
  setup_mark("packed updates");
  // assembly of the original entries into the panels: each KKT entry's
  // supernode and panel position (independent binary searches, split over
  // host threads), then grouped by supernode in entry order
  const int nnzK = T.Kp[N];
  std::vector<int32_t> e_sn(nnzK), e_dst(nnzK);
  {
    std::atomic<bool> outside(false);
    auto work = [&](int j0, int j1) {
      for (int j = j0; j < j1; ++j)
        for (int k = T.Kp[j]; k < T.Kp[j + 1]; ++k) {
          const int a = iperm[T.Ki[k]], b = iperm[j];
          const int c = std::min(a, b), r = std::max(a, b);
          const int sn = snode_of[c];
          const int r0 = T.sn_rowptr[sn], nr = T.sn_rowptr[sn + 1] - r0;
          const int32_t* rows = T.sn_rows.data() + r0;
          const int pos = (int)(std::lower_bound(rows, rows + nr, r) - rows);
          if (pos >= nr || rows[pos] != r) outside = true;
          e_sn[k] = sn;
          e_dst[k] = pos + (c - T.sn_col[sn]) * nr;
        }
    };
    parallel_columns(N, nnzK, T.Kp.data(), work);
    if (outside) throw std::runtime_error("internal: entry outside the symbolic pattern");
  }
  T.asm_ptr.assign(nsn + 1, 0);
  for (int k = 0; k < nnzK; ++k) T.asm_ptr[e_sn[k] + 1]++;
  for (int sn = 0; sn < nsn; ++sn) T.asm_ptr[sn + 1] += T.asm_ptr[sn];
  std::vector<std::array<int, 3>> asm_rec(nnzK);
  {
    std::vector<int32_t> nx(T.asm_ptr.begin(), T.asm_ptr.end() - 1);
    for (int j = 0; j < N; ++j)
      for (int k = T.Kp[j]; k < T.Kp[j + 1]; ++k)
        asm_rec[nx[e_sn[k]]++] = {e_dst[k], k, T.Ki[k] == j ? (int)T.dsign[j] : 0};
  }
  setup_mark("assembly");
  // lanes per supernode at each level: levels with few supernodes give each
  // one more lanes (and so a larger share of the LDS area)
  T.lev_group.assign(T.nlev, T.group_hint);
  for (int l = 0; l < T.nlev && T.opts.group_lanes == 0; ++l) {
    const int cnt = T.lev_ptr[l + 1] - T.lev_ptr[l] - (l == 0 ? T.nleaf : 0);
    const int g = cnt <= 4 ? 64 : (cnt <= 8 ? 32 : 16);
    T.lev_group[l] = std::max(g, T.group_hint);
  }
  T.sn.assign(nsn, SnDesc{});
  for (int s = 0; s < nsn; ++s) {
    SnDesc& r = T.sn[s];
    r.f = T.sn_col[s]; r.ns = T.sn_col[s + 1] - T.sn_col[s];
    r.r0 = T.sn_rowptr[s]; r.nr = T.sn_rowptr[s + 1] - T.sn_rowptr[s];
    r.asm0 = T.asm_ptr[s]; r.asm1 = T.asm_ptr[s + 1];
    r.upd0 = T.upd_ptr[s]; r.upd1 = T.upd_ptr[s + 1];
    r.lptr = T.sn_lptr[s]; r.iptr = T.sn_iptr[s];
  }
  T.sn_lev.resize(nsn);
  for (int t = 0; t < nsn; ++t) T.sn_lev[t] = T.sn[T.lev_sn[t]];
  setup_mark("lanes");
  // row-owner form of the updates: for every row of every target panel, the
  // contributions that land on it, in the order of the update list (so the
  // per-entry summation order is that of the update loop)
  {
    // counted first, then filled in one array: the records of each row keep
    // the order of the update loop
    const size_t nrows = T.sn_rows.size();
    T.rowupd_ptr.assign(nrows + 1, 0);
    for (int s = 0; s < nsn; ++s)
      for (int u = T.upd_ptr[s]; u < T.upd_ptr[s + 1]; ++u) {
        const UpdDesc& ud = T.upd[u];
        for (int a = 0; a < ud.nrd - ud.roff; ++a) ++T.rowupd_ptr[T.sn_rowptr[s] + T.upd_rel[ud.rel + a] + 1];
      }
    for (size_t p = 0; p < nrows; ++p) T.rowupd_ptr[p + 1] += T.rowupd_ptr[p];
    T.rowupd.assign((size_t)T.rowupd_ptr[nrows], RowUpd{});
    std::vector<int32_t> fill(T.rowupd_ptr.begin(), T.rowupd_ptr.end() - 1);
    for (int s = 0; s < nsn; ++s)
      for (int u = T.upd_ptr[s]; u < T.upd_ptr[s + 1]; ++u) {
        const UpdDesc& ud = T.upd[u];
        const int nrem = ud.nrd - ud.roff;
        for (int a = 0; a < nrem; ++a) {
          RowUpd& r = T.rowupd[fill[T.sn_rowptr[s] + T.upd_rel[ud.rel + a]]++];
          r.base = ud.lptr + ud.roff;
          r.a = a; r.nsd = ud.nsd; r.nrd = ud.nrd; r.fd = ud.fd;
          r.bmax = std::min(a + 1, ud.ncol);
          r.rel = ud.rel;
        }
      }
    T.rowupd.push_back(RowUpd{});  // readable record past the end: the device loops load one ahead
  }
  T.asm_dst.resize(nnzK);
  T.asm_src.resize(nnzK);
  T.asm_sgn.resize(nnzK);
  for (int t = 0; t < nnzK; ++t) {
    T.asm_dst[t] = asm_rec[t][0];
    T.asm_src[t] = asm_rec[t][1];
    T.asm_sgn[t] = (int8_t)asm_rec[t][2];
  }
  setup_mark("row-owner updates");
}

// The refinement CSR of K, the ordering and the supernodal symbolic analysis
// of a structure whose triu CSC pattern (Kp, Ki), N, Kdiag and dsign are set.
// matrices up to this dimension are ordered by exact minimum degree
constexpr int kExactOrderMaxN = 4000;

static void finish_kkt_symbolic(Structure& T, bool fp32_order) {
  const int N = T.N;
  // full symmetric CSR of K for the refinement residual
  {
    std::vector<int> rc(N + 1, 0);
    for (int j = 0; j < N; ++j)
      for (int k = T.Kp[j]; k < T.Kp[j + 1]; ++k) {
        int i = T.Ki[k];
        rc[i + 1]++;
        if (i != j) rc[j + 1]++;
      }
    T.Kfp.assign(N + 1, 0);
    for (int i = 0; i < N; ++i) T.Kfp[i + 1] = T.Kfp[i] + rc[i + 1];
    T.Kfi.assign(T.Kfp[N], 0);
    T.Kfk.assign(T.Kfp[N], 0);
    std::vector<int> nx(T.Kfp.begin(), T.Kfp.end() - 1);
    for (int j = 0; j < N; ++j)
      for (int k = T.Kp[j]; k < T.Kp[j + 1]; ++k) {
        int i = T.Ki[k];
        int p0 = nx[i]++;
        T.Kfi[p0] = j; T.Kfk[p0] = k;
        if (i != j) { int p1 = nx[j]++; T.Kfi[p1] = i; T.Kfk[p1] = k; }
      }
  }

  setup_mark("refinement CSR");
  // ---- ordering: minimum degree, and nested dissection when it gives a
  // shallower elimination tree at comparable factor cost
  {
    // CLH_ORDERING=md|amd|nd forces one ordering (diagnostics); default picks
    const char* env = std::getenv("CLH_ORDERING");
    std::string mode = env ? env : "auto";
    // for FP32 factors, pivots that may be tiny (constraint rows, variables
    // without a P diagonal) wait for a neighbour (see min_degree); FP64
    // factors keep the unconstrained order
    std::vector<int8_t> wait(N, 0);
    if (fp32_order)
      for (int i = 0; i < N; ++i) wait[i] = (int)T.weak.size() == N ? T.weak[i] : (T.dsign[i] < 0 ? 1 : 0);
    const int8_t* w = fp32_order ? wait.data() : nullptr;  // (may be dropped below)
    if (mode == "nd") {
      symbolic_factor(T, nested_dissection(N, T.Kp, T.Ki, w));
    } else {
      // minimum degree on the explicit elimination graph for small matrices
      // (the serving shapes); its cost grows with the factor, so larger ones
      // are ordered by approximate minimum degree on the quotient graph
      const bool exact = mode == "md" || (mode != "amd" && N <= kExactOrderMaxN);
      // large matrices take the FP64 order: under the FP32 constraint the
      // nested-dissection candidate won on critical path with twice the fill
      // on the full random sparse SOCP (3.2e8 factor entries against 1.8e8,
      // an 11 987-column root under a chain of 110 levels) and its FP32
      // solves ran five times slower than the FP64 ones; refinement, or the
      // switch to FP64 factors, answers for the pivots the constraint protected
      if (!exact) w = nullptr;
      const std::vector<int32_t> first = exact ? min_degree(N, T.Kp, T.Ki, w) : approx_min_degree(N, T.Kp, T.Ki, w);
      // both candidates are costed first; only the one taken is analysed in full
      SymState first_state;
      bool try_nd = mode != "md" && N >= 128;
      // the nested-dissection candidate is ordered on another host thread
      // while the first order is costed: both only read the KKT pattern, and
      // each result is the one a serial run gives
      std::future<std::vector<int32_t>> nd_pending;
      if (try_nd) nd_pending = std::async(std::launch::async, [&] { return nested_dissection(N, T.Kp, T.Ki, w); });
      symbolic_factor(T, first, try_nd, &first_state);
      setup_mark(try_nd ? "first order costed" : "first order analysed");
      double nnz_nd = 0;
      std::vector<int32_t> nd_order;
      if (try_nd) nd_order = nd_pending.get();
      // a nested-dissection order with far more fill is not analysed (its
      // structure alone could exhaust memory)
      setup_mark("nested dissection order");
      double flops_nd = 0;
      // (an order whose factorization would cost more than 8 times the
      // first's is not taken: its column counts decide that before any
      // further analysis)
      // (for factorizations large enough to fill the GPU, at least 1e10
      // flops, the work itself bounds the time: an order with half as much
      // work again is not analysed further -- on the full random sparse SOCP
      // nested dissection's 5.2e12 flops against 2.2e12, whose row
      // structures alone took 60% of the setup)
      const double flops_cap = T.flops >= 1e10 ? 1.5 * T.flops : 8.0 * T.flops + 1e5;
      if (!nd_order.empty() && count_fill(T, nd_order, 8.0 * (double)T.nnzL + 1e5, nnz_nd, flops_nd) &&
          flops_nd <= flops_cap) {
        // the costs of the nested-dissection order first; the rest of its
        // analysis only when it is taken
        Structure alt = T;
        SymState nd_state;
        symbolic_factor(alt, nd_order, true, &nd_state);
        const bool take = alt.path_cost < T.path_cost && alt.flops <= 8.0 * T.flops + 1e5;
        if (std::getenv("CLH_SETUP_TIMING"))
          std::fprintf(stderr, "[clh setup]   first order: flops %.3g, path %.4g; nested dissection: flops %.3g, path %.4g\n",
                       T.flops, T.path_cost, alt.flops, alt.path_cost);
        setup_mark(take ? "nested dissection costed, taken" : "nested dissection costed, not taken");
        if (take) {
          symbolic_finish(alt, nd_state);
          T = std::move(alt);
          try_nd = false;  // done
        }
      }
      if (try_nd) symbolic_finish(T, first_state);
    }
  }
}

std::shared_ptr<Structure> build_structure(int n, int m, const int32_t* Pp, const int32_t* Pi, const int32_t* Ap,
                                           const int32_t* Ai, const std::vector<ConeSpec>& cones, bool fp32_order,
                                           const SymOpts& opts) {
  auto S = std::make_shared<Structure>();
  Structure& T = *S;
  T.n = n;
  T.m = m;
  T.opts = opts;
  T.key = structure_key(n, m, Pp, Pi, Ap, Ai, cones, fp32_order, opts);
  T.fp32_order = fp32_order;
  T.ucones = cones;
  // ---- validate
  for (int j = 0; j < n; ++j) {
    if (Pp[j + 1] < Pp[j] || Ap[j + 1] < Ap[j]) throw std::runtime_error("invalid column pointers");
    for (int k = Pp[j]; k < Pp[j + 1]; ++k)
      if (Pi[k] < 0 || Pi[k] >= n) throw std::runtime_error("P row index out of range");
    for (int k = Ap[j]; k < Ap[j + 1]; ++k)
      if (Ai[k] < 0 || Ai[k] >= m) throw std::runtime_error("A row index out of range");
  }
  int total = 0;
  for (auto& c : cones) {
    if (c.type < CT_ZERO || c.type > CT_PSD) throw std::runtime_error("unknown cone type");
    if (c.type == CT_PSD && (c.side < 0 || c.side > 32 || c.dim != tri(c.side)))
      throw std::runtime_error("PSD cone side must be in 0..32 (numel n(n+1)/2)");
    if (c.type == CT_SOC && c.dim < 2) throw std::runtime_error("second-order cone dimension must be >= 2");
    if ((c.type == CT_EXP || c.type == CT_POW) && c.dim != 3) throw std::runtime_error("exp/pow cones have dim 3");
    if (c.type == CT_POW && !(c.param > 0 && c.param < 1)) throw std::runtime_error("power cone alpha must be in (0,1)");
    if (c.dim < 0) throw std::runtime_error("negative cone dimension");
    total += c.dim;
  }
  if (total != m) throw std::runtime_error("cone dimensions do not add up to the number of rows of A");
  T.uPp.assign(Pp, Pp + n + 1); T.uPi.assign(Pi, Pi + Pp[n]);
  T.uAp.assign(Ap, Ap + n + 1); T.uAi.assign(Ai, Ai + Ap[n]);

  // ---- P: upper triangle, full symmetric CSR
  std::vector<int> cnt(n + 1, 0);
  for (int j = 0; j < n; ++j)
    for (int k = Pp[j]; k < Pp[j + 1]; ++k) {
      int i = Pi[k];
      if (i > j) continue;
      T.Puser_keep.push_back(k);
      cnt[i + 1]++;
      if (i != j) cnt[j + 1]++;
    }
  T.Pp.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) T.Pp[i + 1] = T.Pp[i] + cnt[i + 1];
  T.Pi.assign(T.Pp[n], 0);
  std::vector<int> nxt(T.Pp.begin(), T.Pp.end() - 1);
  T.Pfull_pos0.resize(T.Puser_keep.size());
  T.Pfull_pos1.resize(T.Puser_keep.size());
  for (size_t t = 0; t < T.Puser_keep.size(); ++t) {
    int k = T.Puser_keep[t];
    int i = Pi[k];
    int j = (int)(std::upper_bound(Pp, Pp + n + 1, k) - Pp) - 1;
    int p0 = nxt[i]++;
    T.Pi[p0] = j;
    T.Pfull_pos0[t] = p0;
    if (i != j) {
      int p1 = nxt[j]++;
      T.Pi[p1] = i;
      T.Pfull_pos1[t] = p1;
    } else {
      T.Pfull_pos1[t] = -1;
    }
  }
  // ---- A: CSR + CSC
  T.Atp.assign(Ap, Ap + n + 1);
  T.Ati.assign(Ai, Ai + Ap[n]);
  int nnzA = Ap[n];
  T.Ap.assign(m + 1, 0);
  for (int k = 0; k < nnzA; ++k) T.Ap[Ai[k] + 1]++;
  for (int i = 0; i < m; ++i) T.Ap[i + 1] += T.Ap[i];
  T.Ai.assign(nnzA, 0);
  T.csc_to_csr.assign(nnzA, 0);
  nxt.assign(T.Ap.begin(), T.Ap.end() - 1);
  for (int j = 0; j < n; ++j)
    for (int k = Ap[j]; k < Ap[j + 1]; ++k) {
      int p = nxt[Ai[k]]++;
      T.Ai[p] = j;
      T.csc_to_csr[k] = p;
    }

  // ---- cones
  T.etype.assign(m, 0);
  int off = 0, nsp = 0;
  for (auto& c : cones) {
    for (int i = 0; i < c.dim; ++i) T.etype[off + i] = (int8_t)c.type;
    if (c.type == CT_NONNEG) T.degree += c.dim;
    if (c.type == CT_SOC) {
      T.degree += 1;
      T.soc_off.push_back(off);
      T.soc_dim.push_back(c.dim);
      bool sparse = c.dim > 5;  // SOC_NO_EXPANSION_MAX_SIZE (src/Clarabel.jl:9)
      T.soc_sp.push_back(sparse ? nsp : -1);
      if (sparse) nsp++;
    }
    if (c.type == CT_EXP) { T.degree += 3; T.exp_off.push_back(off); T.symmetric = false; }
    if (c.type == CT_POW) {
      T.degree += 3;
      T.pow_off.push_back(off);
      T.pow_alpha.push_back(c.param);
      T.symmetric = false;
    }
    if (c.type == CT_PSD) {
      T.degree += c.side;
      T.psd_off.push_back(off);
      T.psd_n.push_back(c.side);
      T.psd_soff.push_back(T.psd_state);
      T.psd_woff.push_back(T.psd_work);
      T.psd_state += 2 * c.side * c.side + 2 * c.side;
      T.psd_work += 8 * c.side * c.side;
    }
    off += c.dim;
  }
  T.p = 2 * nsp;
  T.N = n + m + T.p;
  const int N = T.N;

  // ---- KKT upper triangle, column by column, in the reference's layout
  // (src/kktsolvers/direct-ldl/directldl_kkt_assembly.jl:13-200)
  struct Ent { int row; int tag; int a; int b; };
  enum { E_P, E_PDIAG, E_A, E_LIN, E_SOC, E_NSYM, E_SPV, E_SPU, E_SPD, E_PSD };
  std::vector<std::vector<Ent>> col(N);
  for (int j = 0; j < n; ++j) {
    bool hasdiag = false;
    for (int k = T.Pp[j]; k < T.Pp[j + 1]; ++k) {
      int c = T.Pi[k];  // full row j holds P(j,c) = P(c,j); keep rows c <= j of column j
      if (c <= j) { col[j].push_back({c, E_P, k, 0}); if (c == j) hasdiag = true; }
    }
    if (!hasdiag) col[j].push_back({j, E_PDIAG, 0, 0});
  }
  for (int i = 0; i < m; ++i)
    for (int k = T.Ap[i]; k < T.Ap[i + 1]; ++k) col[n + i].push_back({T.Ai[k], E_A, k, 0});
  off = 0;
  int isoc = 0, iexp = 0, ipow = 0, ipsd = 0, pcol = n + m;
  for (auto& c : cones) {
    int r0 = n + off;
    if (c.type == CT_ZERO || c.type == CT_NONNEG) {
      for (int i = 0; i < c.dim; ++i) col[r0 + i].push_back({r0 + i, E_LIN, off + i, 0});
    } else if (c.type == CT_SOC && T.soc_sp[isoc] < 0) {
      int h = 0;
      for (int cc = 0; cc < c.dim; ++cc)
        for (int rr = 0; rr <= cc; ++rr) col[r0 + cc].push_back({r0 + rr, E_SOC, isoc, h++});
    } else if (c.type == CT_SOC) {
      for (int i = 0; i < c.dim; ++i) col[r0 + i].push_back({r0 + i, E_SOC, isoc, i});
      int sp = T.soc_sp[isoc];
      for (int i = 0; i < c.dim; ++i) col[pcol].push_back({r0 + i, E_SPV, sp, i});
      col[pcol].push_back({pcol, E_SPD, sp, 0});
      for (int i = 0; i < c.dim; ++i) col[pcol + 1].push_back({r0 + i, E_SPU, sp, i});
      col[pcol + 1].push_back({pcol + 1, E_SPD, sp, 1});
      pcol += 2;
    } else if (c.type == CT_PSD) {  // dense packed triu t x t block
      int h = 0;
      for (int cc = 0; cc < c.dim; ++cc)
        for (int rr = 0; rr <= cc; ++rr) col[r0 + cc].push_back({r0 + rr, E_PSD, ipsd, h++});
    } else {  // exp / pow: packed triu 3x3
      int which = (c.type == CT_EXP) ? iexp : (int)(cones.size() + ipow);
      int h = 0;
      for (int cc = 0; cc < 3; ++cc)
        for (int rr = 0; rr <= cc; ++rr) col[r0 + cc].push_back({r0 + rr, E_NSYM, which, h++});
    }
    if (c.type == CT_PSD) ipsd++;
    if (c.type == CT_SOC) isoc++;
    if (c.type == CT_EXP) iexp++;
    if (c.type == CT_POW) ipow++;
    off += c.dim;
  }
  // sizes of the cone maps
  T.soc_kblk.resize(T.soc_off.size());
  {
    int acc = 0;
    for (size_t k = 0; k < T.soc_off.size(); ++k) {
      T.soc_kblk[k] = acc;
      acc += T.soc_sp[k] < 0 ? tri(T.soc_dim[k]) : T.soc_dim[k];
    }
    T.soc_kidx.assign(acc, -1);
    int su = 0;
    T.sp_off.assign(nsp, 0);
    for (size_t k = 0; k < T.soc_off.size(); ++k)
      if (T.soc_sp[k] >= 0) { T.sp_off[T.soc_sp[k]] = su; su += T.soc_dim[k]; }
    T.sp_ku.assign(su, -1);
    T.sp_kv.assign(su, -1);
    T.sp_kD.assign(2 * nsp, -1);
  }
  {
    int acc = 0;
    T.psd_kblk.resize(T.psd_off.size());
    for (size_t k = 0; k < T.psd_off.size(); ++k) {
      T.psd_kblk[k] = acc;
      acc += tri(tri(T.psd_n[k]));
    }
    T.psd_kidx.assign(acc, -1);
  }
  T.exp_kidx.assign(6 * T.exp_off.size(), -1);
  T.pow_kidx.assign(6 * T.pow_off.size(), -1);
  T.lin_kdiag.assign(m, -1);
  T.Pk.assign(T.Pp[n], -1);
  T.Ak.assign(nnzA, -1);
  T.Kdiag.assign(N, -1);
  T.Kp.assign(N + 1, 0);
  for (int j = 0; j < N; ++j) T.Kp[j + 1] = T.Kp[j] + (int)col[j].size();
  T.nnzK = T.Kp[N];
  T.Ki.assign(T.nnzK, 0);
  for (int j = 0; j < N; ++j) {
    auto& cj = col[j];
    std::stable_sort(cj.begin(), cj.end(), [](const Ent& x, const Ent& y) { return x.row < y.row; });
    for (size_t t = 0; t < cj.size(); ++t) {
      int kk = T.Kp[j] + (int)t;
      const Ent& e = cj[t];
      if (t > 0 && cj[t - 1].row == e.row) throw std::runtime_error("duplicate KKT entry (repeated index in P or A?)");
      T.Ki[kk] = e.row;
      if (e.row == j) T.Kdiag[j] = kk;
      switch (e.tag) {
        case E_P: T.Pk[e.a] = kk; break;
        case E_A: T.Ak[e.a] = kk; break;
        case E_LIN: T.lin_kdiag[e.a] = kk; break;
        case E_SOC: T.soc_kidx[T.soc_kblk[e.a] + e.b] = kk; break;
        case E_NSYM:
          if (e.a < (int)cones.size()) T.exp_kidx[6 * e.a + e.b] = kk;
          else T.pow_kidx[6 * (e.a - (int)cones.size()) + e.b] = kk;
          break;
        case E_SPV: T.sp_kv[T.sp_off[e.a] + e.b] = kk; break;
        case E_SPU: T.sp_ku[T.sp_off[e.a] + e.b] = kk; break;
        case E_SPD: T.sp_kD[2 * e.a + e.b] = kk; break;
        case E_PSD: T.psd_kidx[T.psd_kblk[e.a] + e.b] = kk; break;
        default: break;
      }
    }
  }
  for (int j = 0; j < N; ++j)
    if (T.Kdiag[j] < 0) throw std::runtime_error("internal: KKT diagonal missing");
  // D signs (kktsolver_directldl.jl:120-137)
  T.dsign.assign(N, 1);
  for (int i = n; i < n + m; ++i) T.dsign[i] = -1;
  // pivots that are the bare regularization until a neighbour is eliminated:
  // equality rows (Hs = 0) and variables without a P diagonal
  T.weak.assign(N, 0);
  for (int j = 0; j < n; ++j) {
    T.weak[j] = 1;
    for (int k = T.Pp[j]; k < T.Pp[j + 1]; ++k)
      if (T.Pi[k] == j) T.weak[j] = 0;
  }
  for (int i = 0; i < m; ++i) T.weak[n + i] = T.etype[i] == CT_ZERO ? 1 : 0;
  for (int sp = 0; sp < nsp; ++sp) { T.dsign[n + m + 2 * sp] = -1; T.dsign[n + m + 2 * sp + 1] = 1; }
  finish_kkt_symbolic(T, fp32_order);
  return S;
}

std::shared_ptr<Structure> build_ldl_structure(int N, const int32_t* rowptr, const int32_t* colval, int base,
                                               const int32_t* dsigns, std::vector<int32_t>& csr_of_K,
                                               bool fp32_order, const SymOpts& opts) {
  if (N <= 0 || !rowptr || !colval || !dsigns) throw std::runtime_error("invalid matrix");
  if (base != 0 && base != 1) throw std::runtime_error("index base must be 0 or 1");
  const int nnz = rowptr[N] - base;
  if (rowptr[0] != base || nnz < N) throw std::runtime_error("invalid row pointers");
  auto S = std::make_shared<Structure>();
  Structure& T = *S;
  T.N = N;
  T.opts = opts;
  T.Pp.assign(1, 0);
  T.Ap.assign(1, 0);
  T.Atp.assign(1, 0);
  // upper triangle (row <= col) by columns, with the CSR position of each entry
  std::vector<std::vector<std::pair<int, int>>> col(N);
  for (int i = 0; i < N; ++i) {
    if (rowptr[i + 1] < rowptr[i]) throw std::runtime_error("invalid row pointers");
    for (int p = rowptr[i] - base; p < rowptr[i + 1] - base; ++p) {
      const int j = colval[p] - base;
      if (j < 0 || j >= N) throw std::runtime_error("column index out of range");
      if (i <= j) col[j].push_back({i, p});
    }
  }
  T.Kp.assign(N + 1, 0);
  for (int j = 0; j < N; ++j) T.Kp[j + 1] = T.Kp[j] + (int)col[j].size();
  T.nnzK = T.Kp[N];
  T.Ki.assign(T.nnzK, 0);
  csr_of_K.assign(T.nnzK, 0);
  T.Kdiag.assign(N, -1);
  for (int j = 0; j < N; ++j) {
    auto& c = col[j];
    std::sort(c.begin(), c.end());
    for (size_t t = 0; t < c.size(); ++t) {
      if (t > 0 && c[t - 1].first == c[t].first) throw std::runtime_error("duplicate matrix entry");
      const int k = T.Kp[j] + (int)t;
      T.Ki[k] = c[t].first;
      csr_of_K[k] = c[t].second;
      if (c[t].first == j) T.Kdiag[j] = k;
    }
    if (T.Kdiag[j] < 0) throw std::runtime_error("every diagonal entry must be present");
  }
  T.dsign.assign(N, 1);
  for (int i = 0; i < N; ++i) {
    if (dsigns[i] != 1 && dsigns[i] != -1) throw std::runtime_error("dsigns must be +1 or -1");
    T.dsign[i] = (int8_t)dsigns[i];
  }
  finish_kkt_symbolic(T, fp32_order);
  return S;
}

}  // namespace clh
