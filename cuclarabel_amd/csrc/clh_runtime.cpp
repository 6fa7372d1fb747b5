// Runtime of libclarabel_hip: device structures, slot arenas, the single-
// problem solver, the serving loop and the C ABI of include/clarabel_hip.h.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include "../../include/clarabel_hip.h"
#include "clh_backend.h"
#include "clh_symbolic.h"
#include "clh_types.h"

namespace clh {

static thread_local std::string g_last_error;

using Clock = std::chrono::steady_clock;
static double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// ---------------------------------------------------------------------------
// settings
// ---------------------------------------------------------------------------
static void default_settings(clh_settings* s) {  // src/settings.jl:70-148
  const double eps = std::numeric_limits<double>::epsilon();
  std::memset(s, 0, sizeof(*s));
  s->max_iter = 200;
  s->time_limit = std::numeric_limits<double>::infinity();
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
  s->static_regularization_proportional = eps * eps;
  s->dynamic_regularization_enable = 1; s->dynamic_regularization_eps = 1e-13; s->dynamic_regularization_delta = 2e-7;
  s->dynamic_regularization_keep_magnitude = 1;
  s->iterative_refinement_enable = 1; s->iterative_refinement_reltol = 1e-12; s->iterative_refinement_abstol = 1e-12;
  s->iterative_refinement_max_iter = 10; s->iterative_refinement_stop_ratio = 5;
  s->direct_solve_method = CLH_METHOD_HIP;
  s->soc_threshold = 5;
  s->mixed_switch_reltol = 1e-10;
  s->mixed_fp64_static_reg = 1e-8;
  s->serve_engine = CLH_ENGINE_AUTO;
}

static DevSettings to_dev(const clh_settings& s) {
  DevSettings d;
  std::memset(&d, 0, sizeof(d));
  d.max_iter = s.max_iter;
  d.equilibrate_enable = s.equilibrate_enable;
  d.equilibrate_max_iter = s.equilibrate_max_iter;
  d.static_regularization_enable = s.static_regularization_enable;
  d.dynamic_regularization_enable = s.dynamic_regularization_enable;
  d.iterative_refinement_enable = s.iterative_refinement_enable;
  d.iterative_refinement_max_iter = s.iterative_refinement_max_iter;
  d.time_limit = s.time_limit;
  d.max_step_fraction = s.max_step_fraction;
  d.tol_gap_abs = s.tol_gap_abs; d.tol_gap_rel = s.tol_gap_rel; d.tol_feas = s.tol_feas;
  d.tol_infeas_abs = s.tol_infeas_abs; d.tol_infeas_rel = s.tol_infeas_rel; d.tol_ktratio = s.tol_ktratio;
  d.reduced_tol_gap_abs = s.reduced_tol_gap_abs; d.reduced_tol_gap_rel = s.reduced_tol_gap_rel;
  d.reduced_tol_feas = s.reduced_tol_feas; d.reduced_tol_infeas_abs = s.reduced_tol_infeas_abs;
  d.reduced_tol_infeas_rel = s.reduced_tol_infeas_rel; d.reduced_tol_ktratio = s.reduced_tol_ktratio;
  d.equilibrate_min_scaling = s.equilibrate_min_scaling; d.equilibrate_max_scaling = s.equilibrate_max_scaling;
  d.linesearch_backtrack_step = s.linesearch_backtrack_step;
  d.min_switch_step_length = s.min_switch_step_length;
  d.min_terminate_step_length = s.min_terminate_step_length;
  d.static_regularization_constant = s.static_regularization_constant;
  d.static_regularization_proportional = s.static_regularization_proportional;
  d.dynamic_regularization_eps = s.dynamic_regularization_eps;
  d.dynamic_regularization_delta = s.dynamic_regularization_delta;
  d.dyn_keep_magnitude = s.dynamic_regularization_keep_magnitude ? 1 : 0;
  d.iterative_refinement_reltol = s.iterative_refinement_reltol;
  d.iterative_refinement_abstol = s.iterative_refinement_abstol;
  d.iterative_refinement_stop_ratio = s.iterative_refinement_stop_ratio;
  d.mixed_switch_reltol = s.mixed_switch_reltol;
  d.mixed_fp64_static_reg = s.mixed_fp64_static_reg;
  // CLH_GRID_ROWWISE=1: the grid engine's wide supernodes take their updates
  // row by row instead of as tiles (diagnostics)
  const char* rw = std::getenv("CLH_GRID_ROWWISE");
  d.grid_rowwise = rw && rw[0] == '1' ? 1 : 0;
  d.tile_min_cols = s.schedule.tile_min_cols;
  d.dense_min_cols = s.schedule.dense_min_cols;
  return d;
}

// the static regularizer of FP32 factors: sqrt(eps(Float32)) (src/solver.jl:609-611)
static const double kMixedStaticReg = std::sqrt((double)std::numeric_limits<float>::epsilon());

static void check_settings(const clh_settings& s) {
  if (s.direct_solve_method != CLH_METHOD_HIP && s.direct_solve_method != CLH_METHOD_HIPMIXED)
    throw std::invalid_argument("direct_solve_method must be :hip or :hipmixed");
  if (s.soc_threshold != 5) throw std::invalid_argument("soc_threshold other than 5 is not supported");
  if (s.serve_engine < CLH_ENGINE_AUTO || s.serve_engine > CLH_ENGINE_RESIDENT)
    throw std::invalid_argument("serve_engine must be CLH_ENGINE_AUTO, _WORKGROUP, _LANES, _GRID or _RESIDENT");
  if (s.max_iter < 0) throw std::invalid_argument("max_iter must be >= 0");
  if (s.iterative_refinement_max_iter < 0) throw std::invalid_argument("iterative_refinement_max_iter must be >= 0");
  const clh_schedule& c = s.schedule;
  if (c.group_lanes != 0 && c.group_lanes != 16 && c.group_lanes != 32 && c.group_lanes != 64)
    throw std::invalid_argument("schedule.group_lanes must be 0, 16, 32 or 64");
  if (c.blocks_per_cu < 0 || c.grid_blocks < 0 || c.tile_min_cols < 0 || c.dense_min_cols < 0 || c.relax_zero_pct > 100)
    throw std::invalid_argument("schedule fields out of range");
  if ((c.tile_min_cols > 0 && c.tile_min_cols < 16) || (c.dense_min_cols > 0 && c.dense_min_cols < 32))
    throw std::invalid_argument("schedule.tile_min_cols must be >= 16 and dense_min_cols >= 32");
}

// ---------------------------------------------------------------------------
// schedules and the autotuner's cache
// ---------------------------------------------------------------------------
// the symbolic analysis' knobs of a schedule (0 keeps the built-in value)
static SymOpts sym_opts(const clh_schedule& c) {
  SymOpts o;
  if (c.relax_max_cols != 0) o.relax_max_cols = c.relax_max_cols < 0 ? 0 : c.relax_max_cols;
  if (c.relax_zero_pct != 0) o.relax_zero_pct = std::max(0, c.relax_zero_pct);
  o.group_lanes = c.group_lanes;
  return o;
}
// problem-size class: floor(log2(v)) (servers: v = n + m; the grid engine:
// v = factorization flops)
static int size_class(double v) { return v < 2 ? 0 : (int)std::floor(std::log2(v)); }

// The cache of tuned schedules: one text line per (device, engine, method,
// size class): the key, then the seven schedule fields and the default and
// tuned rates the tuner measured.  $CLH_TUNE_CACHE names the file, else
// $HOME/.cache/clarabel_hip/schedules.txt.
static std::string tune_cache_path() {
  if (const char* e = std::getenv("CLH_TUNE_CACHE")) return e;
  const char* home = std::getenv("HOME");
  return std::string(home ? home : ".") + "/.cache/clarabel_hip/schedules.txt";
}
static std::string tune_key(const char* kind, bool fp32, int cls) {
  char name[256] = {0};
  int32_t cus = 0;
  backend::device_info(name, sizeof(name), &cus);
  for (char* c = name; *c; ++c)
    if (*c == ' ' || *c == '\t' || *c == '|') *c = '_';
  return std::string(name) + "|" + std::to_string(cus) + "|" + kind + "|" + (fp32 ? "f32" : "f64") + "|" +
         std::to_string(cls);
}
static std::mutex g_tune_mu;
static std::map<std::string, std::string> tune_cache_read() {
  std::map<std::string, std::string> out;
  FILE* f = std::fopen(tune_cache_path().c_str(), "r");
  if (!f) return out;
  char line[1024];
  while (std::fgets(line, sizeof(line), f)) {
    std::string l(line);
    while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
    const size_t t = l.find('\t');
    if (t != std::string::npos) out[l.substr(0, t)] = l.substr(t + 1);
  }
  std::fclose(f);
  return out;
}
static bool tune_cache_lookup(const std::string& key, clh_schedule* out) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto all = tune_cache_read();
  auto it = all.find(key);
  if (it == all.end()) return false;
  clh_schedule c;
  std::memset(&c, 0, sizeof(c));
  if (std::sscanf(it->second.c_str(), "%d %d %d %d %d %d %d", &c.relax_max_cols, &c.relax_zero_pct, &c.group_lanes,
                  &c.blocks_per_cu, &c.grid_blocks, &c.tile_min_cols, &c.dense_min_cols) != 7)
    return false;
  *out = c;
  return true;
}
static void tune_cache_store(const std::string& key, const clh_schedule& c, double rate0, double rate1) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto all = tune_cache_read();
  char v[256];
  std::snprintf(v, sizeof(v), "%d %d %d %d %d %d %d %.6g %.6g", c.relax_max_cols, c.relax_zero_pct, c.group_lanes,
                c.blocks_per_cu, c.grid_blocks, c.tile_min_cols, c.dense_min_cols, rate0, rate1);
  all[key] = v;
  const std::string path = tune_cache_path();
  for (size_t p = path.find('/', 1); p != std::string::npos; p = path.find('/', p + 1))
    ::mkdir(path.substr(0, p).c_str(), 0755);  // (existing directories are fine)
  const std::string tmp = path + ".tmp" + std::to_string((long)::getpid());
  FILE* f = std::fopen(tmp.c_str(), "w");
  if (!f) return;  // an unwritable cache leaves the tuned schedule in this process only
  for (auto& kv : all) std::fprintf(f, "%s\t%s\n", kv.first.c_str(), kv.second.c_str());
  std::fclose(f);
  std::rename(tmp.c_str(), path.c_str());
}

// ---------------------------------------------------------------------------
// arena layout helper
// ---------------------------------------------------------------------------
struct Layout {
  size_t bytes = 0;
  size_t add(size_t n) {
    size_t off = (bytes + 255) & ~size_t(255);
    bytes = off + n;
    return off;
  }
};

// device copy of one Structure (read-only, shared by all slots of the shape)
struct DevStructure {
  std::shared_ptr<Structure> st;
  char* base = nullptr;
  size_t bytes = 0;
  std::map<std::string, size_t> off;
  template <class T> const T* ptr(const char* k) const {
    auto it = off.find(k);
    return it == off.end() ? nullptr : (const T*)(base + it->second);
  }
  ~DevStructure() { if (base) backend::dev_free(base); }
};

static std::shared_ptr<DevStructure> upload_structure(const std::shared_ptr<Structure>& sp, void* stream) {
  const Structure& S = *sp;
  auto D = std::make_shared<DevStructure>();
  D->st = sp;
  Layout L;
  struct Item { const char* k; const void* p; size_t n; };
  std::vector<Item> items;
  auto I = [&](const char* k, const auto& v) {
    items.push_back({k, v.data(), v.size() * sizeof(v[0])});
  };
  I("Pp", S.Pp); I("Pi", S.Pi); I("Pk", S.Pk);
  I("Ap", S.Ap); I("Ai", S.Ai); I("Atp", S.Atp); I("Ati", S.Ati); I("Ak", S.Ak);
  I("etype", S.etype); I("lin_kdiag", S.lin_kdiag);
  I("soc_off", S.soc_off); I("soc_dim", S.soc_dim); I("soc_kblk", S.soc_kblk); I("soc_sp", S.soc_sp);
  I("soc_kidx", S.soc_kidx); I("sp_ku", S.sp_ku); I("sp_kv", S.sp_kv); I("sp_kD", S.sp_kD); I("sp_off", S.sp_off);
  I("exp_off", S.exp_off); I("exp_kidx", S.exp_kidx); I("pow_off", S.pow_off); I("pow_kidx", S.pow_kidx);
  I("pow_alpha", S.pow_alpha);
  I("psd_off", S.psd_off); I("psd_n", S.psd_n); I("psd_kblk", S.psd_kblk); I("psd_kidx", S.psd_kidx);
  I("psd_soff", S.psd_soff); I("psd_woff", S.psd_woff);
  I("Kdiag", S.Kdiag); I("dsign", S.dsign); I("Kfp", S.Kfp); I("Kfi", S.Kfi); I("Kfk", S.Kfk);
  I("perm", S.perm); I("sn_rows", S.sn_rows); I("lev_ptr", S.lev_ptr); I("lev_sn", S.lev_sn);
  I("lev_group", S.lev_group);
  I("sn", S.sn); I("upd", S.upd); I("upd_rel", S.upd_rel);
  I("sn_lev", S.sn_lev); I("rowupd_ptr", S.rowupd_ptr); I("rowupd", S.rowupd);
  I("asm_dst", S.asm_dst); I("asm_src", S.asm_src); I("asm_sgn", S.asm_sgn);
  for (auto& it : items) D->off[it.k] = L.add(std::max<size_t>(it.n, 8));
  D->bytes = L.bytes;
  D->base = (char*)backend::dev_alloc(L.bytes);
  std::vector<char> host(L.bytes, 0);
  for (auto& it : items)
    if (it.n) std::memcpy(host.data() + D->off[it.k], it.p, it.n);
  backend::h2d(D->base, host.data(), L.bytes, stream);
  backend::sync(stream);
  return D;
}

// Arena of one workgroup slot (lanes = 1) or of one lane batch
// (lanes = kBatchLanes): values, cone state, factors, work vectors, outputs.
// Lane batches interleave every per-problem value array (element i of lane l
// at i * lanes + l) and keep the uploaded values, outputs, trace, timers and
// scalars contiguous per lane.
struct ArenaPlan {
  size_t bytes = 0;
  int lanes = 1;
  // offsets
  size_t vals, Px, Ax, Atx, q, b, vals_end;  // workgroup slot: the upload block
  size_t inbox;                               // lane batch: lanes x nvals uploaded values
  size_t nvals = 0;                           // doubles uploaded per problem
  size_t d, dinv, e, einv, w, lam, soc_eta, soc_d, sp_u, sp_v, nsym, psd_st, psd_wk, Kx, L, Dv, ly, Linv, gscr;
  int group = 1, ngroups = 1;
  size_t vn[10], vm[14], vN[4];
  size_t out, out_lane;                       // lanes x (n + 2m) doubles: x, s, z
  size_t trace, trace_lane, prof, sc, st;
};

// threads: of the engine that runs the slot (0: one workgroup)
static ArenaPlan plan_arena(const Structure& S, int max_iter, int lanes, int threads = 0) {
  ArenaPlan A;
  A.lanes = lanes;
  Layout L;
  const size_t W = (size_t)lanes;
  if (lanes == 1) {
    const int nthr = threads > 0 ? threads : backend::block_threads();
    A.group = nthr >= 64 ? S.group_hint : 1;
    A.ngroups = std::max(1, nthr / A.group);
  }
  // factors are sized for FP64: a mixed-precision slot may move to FP64
  const size_t fs = 8;
  const size_t nP = S.Pp.back(), nA = S.Ai.size();
  A.nvals = nP + 2 * nA + S.n + S.m;
  if (lanes == 1) {  // values uploaded per solve: contiguous, in place
    A.vals = L.add(0);
    A.Px = L.bytes; L.bytes += 8 * nP;
    A.Ax = L.bytes; L.bytes += 8 * nA;
    A.Atx = L.bytes; L.bytes += 8 * nA;
    A.q = L.bytes; L.bytes += 8 * S.n;
    A.b = L.bytes; L.bytes += 8 * S.m;
    A.vals_end = L.bytes;
    A.inbox = 0;
  } else {
    A.inbox = L.add(8 * A.nvals * W);
    A.Px = L.add(8 * nP * W + 8); A.Ax = L.add(8 * nA * W + 8); A.Atx = L.add(8 * nA * W + 8);
    A.q = L.add(8 * S.n * W + 8); A.b = L.add(8 * S.m * W + 8);
    A.vals = A.vals_end = 0;
  }
  A.d = L.add(8 * S.n * W); A.dinv = L.add(8 * S.n * W); A.e = L.add(8 * S.m * W); A.einv = L.add(8 * S.m * W);
  A.w = L.add(8 * S.m * W); A.lam = L.add(8 * S.m * W);
  A.soc_eta = L.add(8 * (S.soc_off.size() + 1) * W); A.soc_d = L.add(8 * (S.soc_off.size() + 1) * W);
  A.sp_u = L.add(8 * (S.sp_ku.size() + 1) * W); A.sp_v = L.add(8 * (S.sp_kv.size() + 1) * W);
  A.nsym = L.add(8 * (24 * (S.exp_off.size() + S.pow_off.size()) + 1) * W);
  A.psd_st = L.add(8 * ((size_t)S.psd_state + 1) * W);
  A.psd_wk = L.add(8 * ((size_t)S.psd_work + 1) * W);
  A.Kx = L.add(8 * S.nnzK * W);
  A.L = L.add(fs * (S.Lsize + 1) * W);
  A.Dv = L.add(fs * S.N * W);
  A.ly = L.add(fs * S.N * W);
  A.Linv = L.add(fs * (S.sn_iptr.back() + 1) * W);
  A.gscr = L.add(fs * (size_t)A.ngroups * S.maxns * W);
  for (auto& o : A.vn) o = L.add(8 * (S.n + 1) * W);
  for (auto& o : A.vm) o = L.add(8 * (S.m + 1) * W);
  for (auto& o : A.vN) o = L.add(8 * S.N * W);
  A.out_lane = (size_t)S.n + 2 * (size_t)S.m;
  A.out = L.add(8 * A.out_lane * W);
  A.trace_lane = 8 * (size_t)(max_iter + 2);
  A.trace = L.add(8 * A.trace_lane * W);
  A.prof = L.add(8 * 8 * W);
  A.sc = L.add(sizeof(SlotScalars) * W);
  A.st = L.add(sizeof(DevSettings));
  A.bytes = (L.bytes + 65535) & ~size_t(65535);
  return A;
}

static SlotDesc make_desc(const DevStructure& DS, const ArenaPlan& A, char* base, bool fp32) {
  const Structure& S = *DS.st;
  SlotDesc D;
  std::memset(&D, 0, sizeof(D));
  D.n = S.n; D.m = S.m; D.p = S.p; D.N = S.N;
  D.nnzP = (int)S.Pp.back(); D.nnzA = (int)S.Ai.size(); D.nnzK = S.nnzK; D.nnzKf = (int)S.Kfi.size();
  D.nsoc = (int)S.soc_off.size(); D.nexp = (int)S.exp_off.size(); D.npow = (int)S.pow_off.size();
  D.nsparse = (int)S.sp_off.size();
  D.degree = S.degree; D.symmetric = S.symmetric ? 1 : 0; D.nsn = S.nsn; D.nlev = S.nlev;
  D.factor_fp32 = fp32 ? 1 : 0;
  D.ngroups = A.ngroups;
  D.nleaf = S.nleaf;
  D.group = A.group;
  D.maxns = S.maxns;
  D.npsd = (int)S.psd_off.size();
  D.inbox_lstride = (int32_t)A.nvals;
  D.out_lstride = (int32_t)A.out_lane;
  D.trace_lstride = (int32_t)A.trace_lane;
  auto dp = [&](size_t o) { return (double*)(base + o); };
  D.Pp = DS.ptr<int32_t>("Pp"); D.Pi = DS.ptr<int32_t>("Pi"); D.Px = dp(A.Px);
  D.Ap = DS.ptr<int32_t>("Ap"); D.Ai = DS.ptr<int32_t>("Ai"); D.Ax = dp(A.Ax);
  D.Atp = DS.ptr<int32_t>("Atp"); D.Ati = DS.ptr<int32_t>("Ati"); D.Atx = dp(A.Atx);
  D.q = dp(A.q); D.b = dp(A.b);
  D.inbox = A.lanes > 1 ? dp(A.inbox) : nullptr;
  D.d = dp(A.d); D.dinv = dp(A.dinv); D.e = dp(A.e); D.einv = dp(A.einv);
  D.etype = DS.ptr<int8_t>("etype"); D.lin_kdiag = DS.ptr<int32_t>("lin_kdiag");
  D.soc_off = DS.ptr<int32_t>("soc_off"); D.soc_dim = DS.ptr<int32_t>("soc_dim");
  D.soc_kblk = DS.ptr<int32_t>("soc_kblk"); D.soc_sp = DS.ptr<int32_t>("soc_sp");
  D.soc_kidx = DS.ptr<int32_t>("soc_kidx");
  D.sp_ku = DS.ptr<int32_t>("sp_ku"); D.sp_kv = DS.ptr<int32_t>("sp_kv"); D.sp_kD = DS.ptr<int32_t>("sp_kD");
  D.sp_off = DS.ptr<int32_t>("sp_off");
  D.exp_off = DS.ptr<int32_t>("exp_off"); D.exp_kidx = DS.ptr<int32_t>("exp_kidx");
  D.pow_off = DS.ptr<int32_t>("pow_off"); D.pow_kidx = DS.ptr<int32_t>("pow_kidx");
  D.pow_alpha = DS.ptr<double>("pow_alpha");
  D.psd_off = DS.ptr<int32_t>("psd_off"); D.psd_n = DS.ptr<int32_t>("psd_n");
  D.psd_kblk = DS.ptr<int32_t>("psd_kblk"); D.psd_kidx = DS.ptr<int32_t>("psd_kidx");
  D.psd_soff = DS.ptr<int32_t>("psd_soff"); D.psd_woff = DS.ptr<int32_t>("psd_woff");
  D.psd_st = dp(A.psd_st); D.psd_wk = dp(A.psd_wk);
  D.w = dp(A.w); D.lam = dp(A.lam); D.soc_eta = dp(A.soc_eta); D.soc_d = dp(A.soc_d);
  D.sp_u = dp(A.sp_u); D.sp_v = dp(A.sp_v); D.nsym = dp(A.nsym);
  D.Kx = dp(A.Kx);
  D.Kdiag = DS.ptr<int32_t>("Kdiag"); D.Pk = DS.ptr<int32_t>("Pk"); D.Ak = DS.ptr<int32_t>("Ak");
  D.dsign = DS.ptr<int8_t>("dsign");
  D.Kfp = DS.ptr<int32_t>("Kfp"); D.Kfi = DS.ptr<int32_t>("Kfi"); D.Kfk = DS.ptr<int32_t>("Kfk");
  D.perm = DS.ptr<int32_t>("perm"); D.sn_rows = DS.ptr<int32_t>("sn_rows");
  D.lev_ptr = DS.ptr<int32_t>("lev_ptr"); D.lev_sn = DS.ptr<int32_t>("lev_sn");
  D.lev_group = DS.ptr<int32_t>("lev_group");
  D.sn = DS.ptr<SnDesc>("sn"); D.upd = DS.ptr<UpdDesc>("upd"); D.upd_rel = DS.ptr<int32_t>("upd_rel");
  D.sn_lev = DS.ptr<SnDesc>("sn_lev"); D.rowupd_ptr = DS.ptr<int32_t>("rowupd_ptr");
  D.rowupd = DS.ptr<RowUpd>("rowupd");
  D.asm_dst = DS.ptr<int32_t>("asm_dst"); D.asm_src = DS.ptr<int32_t>("asm_src");
  D.asm_sgn = DS.ptr<int8_t>("asm_sgn");
  D.L = base + A.L; D.Dv = base + A.Dv; D.ly = base + A.ly;
  D.Linv = base + A.Linv; D.gscr = base + A.gscr;
  double** vn[10] = {&D.x, &D.dx, &D.hx, &D.px, &D.rx, &D.rx_inf, &D.Pxv, &D.x1, &D.x2, &D.workx};
  for (int i = 0; i < 10; ++i) *vn[i] = dp(A.vn[i]);
  double** vm[14] = {&D.s, &D.z, &D.ds, &D.dz, &D.hs, &D.hz, &D.ps, &D.pz, &D.rz, &D.rz_inf, &D.z1, &D.z2,
                     &D.workz, &D.wconic};
  for (int i = 0; i < 14; ++i) *vm[i] = dp(A.vm[i]);
  double** vN[4] = {&D.kb, &D.kx, &D.kw1, &D.kw2};
  for (int i = 0; i < 4; ++i) *vN[i] = dp(A.vN[i]);
  D.out_x = dp(A.out); D.out_s = D.out_x + S.n; D.out_z = D.out_x + S.n + S.m;
  D.trace = dp(A.trace);
  D.prof = (uint64_t*)(base + A.prof);
  D.sc = (SlotScalars*)(base + A.sc);
  D.st = (const DevSettings*)(base + A.st);
  return D;
}

// stage one problem's values in upload order (Px full CSR, Ax CSR, A' CSR, q, b)
static void stage_values(const Structure& S, double* host, const double* Px_user, const double* q,
                         const double* Ax_csc, const double* b) {
  const size_t nP = S.Pp.back(), nA = S.Ai.size();
  double* Px = host;
  double* Ax = Px + nP;
  double* Atx = Ax + nA;
  double* qq = Atx + nA;
  double* bb = qq + S.n;
  for (size_t t = 0; t < S.Puser_keep.size(); ++t) {
    double v = Px_user[S.Puser_keep[t]];
    Px[S.Pfull_pos0[t]] = v;
    if (S.Pfull_pos1[t] >= 0) Px[S.Pfull_pos1[t]] = v;
  }
  for (size_t k = 0; k < nA; ++k) { Atx[k] = Ax_csc[k]; Ax[S.csc_to_csr[k]] = Ax_csc[k]; }
  std::memcpy(qq, q, 8 * S.n);
  std::memcpy(bb, b, 8 * S.m);
}

// shape cache shared by solvers and servers of the process
struct ShapeCache {
  std::mutex mu;
  std::unordered_map<uint64_t, std::shared_ptr<Structure>> host;
  std::shared_ptr<Structure> get(int n, int m, const int32_t* Pp, const int32_t* Pi, const int32_t* Ap,
                                 const int32_t* Ai, const std::vector<ConeSpec>& cones, bool fp32_order,
                                 const SymOpts& opts = SymOpts()) {
    uint64_t key = structure_key(n, m, Pp, Pi, Ap, Ai, cones, fp32_order, opts);
    {
      std::lock_guard<std::mutex> g(mu);
      auto it = host.find(key);
      if (it != host.end()) return it->second;
    }
    auto S = build_structure(n, m, Pp, Pi, Ap, Ai, cones, fp32_order, opts);
    std::lock_guard<std::mutex> g(mu);
    if (host.size() > 4096) host.clear();
    host[key] = S;
    return S;
  }
};
static ShapeCache g_shapes;
// the same shape analysed under another schedule
static std::shared_ptr<Structure> reshape(const Structure& st, const SymOpts& o) {
  return g_shapes.get(st.n, st.m, st.uPp.data(), st.uPi.data(), st.uAp.data(), st.uAi.data(), st.ucones,
                      st.fp32_order, o);
}

static std::vector<ConeSpec> cone_specs(int ncones, const int32_t* ct, const int32_t* cd, const double* cp) {
  std::vector<ConeSpec> v(ncones);
  for (int k = 0; k < ncones; ++k) {
    v[k].type = ct[k];
    v[k].dim = (ct[k] == CT_EXP || ct[k] == CT_POW) ? 3 : (ct[k] == CT_PSD ? cd[k] * (cd[k] + 1) / 2 : cd[k]);
    v[k].side = ct[k] == CT_PSD ? cd[k] : 0;
    v[k].param = (ct[k] == CT_POW && cp) ? cp[k] : 0.0;
  }
  return v;
}

static void fill_info(const SlotScalars& S, const Structure& st, clh_info* info) {
  info->status = S.status;
  info->iterations = S.info_iterations;
  info->obj_val = S.obj_val;
  info->obj_val_dual = S.obj_val_dual;
  info->r_prim = S.res_primal;
  info->r_dual = S.res_dual;
  info->gap_abs = S.gap_abs;
  info->gap_rel = S.gap_rel;
  info->ktratio = S.ktratio;
  info->mu = S.mu;
  info->ir_steps = S.ir_steps;
  info->kkt_solves = S.kkt_solves;
  info->nnz_L = st.nnzL;
  info->factor_flops = st.flops;
  info->fp32_factors = S.n_fp32_factors;
  info->switched_fp64 = S.prec64;
  info->nnz_K = st.nnzK;
  info->kkt_dim = st.N;
  info->n_supernodes = st.nsn;
  info->n_levels = st.nlev;
}

}  // namespace clh

using namespace clh;

// ---------------------------------------------------------------------------
// single-problem solver
// ---------------------------------------------------------------------------
// the single-problem solver moves to the grid engine from this factorization
// cost (CLH_ENGINE_AUTO), over at most this many workgroups
constexpr double kGridMinFlops = 1e8;
constexpr int kGridMaxBlocks = 512;
// factorizations of at least this cost are dominated by the rank-256 updates
// of a dense root, whose tiles wait on global loads: they take up to 1024
// workgroups (4 per CU) to hide that latency.  Measured on MI355X: the random
// sparse SOCP with n = 20 000 (2.2e12 flops) solves in 12.7 s instead of
// 14.3 s, while the 500-cone PSD relaxation (2.4e11) and the exp/pow logistic
// problem (1.6e9) lose 4% and 33% to the dearer grid barriers.
constexpr double kGridWideFlops = 1e12;
constexpr int kGridWideBlocks = 1024;

struct clh_solver {
  clh_settings settings;
  clh_schedule sched;         // the schedule in use (settings.schedule, or as tuned)
  bool fp32 = false;
  std::shared_ptr<Structure> st;
  std::shared_ptr<DevStructure> dst;
  ArenaPlan plan;
  char* arena = nullptr;
  char* host_vals = nullptr;   // staging for uploads
  char* host_out = nullptr;    // staging for downloads
  SlotDesc desc;
  SlotDesc* desc_dev = nullptr;
  int32_t* phase_dev = nullptr;
  std::vector<int32_t> Pp, Pi, Ap, Ai;
  std::vector<double> Px, q, Ax, b;
  bool solved_once = false;
  bool reuse_equil = false;
  SlotScalars last{};
  std::vector<double> x, s, z, trace;
  double setup_time = 0, solve_time = 0;
  uint64_t prof[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  void* stream = nullptr;
  bool grid = false;          // the grid engine: one problem over a cooperative grid
  int grid_nb = 1;
  double* gred = nullptr;     // its reduction partials
  size_t gred_bytes = 0;
  int32_t* abort_word = nullptr;  // host-mapped: ends a grid launch when set
  ~clh_solver() {
    if (abort_word) backend::host_free(abort_word);
    if (gred) backend::dev_free(gred);
    if (arena) backend::dev_free(arena);
    if (desc_dev) backend::dev_free(desc_dev);
    if (phase_dev) backend::dev_free(phase_dev);
    if (host_vals) backend::host_free(host_vals);
    if (host_out) backend::host_free(host_out);
  }
};

#define CLH_TRY(...)                                  \
  try {                                               \
    __VA_ARGS__                                       \
  } catch (const std::invalid_argument& e) {          \
    g_last_error = e.what();                          \
    return -1;                                        \
  } catch (const std::exception& e) {                 \
    g_last_error = e.what();                          \
    return 1;                                         \
  }

// the grid engine's workgroups: the schedule's count, else 512 (1024 from
// 1e12 flops), capped by what the GPU keeps resident (CLH_GRID_BLOCKS caps
// it too: diagnostics)
static int grid_block_rule(const clh_solver* S) {
  int cap = S->sched.grid_blocks > 0 ? S->sched.grid_blocks
                                     : (S->st->flops >= kGridWideFlops ? kGridWideBlocks : kGridMaxBlocks);
  if (const char* e = std::getenv("CLH_GRID_BLOCKS")) cap = std::max(1, std::atoi(e));
  return backend::grid_blocks(cap);
}
// the device settings in the arena (max_iter_override >= 0: a tuning trial's
// iteration cap)
static void solver_write_settings(clh_solver* S, int max_iter_override = -1) {
  clh_settings cs = S->settings;
  cs.schedule = S->sched;
  DevSettings ds = to_dev(cs);
  // gpu_preprocess (src/solver.jl:604-615): the FP32-factor method always
  // takes sqrt(eps(Float32)) as its static regularizer, whatever the settings say
  if (S->fp32) ds.static_regularization_constant = kMixedStaticReg;
  if (max_iter_override >= 0) ds.max_iter = max_iter_override;
  backend::h2d(S->arena + S->plan.st, &ds, sizeof(ds), nullptr);
  backend::sync(nullptr);
}
// arena, descriptor, staging buffers and (grid engine) reduction partials for
// a launch of nb workgroups (the group scratch is sized by the threads)
static void solver_layout(clh_solver* S, int nb) {
  if (S->gred) { backend::dev_free(S->gred); S->gred = nullptr; }
  if (S->arena) { backend::dev_free(S->arena); S->arena = nullptr; }
  if (S->host_vals) { backend::host_free(S->host_vals); S->host_vals = nullptr; }
  if (S->host_out) { backend::host_free(S->host_out); S->host_out = nullptr; }
  int threads = 0;
  if (S->grid) {
    S->grid_nb = nb;
    threads = nb * backend::grid_threads();
    S->gred_bytes = sizeof(double) * 2 * nb + backend::kGridBarBytes;
    S->gred = (double*)backend::dev_alloc(S->gred_bytes);
    const std::vector<char> zero(S->gred_bytes, 0);
    backend::h2d(S->gred, zero.data(), S->gred_bytes, nullptr);
  }
  S->plan = plan_arena(*S->st, S->settings.max_iter, 1, threads);
  S->arena = (char*)backend::dev_alloc(S->plan.bytes);
  S->desc = make_desc(*S->dst, S->plan, S->arena, S->fp32);
  S->desc.abort_flag = S->abort_word;
  S->host_vals = (char*)backend::host_alloc(S->plan.vals_end - S->plan.vals + 8);
  S->host_out = (char*)backend::host_alloc(8 * S->plan.out_lane + 8);
  backend::h2d(S->desc_dev, &S->desc, sizeof(SlotDesc), nullptr);
  solver_write_settings(S);
}

extern "C" {

const char* clh_last_error(void) { return g_last_error.c_str(); }
const char* clh_version(void) { return "clarabel_hip 0.1.0 (Clarabel.jl 0.10.0 hot path)"; }
int clh_device_info(char* name, size_t len, int32_t* num_cu) { return backend::device_info(name, len, num_cu); }
int clh_default_settings(clh_settings* s) {
  if (!s) return -1;
  default_settings(s);
  return 0;
}

int clh_solver_create(clh_solver** out, int32_t n, int32_t m, const int32_t* Pp, const int32_t* Pi, const double* Px,
                      const double* q, const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* b,
                      int32_t ncones, const int32_t* cone_type, const int32_t* cone_dim, const double* cone_param,
                      const clh_settings* settings) {
  CLH_TRY({
    if (!out || n < 0 || m < 0 || !Pp || !Ap || !q || (m > 0 && !b) || ncones < 0)
      throw std::invalid_argument("invalid arguments to clh_solver_create");
    auto t0 = Clock::now();
    std::unique_ptr<clh_solver> S(new clh_solver);
    if (settings) S->settings = *settings;
    else default_settings(&S->settings);
    check_settings(S->settings);
    S->fp32 = S->settings.direct_solve_method == CLH_METHOD_HIPMIXED;
    S->sched = S->settings.schedule;
    auto cones = cone_specs(ncones, cone_type, cone_dim, cone_param);
    try {
      S->st = g_shapes.get(n, m, Pp, Pi, Ap, Ai, cones, S->fp32, sym_opts(S->sched));
    } catch (const std::runtime_error& e) {
      throw std::invalid_argument(e.what());
    }
    S->Pp.assign(Pp, Pp + n + 1); S->Pi.assign(Pi, Pi + Pp[n]); S->Px.assign(Px, Px + Pp[n]);
    S->Ap.assign(Ap, Ap + n + 1); S->Ai.assign(Ai, Ai + Ap[n]); S->Ax.assign(Ax, Ax + Ap[n]);
    S->q.assign(q, q + n); S->b.assign(b, b + m);
    // (CLH_ENGINE_LANES is a serving choice: a single solve runs on one workgroup then)
    S->grid = S->settings.serve_engine == CLH_ENGINE_GRID ||
              (S->settings.serve_engine == CLH_ENGINE_AUTO && S->st->flops >= kGridMinFlops);
    if (S->grid && S->sched.use_tuned) {  // the grid engine's tuned launch for this size class
      clh_schedule t;
      if (tune_cache_lookup(tune_key("grid", S->fp32, size_class(S->st->flops)), &t)) {
        S->sched.grid_blocks = t.grid_blocks;
        S->sched.tile_min_cols = t.tile_min_cols;
        S->sched.dense_min_cols = t.dense_min_cols;
      }
    }
    S->dst = upload_structure(S->st, nullptr);
    S->desc_dev = (SlotDesc*)backend::dev_alloc(sizeof(SlotDesc));
    S->phase_dev = (int32_t*)backend::dev_alloc(256);
    if (S->grid) {
      S->abort_word = (int32_t*)backend::host_alloc_coherent(sizeof(int32_t));
      *(volatile int32_t*)S->abort_word = 0;
    }
    solver_layout(S.get(), S->grid ? grid_block_rule(S.get()) : 0);
    S->setup_time = seconds_since(t0);
    *out = S.release();
    return 0;
  })
}

int clh_solve(clh_solver* S, void* stream) {
  CLH_TRY({
    if (!S) throw std::invalid_argument("null solver");
    auto t0 = Clock::now();
    const Structure& st = *S->st;
    stage_values(st, (double*)S->host_vals, S->Px.data(), S->q.data(), S->Ax.data(), S->b.data());
    backend::h2d(S->arena + S->plan.vals, S->host_vals, S->plan.vals_end - S->plan.vals, stream);
    SlotScalars sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.phase = PH_SETUP;
    // info.total_time > time_limit (src/info.jl:182-183): the total counts
    // setup! on the first solve after setup only (info_reset!, :242-260)
    sc.time_left = S->settings.time_limit - (S->solved_once ? 0.0 : S->setup_time);
    if (S->reuse_equil) {
      sc.reuse_equil = 1;
      sc.c = S->last.c;
    }
    backend::h2d(S->arena + S->plan.sc, &sc, sizeof(sc), stream);
    static const uint64_t zeros[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    backend::h2d(S->arena + S->plan.prof, zeros, sizeof(zeros), stream);
    const int passes = std::max(1, S->settings.max_iter + 8);
    int32_t phase = -1;
    // the host's watchdog on a grid launch: with a finite time limit, a launch
    // still running a minute past it is aborted (the device checks the limit
    // only between iterations); CLH_ABORT_AFTER_S sets another delay
    // (diagnostics: 0 aborts before the launch, which tests the abort path)
    double watchdog = S->settings.time_limit + 60.0;
    if (const char* e = std::getenv("CLH_ABORT_AFTER_S")) watchdog = std::atof(e);
    int round = 0;
    for (; round < 64; ++round) {
      if (S->grid) {
        *(volatile int32_t*)S->abort_word = watchdog <= 0 ? 1 : 0;
        backend::launch_grid(S->desc_dev, passes, S->phase_dev, S->gred, S->grid_nb, stream);
        backend::d2h(&phase, S->phase_dev, sizeof(int32_t), stream);
        while (!backend::stream_idle(stream)) {
          if (seconds_since(t0) > watchdog) *(volatile int32_t*)S->abort_word = 1;
          std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
        if (*(volatile int32_t*)S->abort_word) {
          // the barrier counters and partials are left in no usable state
          const std::vector<char> zero(S->gred_bytes, 0);
          backend::h2d(S->gred, zero.data(), S->gred_bytes, stream);
          backend::sync(stream);
          *(volatile int32_t*)S->abort_word = 0;
          throw std::runtime_error(watchdog <= 0 || seconds_since(t0) > watchdog
                                       ? "grid launch aborted by the host (time limit exceeded by its margin)"
                                       : "grid launch aborted: a grid barrier waited 30 s for a workgroup");
        }
      } else {
        backend::launch_ipm(backend::LaunchSet{S->desc_dev, 1, S->phase_dev, nullptr, 0, nullptr}, passes, stream);
        backend::d2h(&phase, S->phase_dev, sizeof(int32_t), stream);
        backend::sync(stream);
      }
      if (phase == PH_DONE) break;
    }
    if (phase != PH_DONE)
      throw std::runtime_error("solve did not finish in " + std::to_string(round) + " launches of " +
                               std::to_string(passes) + " passes (phase " + std::to_string(phase) + ")");
    backend::d2h(&S->last, S->arena + S->plan.sc, sizeof(SlotScalars), stream);
    backend::d2h(S->prof, S->arena + S->plan.prof, sizeof(S->prof), stream);
    backend::d2h(S->host_out, S->arena + S->plan.out, 8 * S->plan.out_lane, stream);
    S->trace.assign(8 * (size_t)(S->settings.max_iter + 2), 0.0);
    backend::d2h(S->trace.data(), S->arena + S->plan.trace, 8 * S->trace.size(), stream);
    backend::sync(stream);
    const double* ho = (const double*)S->host_out;
    S->x.assign(ho, ho + st.n);
    S->s.assign(ho + st.n, ho + st.n + st.m);
    S->z.assign(ho + st.n + st.m, ho + st.n + 2 * st.m);
    S->solve_time = seconds_since(t0);
    S->solved_once = true;
    return 0;
  })
}

int clh_get_solution(clh_solver* S, double* x, double* sv, double* z, clh_info* info) {
  if (!S || !S->solved_once) { g_last_error = "no solution available"; return -1; }
  if (x) std::memcpy(x, S->x.data(), 8 * S->x.size());
  if (sv) std::memcpy(sv, S->s.data(), 8 * S->s.size());
  if (z) std::memcpy(z, S->z.data(), 8 * S->z.size());
  if (info) {
    std::memset(info, 0, sizeof(*info));
    fill_info(S->last, *S->st, info);
    info->solve_time = S->solve_time;
    info->setup_time = S->setup_time;
    for (int k = 0; k < 8; ++k) info->phase_seconds[k] = S->prof[k] * 1e-8;  // 100 MHz device clock
  }
  return 0;
}

int clh_get_trace(clh_solver* S, double* trace, int32_t max_rows, int32_t* rows) {
  if (!S || !S->solved_once) { g_last_error = "no solution available"; return -1; }
  int r = std::min<int>(S->last.trace_len, S->settings.max_iter + 2);
  r = std::min(r, (int)max_rows);
  if (trace) std::memcpy(trace, S->trace.data(), 8 * 8 * (size_t)r);
  if (rows) *rows = r;
  return 0;
}

int clh_update_P(clh_solver* S, const double* Px) {
  if (!S || !Px) return -1;
  std::memcpy(S->Px.data(), Px, 8 * S->Px.size());
  S->reuse_equil = S->solved_once;
  return 0;
}
int clh_update_q(clh_solver* S, const double* q) {
  if (!S || !q) return -1;
  std::memcpy(S->q.data(), q, 8 * S->q.size());
  S->reuse_equil = S->solved_once;
  return 0;
}
int clh_update_A(clh_solver* S, const double* Ax) {
  if (!S || !Ax) return -1;
  std::memcpy(S->Ax.data(), Ax, 8 * S->Ax.size());
  S->reuse_equil = S->solved_once;
  return 0;
}
int clh_update_b(clh_solver* S, const double* b) {
  if (!S || !b) return -1;
  std::memcpy(S->b.data(), b, 8 * S->b.size());
  S->reuse_equil = S->solved_once;
  return 0;
}

// ---- device-pointer variants (the counterpart of the reference's Python
// bridge, src/python/py2jl.jl:1-26 / jl2py.py: CuPy arrays handed to the
// solver and its solution handed back as device arrays).  The problem's
// arrays come from device memory: the patterns are brought to the host for
// the structure analysis (which runs there, as cuDSS's analysis does), the
// values are kept as the host path keeps them and uploaded by each solve.
int clh_solver_create_device(clh_solver** out, int32_t n, int32_t m, const int32_t* Pp_dev, const int32_t* Pi_dev,
                             const double* Px_dev, const double* q_dev, const int32_t* Ap_dev, const int32_t* Ai_dev,
                             const double* Ax_dev, const double* b_dev, int32_t ncones, const int32_t* cone_type,
                             const int32_t* cone_dim, const double* cone_param, const clh_settings* settings) {
  CLH_TRY({
    if (!out || n < 0 || m < 0 || !Pp_dev || !Ap_dev || !q_dev || (m > 0 && !b_dev))
      throw std::invalid_argument("invalid arguments to clh_solver_create_device");
    std::vector<int32_t> Pp(n + 1), Ap(n + 1);
    backend::d2h(Pp.data(), Pp_dev, 4 * (size_t)(n + 1), nullptr);
    backend::d2h(Ap.data(), Ap_dev, 4 * (size_t)(n + 1), nullptr);
    backend::sync(nullptr);
    if (Pp[0] != 0 || Ap[0] != 0 || Pp[n] < 0 || Ap[n] < 0) throw std::invalid_argument("invalid column pointers");
    std::vector<int32_t> Pi(Pp[n]), Ai(Ap[n]);
    std::vector<double> Px(Pp[n]), Ax(Ap[n]), q(n), b(m);
    if (Pp[n]) { backend::d2h(Pi.data(), Pi_dev, 4 * (size_t)Pp[n], nullptr); backend::d2h(Px.data(), Px_dev, 8 * (size_t)Pp[n], nullptr); }
    if (Ap[n]) { backend::d2h(Ai.data(), Ai_dev, 4 * (size_t)Ap[n], nullptr); backend::d2h(Ax.data(), Ax_dev, 8 * (size_t)Ap[n], nullptr); }
    if (n) backend::d2h(q.data(), q_dev, 8 * (size_t)n, nullptr);
    if (m) backend::d2h(b.data(), b_dev, 8 * (size_t)m, nullptr);
    backend::sync(nullptr);
    return clh_solver_create(out, n, m, Pp.data(), Pi.data(), Px.data(), q.data(), Ap.data(), Ai.data(), Ax.data(),
                             b.data(), ncones, cone_type, cone_dim, cone_param, settings);
  })
}

// the last solution straight from the solver's arena into device buffers
// (device to device; NULL skips one)
int clh_get_solution_device(clh_solver* S, double* x_dev, double* s_dev, double* z_dev, void* stream) {
  CLH_TRY({
    if (!S) throw std::invalid_argument("null solver");
    if (!S->solved_once) throw std::invalid_argument("no solution available");
    const size_t n = S->st->n, m = S->st->m;
    const char* o = S->arena + S->plan.out;
    if (x_dev && n) backend::d2d(x_dev, o, 8 * n, stream);
    if (s_dev && m) backend::d2d(s_dev, o + 8 * n, 8 * m, stream);
    if (z_dev && m) backend::d2d(z_dev, o + 8 * (n + m), 8 * m, stream);
    backend::sync(stream);
    return 0;
  })
}

// update_P!/q!/A!/b! from device memory
static int update_from_device(clh_solver* S, std::vector<double>& dst, const double* src_dev) {
  CLH_TRY({
    if (!S || !src_dev) throw std::invalid_argument("null argument");
    if (!dst.empty()) backend::d2h(dst.data(), src_dev, 8 * dst.size(), nullptr);
    backend::sync(nullptr);
    S->reuse_equil = S->solved_once;
    return 0;
  })
}
int clh_update_P_device(clh_solver* S, const double* Px_dev) { return S ? update_from_device(S, S->Px, Px_dev) : -1; }
int clh_update_q_device(clh_solver* S, const double* q_dev) { return S ? update_from_device(S, S->q, q_dev) : -1; }
int clh_update_A_device(clh_solver* S, const double* Ax_dev) { return S ? update_from_device(S, S->Ax, Ax_dev) : -1; }
int clh_update_b_device(clh_solver* S, const double* b_dev) { return S ? update_from_device(S, S->b, b_dev) : -1; }

// The grid engine's schedule by coordinate descent over the workgroup count
// and the tile and dense thresholds, each candidate timed over a solve
// capped at a few iterations (the first, untimed, warms the code and the
// caches); a candidate replaces the best so far when it is 3% faster (the
// spread of repeated trials), and the final choice is timed again beside the
// starting schedule, which it must beat.  The winner stays in the solver and
// goes to the cache.  Before the first solve only: the trials overwrite what a warm
// re-solve reuses.
int clh_solver_autotune(clh_solver* S, double max_seconds, clh_tune_report* report) {
  CLH_TRY({
    if (!S) throw std::invalid_argument("null solver");
    if (S->solved_once) throw std::invalid_argument("clh_solver_autotune must run before the first clh_solve");
    clh_tune_report R;
    std::memset(&R, 0, sizeof(R));
    const auto t0 = Clock::now();
    if (S->grid) {
      constexpr int kTrialIters = 3;
      auto trial = [&]() {  // IPM iterations per second
        solver_write_settings(S, kTrialIters);
        const auto t = Clock::now();
        if (clh_solve(S, nullptr) != 0) throw std::runtime_error("autotune trial: " + g_last_error);
        const double dt = seconds_since(t);
        R.trials++;
        const double rate = (S->last.iter + 1) / std::max(dt, 1e-9);
        if (S->settings.verbose) {
          std::fprintf(stderr, "[clh autotune] trial %d: %d workgroups, tiles from %d, dense from %d columns: "
                       "%.2f iterations/s (%.1f s in)\n", R.trials, S->grid_nb, S->sched.tile_min_cols,
                       S->sched.dense_min_cols, rate, seconds_since(t0));
          std::fflush(stderr);
        }
        return rate;
      };
      auto apply = [&](const clh_schedule& c) {
        const int nb_old = S->grid_nb;
        S->sched = c;
        const int nb = grid_block_rule(S);
        if (nb != nb_old) solver_layout(S, nb);
      };
      clh_schedule best = S->sched;
      apply(best);
      trial();  // warm-up
      double rbest = trial();
      R.default_rate = rbest;
      const int nbmax = backend::grid_blocks(1 << 20);
      const std::vector<std::pair<int32_t clh_schedule::*, std::vector<int>>> knobs = {
          {&clh_schedule::grid_blocks, {256, 512, 768, 1024}},
          {&clh_schedule::tile_min_cols, {64, 128, 256}},
          {&clh_schedule::dense_min_cols, {64, 128, 256}}};
      best.grid_blocks = S->grid_nb;
      best.tile_min_cols = best.tile_min_cols > 0 ? best.tile_min_cols : 128;
      best.dense_min_cols = best.dense_min_cols > 0 ? best.dense_min_cols : 128;
      const clh_schedule start = best;
      for (const auto& kv : knobs)
        for (int v : kv.second) {
          if (seconds_since(t0) > max_seconds) break;
          if (best.*(kv.first) == v || (kv.first == &clh_schedule::grid_blocks && v > nbmax)) continue;
          clh_schedule c = best;
          c.*(kv.first) = v;
          apply(c);
          const double r = trial();
          if (r > 1.03 * rbest) { best = c; rbest = r; }
        }
      R.tuned_rate = R.default_rate;
      if (std::memcmp(&best, &start, sizeof(best)) != 0) {  // side by side with the start: never slower
        apply(start);
        const double r0 = trial();
        apply(best);
        const double r1 = trial();
        R.default_rate = 0.5 * (R.default_rate + r0);
        R.tuned_rate = r1;
        if (r1 <= R.default_rate) {
          best = start;
          R.tuned_rate = R.default_rate;
        }
      }
      apply(best);
      best.grid_blocks = S->grid_nb;
      S->sched = best;
      R.classes = 1;
      tune_cache_store(tune_key("grid", S->fp32, size_class(S->st->flops)), best, R.default_rate, R.tuned_rate);
      // the solver as created: no solve yet
      solver_write_settings(S);
      S->solved_once = false;
      S->reuse_equil = false;
      std::memset(&S->last, 0, sizeof(S->last));
    }
    R.seconds = seconds_since(t0);
    if (report) *report = R;
    return 0;
  })
}

int clh_solver_get_schedule(clh_solver* S, clh_schedule* out) {
  if (!S || !out) { g_last_error = "null argument"; return -1; }
  *out = S->sched;
  if (S->grid) out->grid_blocks = S->grid_nb;
  return 0;
}

void clh_solver_destroy(clh_solver* S) { delete S; }

}  // extern "C"

// ---------------------------------------------------------------------------
// serving
// ---------------------------------------------------------------------------
namespace {

struct Request {
  int64_t ticket;
  std::shared_ptr<Structure> st;
  std::vector<double> Px, q, Ax, b;
  Clock::time_point t_submit;
  bool lanes = false;  // runs on the lane engine
};

struct Result {
  std::vector<double> x, s, z;
  clh_info info;
  double latency;
};

// one request on the workgroup engine
struct SlotRec {
  bool busy = false;
  int64_t ticket = -1;
  char* arena = nullptr;
  size_t bytes = 0;
  char* outbox = nullptr;  // host-mapped: x, s, z, then the final SlotScalars
  size_t outbox_bytes = 0;
  std::shared_ptr<DevStructure> dst;
  ArenaPlan plan;
  Clock::time_point t_submit, t_admit;
  // the device-resident loop: the request's inbox record and cancel word
  // (coherent host memory) and its sequence number in the ring
  char* inbox = nullptr;
  size_t inbox_bytes = 0;
  int32_t* cancel = nullptr;
  uint64_t seq = 0;
};

// up to kBatchLanes requests of one shape on the lane engine
struct BatchRec {
  bool live = false;
  uint64_t key = 0;
  int nbusy = 0;
  char* arena = nullptr;
  size_t bytes = 0;
  std::shared_ptr<DevStructure> dst;
  ArenaPlan plan;
  bool busy[kBatchLanes];
  int64_t ticket[kBatchLanes];
  Clock::time_point t_submit[kBatchLanes], t_admit[kBatchLanes];
};

// where a finished request's outputs were copied to in the harvest buffer
struct Pending {
  int64_t ticket;
  const Structure* st;
  size_t off_out, off_sc;
  Clock::time_point t_submit, t_admit;
};

}  // namespace

struct clh_server {
  clh_settings settings;
  clh_schedule base;                    // settings.schedule, with the tuned workgroups per CU
  std::map<int, clh_schedule> cls;      // tuned schedules by size class of n + m
  int64_t trial_ticket = 0;             // autotuning trials use negative tickets
  DevSettings dsettings;
  bool fp32 = false;
  int64_t budget = 0, in_use = 0;
  int max_slots = 0, max_batches = 0, busy_requests = 0;
  double lane_max_flops = 0;  // auto: the lane engine only where set by CLH_LANE_MAX_FLOPS
  void* stream = nullptr;
  bool own_stream = false;
  std::vector<SlotRec> slots;
  std::vector<BatchRec> batches;
  SlotDesc* descs = nullptr;     // pinned: max_slots workgroup slots, then max_batches lane batches
  size_t ndescs = 0;
  SlotDesc* descs_dev = nullptr;
  DevSettings* settings_dev = nullptr;  // one copy for every workgroup slot
  std::vector<int> fresh;        // slots whose descriptor points at this round's inbox
  std::multimap<size_t, char*> free_outboxes;
  int32_t* phase_dev = nullptr;
  int32_t* phase_host = nullptr;
  bool descs_dirty = true;
  bool pending = false;          // a launch is running (progress without wait)
  void* graph = nullptr;
  int graph_passes = -1, graph_ns = -1, graph_nb = -1;
  std::deque<Request> queue;
  std::unordered_map<int64_t, Result> done;
  std::unordered_map<uint64_t, std::shared_ptr<DevStructure>> dev_shapes;
  std::multimap<size_t, char*> free_arenas;
  int64_t pooled_bytes = 0;
  std::vector<char*> deferred_free;  // arenas beyond the pool's cap, freed when no launch runs
  int64_t next_ticket = 1, submitted = 0, completed = 0, rejected = 0, launches = 0, cancelled = 0;
  std::vector<double> latencies;
  double busy_time = 0, kernel_time = 0;
  char* stage = nullptr;
  size_t stage_bytes = 0;
  char* outbuf = nullptr;
  size_t outbuf_bytes = 0;
  // the device-resident request loop (CLH_ENGINE_RESIDENT): a ring of
  // max_slots descriptors with done words and the control words in coherent
  // host memory, the claim counter on the device, structure uploads on a
  // stream of their own (the persistent launch holds `stream`)
  struct ResCtrl {
    uint64_t head;
    uint32_t stop, pad;
  };
  bool resident = false;
  ResCtrl* rctrl = nullptr;
  SlotDesc* rring = nullptr;
  uint64_t* rdone = nullptr;
  uint64_t* rtail = nullptr;
  uint64_t rseq = 0;
  int rblocks = 0;
  bool rlaunched = false;
  void* rtimer = nullptr;
  void* upload_stream = nullptr;
  std::multimap<size_t, char*> free_coherent;

  ~clh_server() {
    if (rctrl) __atomic_store_n(&rctrl->stop, 1u, __ATOMIC_RELEASE);  // the persistent launch leaves
    try {
      backend::sync(stream);  // a destructor must not throw (a failed create lands here too)
    } catch (const std::exception&) {
    }
    pending = false;
    if (graph) backend::graph_destroy(graph);
    for (auto& s : slots) if (s.arena) backend::dev_free(s.arena);
    for (auto& b : batches) if (b.arena) backend::dev_free(b.arena);
    for (auto& f : free_arenas) backend::dev_free(f.second);
    for (char* p : deferred_free) backend::dev_free(p);
    for (auto& s : slots) if (s.outbox) backend::host_free(s.outbox);
    for (auto& f : free_outboxes) backend::host_free(f.second);
    if (descs) backend::host_free(descs);
    if (settings_dev) backend::dev_free(settings_dev);
    if (descs_dev) backend::dev_free(descs_dev);
    if (phase_dev) backend::dev_free(phase_dev);
    if (phase_host) backend::host_free(phase_host);
    if (stage) backend::host_free(stage);
    if (outbuf) backend::host_free(outbuf);
    for (auto& s : slots) {
      if (s.inbox) backend::host_free(s.inbox);
      if (s.cancel) backend::host_free(s.cancel);
    }
    for (auto& f : free_coherent) backend::host_free(f.second);
    if (rctrl) backend::host_free(rctrl);
    if (rring) backend::host_free(rring);
    if (rdone) backend::host_free(rdone);
    if (rtail) backend::dev_free(rtail);
    if (upload_stream) backend::stream_destroy(upload_stream);
    if (rtimer) backend::timer_destroy(rtimer);
    if (own_stream) backend::stream_destroy(stream);
  }

  char* get_coherent(size_t bytes) {
    auto it = free_coherent.find(bytes);
    if (it != free_coherent.end()) {
      char* p = it->second;
      free_coherent.erase(it);
      return p;
    }
    return (char*)backend::host_alloc_coherent(bytes);
  }

  // the schedule of a problem with n + m rows and columns
  clh_schedule sched_for(int n_plus_m) const {
    auto it = cls.find(size_class(n_plus_m));
    return it != cls.end() ? it->second : base;
  }
  // the persistent launch's workgroups: blocks_per_cu per CU, at most what fits
  int resident_block_count() const {
    const int mx = backend::resident_blocks();
    if (base.blocks_per_cu <= 0) return mx;
    int32_t cus = 1;
    backend::device_info(nullptr, 0, &cus);
    return std::max(1, std::min(mx, base.blocks_per_cu * std::max(1, (int)cus)));
  }

  // Autotuning on the queued requests (they stay queued; the trials serve
  // copies and drop the results).  Every trial serves the whole sample, all
  // submitted at once, and is timed as a whole (the rate a saturated server
  // gives, its ramp and tail included): per size class, the relaxation and
  // group knobs by coordinate descent with the other classes on their
  // current schedules; then the resident workgroups per CU.  A candidate
  // replaces the best so far when it is 3% faster (the spread of repeated
  // runs on the MPC stream).  The sample should hold several times as many
  // requests as the server keeps resident, or the ramp and tail dominate.
  clh_tune_report autotune(double max_seconds) {
    if (busy_requests > 0 || pending)
      throw std::invalid_argument("clh_server_autotune needs an idle server (requests are in flight)");
    if (queue.empty()) throw std::invalid_argument("clh_server_autotune tunes on queued requests: submit a sample first");
    clh_tune_report R;
    std::memset(&R, 0, sizeof(R));
    const auto t0 = Clock::now();
    std::deque<Request> real;
    real.swap(queue);
    const int64_t s_completed = completed, s_launches = launches;
    const double s_busy = busy_time, s_kernel = kernel_time;
    const size_t s_lat = latencies.size();
    std::map<int, std::vector<const Request*>> by;
    std::vector<const Request*> all;
    for (const Request& r : real) {
      by[size_class(r.st->n + r.st->m)].push_back(&r);
      all.push_back(&r);
    }
    std::map<int, clh_schedule> cur;
    for (auto& kv : by) cur[kv.first] = sched_for(kv.second.front()->st->n + kv.second.front()->st->m);
    std::vector<int> tuned;  // the classes with enough requests to time
    auto serve_trial = [&](const std::vector<const Request*>& rs, const std::map<int, clh_schedule>& sch) {
      std::vector<Request> copies;
      copies.reserve(rs.size());
      std::map<size_t, int> sizes;  // arenas by size the trial can hold at once
      std::set<uint64_t> keys;
      for (const Request* r : rs) {
        Request c = *r;
        c.st = reshape(*r->st, sym_opts(sch.at(size_class(r->st->n + r->st->m))));
        c.ticket = --trial_ticket;
        keys.insert(c.st->key);
        int& k = sizes[plan_arena(*c.st, settings.max_iter, 1).bytes];
        k = std::min(k + 1, max_slots);
        copies.push_back(std::move(c));
      }
      // before the clock: the candidate's structures analysed and uploaded
      // (the others' device copies dropped) and its arenas allocated (the
      // others' freed), so every trial times the serving alone and memory
      // stays at one candidate's worth however many are tried
      for (auto it = dev_shapes.begin(); it != dev_shapes.end();)
        it = keys.count(it->first) ? std::next(it) : dev_shapes.erase(it);
      for (const Request& c : copies) shape_on_device(c.st);
      shape_pool(sizes);
      rblocks = resident_block_count();
      const auto t = Clock::now();
      for (Request& c : copies) {
        c.t_submit = t;
        queue.push_back(std::move(c));
      }
      while (progress(4, true) > 0) {
      }
      const double dt = seconds_since(t);
      for (auto it = done.begin(); it != done.end();) it = it->first < 0 ? done.erase(it) : std::next(it);
      R.trials++;
      const double rate = (double)rs.size() / std::max(dt, 1e-9);
      if (settings.verbose) {  // one line per trial: what the tuner is doing, as it does it
        std::fprintf(stderr, "[clh autotune] trial %d: %zu requests, %.1f solves/s, %d workgroups/CU (%.1f s in)\n",
                     R.trials, rs.size(), rate, base.blocks_per_cu, seconds_since(t0));
        std::fflush(stderr);
      }
      return rate;
    };
    try {
      serve_trial(all, cur);  // warm-up: arenas, code, caches
      const std::map<int, clh_schedule> start = cur;
      const int bpc_start = base.blocks_per_cu;
      int bpc_cur = bpc_start;
      double rbest = serve_trial(all, cur);
      R.default_rate = rbest;
      // a candidate that times 3% faster than the incumbent is timed again
      // beside a new run of the incumbent, and taken when the mean of its two
      // runs beats the incumbent's by 2% (single runs spread by about 3%)
      auto consider = [&](const std::map<int, clh_schedule>& cand, int bpc_cand) {
        base.blocks_per_cu = bpc_cand;
        const double r = serve_trial(all, cand);
        if (r > 1.03 * rbest) {
          const double r2 = serve_trial(all, cand);
          base.blocks_per_cu = bpc_cur;
          const double i2 = serve_trial(all, cur);
          const double rc = 0.5 * (r + r2), ri = 0.5 * (rbest + i2);
          if (rc > 1.02 * ri) {
            cur = cand;
            bpc_cur = bpc_cand;
            rbest = rc;
          } else {
            rbest = ri;
          }
        }
        base.blocks_per_cu = bpc_cur;
      };
      const std::vector<std::pair<int32_t clh_schedule::*, std::vector<int>>> knobs = {
          {&clh_schedule::relax_max_cols, {128, 32, 64, 256, -1}},
          {&clh_schedule::relax_zero_pct, {25, 10, 50}},
          {&clh_schedule::group_lanes, {0, 16, 32, 64}}};
      for (auto& kv : by) {
        if (kv.second.size() < 4 || seconds_since(t0) > max_seconds) continue;
        R.classes++;
        tuned.push_back(kv.first);
        clh_schedule& b = cur[kv.first];
        if (b.relax_max_cols == 0) b.relax_max_cols = 128;
        if (b.relax_zero_pct == 0) b.relax_zero_pct = 25;
        for (const auto& knob : knobs)
          for (int v : knob.second) {
            if (seconds_since(t0) > max_seconds) break;
            if (cur[kv.first].*(knob.first) == v) continue;
            std::map<int, clh_schedule> cand = cur;
            cand[kv.first].*(knob.first) = v;
            consider(cand, bpc_cur);
          }
      }
      if (resident) {  // workgroups per CU, each class on its schedule
        int32_t cus = 1;
        backend::device_info(nullptr, 0, &cus);
        const int per_cu_max = std::max(1, backend::resident_blocks() / std::max(1, (int)cus));
        for (int v = per_cu_max; v >= 1 && seconds_since(t0) <= max_seconds; --v) {
          if ((bpc_cur == 0 && v == per_cu_max) || v == bpc_cur) continue;
          consider(cur, v);
        }
      }
      // the choice against the starting schedule, side by side: the tuner
      // never leaves a server slower than it found it
      bool changed = bpc_cur != bpc_start;
      for (auto& kv : cur) {
        clh_schedule a = kv.second, z = start.at(kv.first);
        if (a.relax_max_cols == 128) a.relax_max_cols = 0;
        if (a.relax_zero_pct == 25) a.relax_zero_pct = 0;
        if (z.relax_max_cols == 128) z.relax_max_cols = 0;
        if (z.relax_zero_pct == 25) z.relax_zero_pct = 0;
        a.blocks_per_cu = z.blocks_per_cu = 0;
        if (std::memcmp(&a, &z, sizeof(a)) != 0) changed = true;
      }
      R.tuned_rate = R.default_rate;
      if (changed) {
        base.blocks_per_cu = bpc_start;
        const double r0 = serve_trial(all, start);
        base.blocks_per_cu = bpc_cur;
        const double r1 = serve_trial(all, cur);
        R.default_rate = 0.5 * (R.default_rate + r0);
        R.tuned_rate = r1;
        if (r1 <= R.default_rate) {
          cur = start;
          bpc_cur = bpc_start;
          R.tuned_rate = R.default_rate;
        }
      }
      base.blocks_per_cu = bpc_cur;
      for (auto& kv : cur) kv.second.blocks_per_cu = bpc_cur;
    } catch (...) {
      queue.swap(real);
      throw;
    }
    for (auto& kv : cur) cls[kv.first] = kv.second;
    for (int c : tuned) tune_cache_store(tune_key("serve", fp32, c), cur[c], R.default_rate, R.tuned_rate);
    // the queued requests, analysed under their class's schedule
    for (Request& r : real) r.st = reshape(*r.st, sym_opts(sched_for(r.st->n + r.st->m)));
    queue.swap(real);
    completed = s_completed;
    launches = s_launches;
    busy_time = s_busy;
    kernel_time = s_kernel;
    latencies.resize(s_lat);
    R.seconds = seconds_since(t0);
    return R;
  }

  // ---- the device-resident request loop ------------------------------------
  // Admit queued requests while slots and the budget allow: each gets its
  // arena, an inbox record, an outbox and a cancel word, and its descriptor
  // goes to the ring; head is published once for the round.
  int admit_resident() {
    int admitted = 0;
    const auto now = Clock::now();
    while (!queue.empty() && busy_requests < max_slots) {
      Request& r = queue.front();
      int slot = -1;
      for (int i = 0; i < max_slots; ++i)
        if (!slots[i].busy) { slot = i; break; }
      if (slot < 0) break;
      ArenaPlan plan = plan_arena(*r.st, settings.max_iter, 1);
      if (in_use + (int64_t)plan.bytes > budget) {
        if (in_use == 0 && admitted == 0) { reject_front(); continue; }
        break;
      }
      SlotRec& s = slots[slot];
      s.busy = true;
      s.ticket = r.ticket;
      s.bytes = plan.bytes;
      s.arena = get_arena(plan.bytes);
      s.dst = shape_on_device(r.st);
      s.plan = plan;
      s.t_submit = r.t_submit;
      s.t_admit = now;
      in_use += plan.bytes;
      s.outbox_bytes = ((8 * plan.out_lane + 63) & ~size_t(63)) + sizeof(SlotScalars);
      s.outbox = get_coherent(s.outbox_bytes);
      s.inbox_bytes = 8 * (kInboxHeaderDoubles + plan.nvals);
      s.inbox = get_coherent(s.inbox_bytes);
      if (!s.cancel) s.cancel = (int32_t*)backend::host_alloc_coherent(64);
      *(volatile int32_t*)s.cancel = 0;
      SlotScalars* sc = (SlotScalars*)s.inbox;
      std::memset(sc, 0, sizeof(*sc));
      sc->phase = PH_SETUP;
      sc->time_left = time_left(r, now);
      stage_values(*r.st, (double*)s.inbox + kInboxHeaderDoubles, r.Px.data(), r.q.data(), r.Ax.data(), r.b.data());
      SlotDesc D = make_desc(*s.dst, plan, s.arena, fp32);
      D.st = settings_dev;
      D.out_x = (double*)s.outbox;
      D.out_s = D.out_x + r.st->n;
      D.out_z = D.out_x + r.st->n + r.st->m;
      D.out_sc = (SlotScalars*)(s.outbox + ((8 * plan.out_lane + 63) & ~size_t(63)));
      D.out_sc->phase = PH_SETUP;  // PH_DONE only once the device has finished it
      D.inbox = (const double*)s.inbox;
      D.cancel_flag = s.cancel;
      s.seq = rseq++;
      rring[s.seq % (uint64_t)max_slots] = D;
      busy_requests++;
      admitted++;
      queue.pop_front();
    }
    if (admitted) {
      std::atomic_thread_fence(std::memory_order_seq_cst);  // descriptors and records before head
      __atomic_store_n(&rctrl->head, rseq, __ATOMIC_RELEASE);
    }
    return admitted;
  }
  // the finished requests: their done word carries the sequence number + 1
  void harvest_resident() {
    auto now = Clock::now();
    for (int i = 0; i < max_slots; ++i) {
      SlotRec& s = slots[i];
      if (!s.busy) continue;
      if (__atomic_load_n(&rdone[s.seq % (uint64_t)max_slots], __ATOMIC_ACQUIRE) != s.seq + 1) continue;
      const Structure& st = *s.dst->st;
      SlotScalars sc = *(const SlotScalars*)(s.outbox + ((8 * s.plan.out_lane + 63) & ~size_t(63)));
      if (sc.phase != PH_DONE) sc.status = ST_NUMERICAL_ERROR;  // not finished in its launch's passes
      record(s.ticket, st, (const double*)s.outbox, sc, s.t_submit, s.t_admit, now);
      put_arena(s.arena, s.bytes);
      free_coherent.insert({s.outbox_bytes, s.outbox});
      free_coherent.insert({s.inbox_bytes, s.inbox});
      in_use -= s.bytes;
      int32_t* cw = s.cancel;
      s = SlotRec();
      s.cancel = cw;  // kept for the slot's next request
    }
  }
  // a persistent launch is running whenever requests are in flight (its
  // workgroups leave after 100 ms without work; the next round relaunches)
  void ensure_resident_launch() {
    if (busy_requests == 0) return;
    if (rlaunched && !backend::stream_idle(stream)) return;
    backend::ResidentArgs a{&rctrl->head, &rctrl->stop, rring, rdone, rtail, max_slots,
                            std::max(1, settings.max_iter + 8), 10000000ull};
    backend::launch_resident(a, rblocks, stream, &rtimer);
    rlaunched = true;
    launches++;
  }
  // a drained server ends its persistent launch (the stop word lets every
  // workgroup leave at once), which frees the GPU and gives the launch's time
  void end_resident_launch() {
    if (!rlaunched) return;
    __atomic_store_n(&rctrl->stop, 1u, __ATOMIC_RELEASE);
    backend::sync(stream);
    __atomic_store_n(&rctrl->stop, 0u, __ATOMIC_RELEASE);
    kernel_time += backend::resident_seconds(rtimer);
    rlaunched = false;
    free_deferred();
  }
  int progress_resident(bool wait) {
    auto t0 = Clock::now();
    harvest_resident();
    if (busy_requests == 0 && queue.empty()) end_resident_launch();
    admit_resident();
    ensure_resident_launch();
    if (wait) {  // until a request finishes, or 2 ms
      const int before = busy_requests;
      while (busy_requests > 0 && busy_requests == before && seconds_since(t0) < 2e-3) {
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        harvest_resident();
        ensure_resident_launch();
      }
      if (busy_requests == 0 && queue.empty()) end_resident_launch();
    }
    busy_time += seconds_since(t0);
    return (int)queue.size() + busy_requests;
  }

  // the engine of a request: as set, or (auto) lanes for factorizations of at
  // most lane_max_flops whose 64-lane batch takes at most a quarter of the HBM
  // budget.  The default threshold is 0 (workgroup engine): measured on
  // MI355X the lane engine's serial per-lane chains of dependent HBM loads are
  // slower than the workgroup engine for the MPC stream.
  bool use_lanes(const Structure& st) const {
    if (resident) return false;
    if (settings.serve_engine == CLH_ENGINE_LANES) return true;
    if (settings.serve_engine == CLH_ENGINE_WORKGROUP) return false;
    return st.flops <= lane_max_flops &&
           (int64_t)plan_arena(st, settings.max_iter, kBatchLanes).bytes <= budget / 4;
  }
  // Finished requests' arenas are pooled by size for the next request of
  // that size, up to the HBM budget in pooled bytes (a server whose request
  // sizes keep changing would otherwise keep an arena of every size it has
  // seen).
  char* get_arena(size_t bytes) {
    auto it = free_arenas.find(bytes);
    if (it != free_arenas.end()) {
      char* p = it->second;
      free_arenas.erase(it);
      pooled_bytes -= bytes;
      return p;
    }
    return (char*)backend::dev_alloc(bytes);
  }
  void put_arena(char* p, size_t bytes) {
    if (pooled_bytes + (int64_t)bytes > budget) {
      // a device free waits for the device to go idle: deferred to the end
      // of the launch (free_deferred) rather than stalling a running one
      deferred_free.push_back(p);
      return;
    }
    free_arenas.insert({bytes, p});
    pooled_bytes += bytes;
  }
  void free_deferred() {
    for (char* p : deferred_free) backend::dev_free(p);
    deferred_free.clear();
  }
  // the pool holding exactly `need` arenas of each size (the others freed)
  void shape_pool(const std::map<size_t, int>& need) {
    for (auto it = free_arenas.begin(); it != free_arenas.end();) {
      auto n = need.find(it->first);
      const int want = n == need.end() ? 0 : n->second;
      if ((int)free_arenas.count(it->first) > want) {
        backend::dev_free(it->second);
        pooled_bytes -= it->first;
        it = free_arenas.erase(it);
      } else {
        ++it;
      }
    }
    for (auto& kv : need)
      for (int k = (int)free_arenas.count(kv.first); k < kv.second && pooled_bytes + (int64_t)kv.first <= budget; ++k) {
        free_arenas.insert({kv.first, (char*)backend::dev_alloc(kv.first)});
        pooled_bytes += kv.first;
      }
  }
  char* get_outbox(size_t bytes) {
    auto it = free_outboxes.find(bytes);
    if (it != free_outboxes.end()) {
      char* p = it->second;
      free_outboxes.erase(it);
      return p;
    }
    return (char*)backend::host_alloc(bytes);
  }
  // grow the upload staging buffer; the caller has synchronized the stream
  void ensure_stage(size_t bytes) {
    if (bytes <= stage_bytes) return;
    if (stage) backend::host_free(stage);
    stage_bytes = std::max(bytes, stage_bytes * 2);
    stage = (char*)backend::host_alloc(stage_bytes);
  }
  void ensure_out(size_t bytes) {
    if (bytes <= outbuf_bytes) return;
    if (outbuf) backend::host_free(outbuf);
    outbuf_bytes = std::max(bytes, outbuf_bytes * 2);
    outbuf = (char*)backend::host_alloc(outbuf_bytes);
  }
  std::shared_ptr<DevStructure>& shape_on_device(const std::shared_ptr<Structure>& st) {
    std::shared_ptr<DevStructure>& dst = dev_shapes[st->key];
    if (!dst) {
      dst = upload_structure(st, resident ? upload_stream : stream);
      if (resident) backend::sync(upload_stream);  // before any descriptor that points at it is published
    }
    return dst;
  }
  void reject_front() {
    Request& r = queue.front();
    rejected++;
    Result res;
    std::memset(&res.info, 0, sizeof(res.info));
    res.info.status = -1;
    res.latency = 0;
    done[r.ticket] = std::move(res);
    queue.pop_front();
  }

  // A request's time limit counts from its submission: what is left of it
  // when its first pass starts (the queueing time is spent already).
  double time_left(const Request& r, Clock::time_point now) const {
    return settings.time_limit - std::chrono::duration<double>(now - r.t_submit).count();
  }

  // Cancel a request: a queued one is answered at once (MAX_TIME, no
  // iterations); a resident one gets a negative time limit, so it ends with
  // MAX_TIME (or the status its last iterate earns under the reduced
  // tolerances) at its next termination check.  The stream is idle between
  // progress() calls.  Returns 0 when cancelled, 1 when already finished.
  int cancel(int64_t ticket) {
    finish_launch();
    if (resident) harvest_resident();
    if (done.count(ticket)) return 1;
    for (auto it = queue.begin(); it != queue.end(); ++it) {
      if (it->ticket != ticket) continue;
      Result res;
      const Structure& st = *it->st;
      res.x.assign(st.n, 0.0);
      res.s.assign(st.m, 0.0);
      res.z.assign(st.m, 0.0);
      std::memset(&res.info, 0, sizeof(res.info));
      res.info.status = CLH_MAX_TIME;
      res.info.obj_val = res.info.obj_val_dual = std::numeric_limits<double>::quiet_NaN();
      res.info.kkt_dim = st.N;
      res.latency = std::chrono::duration<double>(Clock::now() - it->t_submit).count();
      done[ticket] = std::move(res);
      queue.erase(it);
      cancelled++;
      return 0;
    }
    if (resident) {
      for (int i = 0; i < max_slots; ++i)
        if (slots[i].busy && slots[i].ticket == ticket) {
          *(volatile int32_t*)slots[i].cancel = 1;  // read by the slot at its next termination check
          cancelled++;
          return 0;
        }
      throw std::invalid_argument("unknown ticket");
    }
    char* dst = nullptr;
    for (int i = 0; i < max_slots && !dst; ++i)
      if (slots[i].busy && slots[i].ticket == ticket) dst = slots[i].arena + slots[i].plan.sc;
    for (int b = 0; b < max_batches && !dst; ++b) {
      BatchRec& B = batches[b];
      if (!B.live) continue;
      for (int l = 0; l < kBatchLanes; ++l)
        if (B.busy[l] && B.ticket[l] == ticket) { dst = B.arena + B.plan.sc + sizeof(SlotScalars) * l; break; }
    }
    if (!dst) throw std::invalid_argument("unknown ticket");
    ensure_stage(sizeof(double));
    *(double*)stage = -1.0;
    backend::h2d(dst + offsetof(SlotScalars, time_left), stage, sizeof(double), stream);
    backend::sync(stream);
    cancelled++;
    return 0;
  }

  // Admit queued requests (in order) while request capacity, table entries
  // and the HBM budget allow.  All uploads of one round go through one staging
  // buffer and are ordered before the next launch on the stream; the stream
  // is idle when admit() runs.
  int admit() {
    struct Up { char* dst; size_t off, bytes; };
    std::vector<Up> ups;
    size_t need = 0;
    auto reserve = [&](char* dst, size_t bytes) {
      size_t off = (need + 63) & ~size_t(63);
      need = off + bytes;
      ups.push_back({dst, off, bytes});
      return off;
    };
    // pass 1: decide placements and reserve staging space; pass 2 fills it
    struct Place { Request* r; bool lanes; int idx, lane; };
    std::vector<Place> places;     // lane engine: uploads
    std::vector<Place> wg_places;  // workgroup engine: inbox records
    std::vector<size_t> inbox_off;
    auto inbox_reserve = [&](size_t bytes) {
      size_t off = (need + 63) & ~size_t(63);
      need = off + bytes;
      inbox_off.push_back(off);
    };
    std::unordered_map<uint64_t, std::vector<int>> open;  // shape -> batches with a free lane
    for (int b = 0; b < max_batches; ++b)
      if (batches[b].live && batches[b].nbusy < kBatchLanes) open[batches[b].key].push_back(b);
    int admitted = 0;
    size_t qi = 0;
    std::vector<Request*> taken;
    while (qi < queue.size() && busy_requests < max_slots) {
      Request& r = queue[qi];
      if (r.lanes) {
        int b = -1;
        auto it = open.find(r.st->key);
        if (it != open.end() && !it->second.empty()) b = it->second.back();
        if (b < 0) {  // a new batch for this shape
          for (int i = 0; i < max_batches; ++i)
            if (!batches[i].live) { b = i; break; }
          if (b < 0) break;
          ArenaPlan plan = plan_arena(*r.st, settings.max_iter, kBatchLanes);
          if (in_use + (int64_t)plan.bytes > budget) {
            if (in_use == 0 && admitted == 0 && qi == 0) { reject_front(); continue; }
            break;
          }
          BatchRec& B = batches[b];
          B = BatchRec();
          B.live = true;
          B.key = r.st->key;
          B.plan = plan;
          B.bytes = plan.bytes;
          B.arena = get_arena(plan.bytes);
          B.dst = shape_on_device(r.st);
          for (int l = 0; l < kBatchLanes; ++l) B.busy[l] = false;
          in_use += plan.bytes;
          descs[max_slots + b] = make_desc(*B.dst, plan, B.arena, fp32);
          descs_dirty = true;
          // every lane starts idle (phase done); settings once per batch
          size_t o = reserve(B.arena + plan.sc, sizeof(SlotScalars) * kBatchLanes);
          places.push_back({nullptr, true, b, -1});
          (void)o;
          reserve(B.arena + plan.st, sizeof(DevSettings));
          places.push_back({nullptr, true, b, -2});
          open[B.key].push_back(b);
        }
        BatchRec& B = batches[b];
        int lane = -1;
        for (int l = 0; l < kBatchLanes; ++l)
          if (!B.busy[l]) { lane = l; break; }
        B.busy[lane] = true;
        B.nbusy++;
        B.ticket[lane] = r.ticket;
        B.t_submit[lane] = r.t_submit;
        B.t_admit[lane] = Clock::now();
        if (B.nbusy == kBatchLanes) {
          auto& v = open[B.key];
          v.erase(std::remove(v.begin(), v.end(), b), v.end());
        }
        places.push_back({&r, true, b, lane});
        reserve(B.arena + B.plan.inbox + 8 * B.plan.nvals * lane, 8 * B.plan.nvals);
        reserve(B.arena + B.plan.sc + sizeof(SlotScalars) * lane, sizeof(SlotScalars));
      } else {
        int slot = -1;
        for (int i = 0; i < max_slots; ++i)
          if (!slots[i].busy) { slot = i; break; }
        if (slot < 0) break;
        ArenaPlan plan = plan_arena(*r.st, settings.max_iter, 1);
        if (in_use + (int64_t)plan.bytes > budget) {
          if (in_use == 0 && admitted == 0 && qi == 0) { reject_front(); continue; }
          break;
        }
        SlotRec& s = slots[slot];
        s.busy = true;
        s.ticket = r.ticket;
        s.bytes = plan.bytes;
        s.arena = get_arena(plan.bytes);
        s.dst = shape_on_device(r.st);
        s.plan = plan;
        s.t_submit = r.t_submit;
        s.t_admit = Clock::now();
        in_use += plan.bytes;
        // outputs go straight to host-mapped memory; values and initial
        // scalars come from this round's inbox (no per-request copies)
        s.outbox_bytes = ((8 * plan.out_lane + 63) & ~size_t(63)) + sizeof(SlotScalars);
        s.outbox = get_outbox(s.outbox_bytes);
        SlotDesc& D = descs[slot];
        D = make_desc(*s.dst, plan, s.arena, fp32);
        D.st = settings_dev;
        D.out_x = (double*)s.outbox;
        D.out_s = D.out_x + r.st->n;
        D.out_z = D.out_x + r.st->n + r.st->m;
        D.out_sc = (SlotScalars*)(s.outbox + ((8 * plan.out_lane + 63) & ~size_t(63)));
        descs_dirty = true;
        wg_places.push_back({&r, false, slot, 0});
        inbox_reserve(8 * (kInboxHeaderDoubles + plan.nvals));
        fresh.push_back(slot);
      }
      busy_requests++;
      admitted++;
      taken.push_back(&r);
      ++qi;
    }
    if (places.empty() && wg_places.empty()) return 0;
    ensure_stage(need);
    const auto now_admit = Clock::now();
    size_t u = 0;
    for (const Place& p : places) {
      if (p.lanes && p.lane == -1) {  // fresh batch: all lanes idle
        SlotScalars* sc = (SlotScalars*)(stage + ups[u++].off);
        std::memset(sc, 0, sizeof(SlotScalars) * kBatchLanes);
        for (int l = 0; l < kBatchLanes; ++l) sc[l].phase = PH_DONE;
        continue;
      }
      if (p.lanes && p.lane == -2) {
        *(DevSettings*)(stage + ups[u++].off) = dsettings;
        continue;
      }
      const Request& r = *p.r;
      stage_values(*r.st, (double*)(stage + ups[u++].off), r.Px.data(), r.q.data(), r.Ax.data(), r.b.data());
      SlotScalars* sc = (SlotScalars*)(stage + ups[u++].off);
      std::memset(sc, 0, sizeof(*sc));
      sc->phase = PH_SETUP;
      sc->time_left = time_left(r, now_admit);
    }
    // workgroup slots: inbox records (header scalars, then the values)
    for (size_t k = 0; k < wg_places.size(); ++k) {
      const Place& p = wg_places[k];
      const Request& r = *p.r;
      double* rec = (double*)(stage + inbox_off[k]);
      SlotScalars* sc = (SlotScalars*)rec;
      std::memset(sc, 0, sizeof(*sc));
      sc->phase = PH_SETUP;
      sc->time_left = time_left(r, now_admit);
      stage_values(*r.st, rec + kInboxHeaderDoubles, r.Px.data(), r.q.data(), r.Ax.data(), r.b.data());
      descs[p.idx].inbox = rec;
    }
    for (const Up& up : ups) backend::h2d(up.dst, stage + up.off, up.bytes, stream);
    // a fresh batch's idle-lane scalars precede its first lane's scalars in
    // the upload order, so the later write wins
    queue.erase(queue.begin(), queue.begin() + qi);
    return admitted;
  }

  void record(int64_t ticket, const Structure& st, const double* o, const SlotScalars& sc,
              Clock::time_point t_submit, Clock::time_point t_admit, Clock::time_point now) {
    Result res;
    res.x.assign(o, o + st.n);
    res.s.assign(o + st.n, o + st.n + st.m);
    res.z.assign(o + st.n + st.m, o + st.n + 2 * st.m);
    std::memset(&res.info, 0, sizeof(res.info));
    fill_info(sc, st, &res.info);
    res.latency = std::chrono::duration<double>(now - t_submit).count();
    res.info.solve_time = std::chrono::duration<double>(now - t_admit).count();
    latencies.push_back(res.latency);
    done[ticket] = std::move(res);
    completed++;
    busy_requests--;
  }

  // record every finished request: workgroup slots wrote their outputs to
  // host-mapped memory; lane batches are copied out (one synchronization)
  void harvest() {
    auto now = Clock::now();
    for (int i = 0; i < max_slots; ++i) {
      SlotRec& s = slots[i];
      if (!s.busy || phase_host[i] != PH_DONE) continue;
      const Structure& st = *s.dst->st;
      const SlotScalars* sc = (const SlotScalars*)(s.outbox + ((8 * s.plan.out_lane + 63) & ~size_t(63)));
      record(s.ticket, st, (const double*)s.outbox, *sc, s.t_submit, s.t_admit, now);
      put_arena(s.arena, s.bytes);
      free_outboxes.insert({s.outbox_bytes, s.outbox});
      in_use -= s.bytes;
      s = SlotRec();
      std::memset(&descs[i], 0, sizeof(SlotDesc));
      descs[i].n = -1;
      descs_dirty = true;
    }
    std::vector<Pending> pend;
    std::vector<std::pair<int, int>> freed;  // (batch, lane)
    size_t need = 0;
    auto take = [&](size_t bytes) {
      size_t off = (need + 63) & ~size_t(63);
      need = off + bytes;
      return off;
    };
    for (int b = 0; b < max_batches; ++b) {
      BatchRec& B = batches[b];
      if (!B.live) continue;
      for (int l = 0; l < kBatchLanes; ++l) {
        if (!B.busy[l] || phase_host[max_slots + b * kBatchLanes + l] != PH_DONE) continue;
        Pending p{B.ticket[l], B.dst->st.get(), take(8 * B.plan.out_lane), take(sizeof(SlotScalars)),
                  B.t_submit[l], B.t_admit[l]};
        pend.push_back(p);
        freed.push_back({b, l});
      }
    }
    if (pend.empty()) return;
    ensure_out(need);
    for (size_t k = 0; k < pend.size(); ++k) {
      const Pending& p = pend[k];
      BatchRec& B = batches[freed[k].first];
      const int l = freed[k].second;
      backend::d2h(outbuf + p.off_out, B.arena + B.plan.out + 8 * B.plan.out_lane * l, 8 * B.plan.out_lane, stream);
      backend::d2h(outbuf + p.off_sc, B.arena + B.plan.sc + sizeof(SlotScalars) * l, sizeof(SlotScalars), stream);
    }
    backend::sync(stream);
    now = Clock::now();
    for (size_t k = 0; k < pend.size(); ++k) {
      const Pending& p = pend[k];
      record(p.ticket, *p.st, (const double*)(outbuf + p.off_out), *(const SlotScalars*)(outbuf + p.off_sc),
             p.t_submit, p.t_admit, now);
      int b = freed[k].first, i = freed[k].second;
      BatchRec& B = batches[b];
      B.busy[i] = false;
      B.nbusy--;
      // the lane stays in phase done on the device and idles until reused
      if (B.nbusy == 0) {
        put_arena(B.arena, B.bytes);
        in_use -= B.bytes;
        B = BatchRec();
        std::memset(&descs[max_slots + b], 0, sizeof(SlotDesc));
        descs[max_slots + b].n = -1;
        descs_dirty = true;
      }
    }
  }

  // the end of an asynchronous launch: wait for it, account for it, harvest
  void finish_launch() {
    if (!pending) return;
    backend::sync(stream);
    pending = false;
    kernel_time += backend::graph_last_seconds(graph);
    launches++;
    // this round's inbox has been read: its slots continue from their arenas
    for (int i : fresh) descs[i].inbox = nullptr;
    if (!fresh.empty()) descs_dirty = true;
    fresh.clear();
    harvest();
    free_deferred();
  }

  // One round: finish the previous launch (if one is pending), harvest,
  // admit, launch `passes` passes over every resident slot.  With wait the
  // launch is finished before returning; without, it runs while the caller
  // goes on (submits more requests, say) and the next call finishes it.  A
  // new admission always waits for the previous launch, which still reads its
  // inbox records from the staging memory.
  int progress(int passes, bool wait = true) {
    if (resident) return progress_resident(wait);
    auto t0 = Clock::now();
    finish_launch();
    harvest();
    admit();
    int nslots_used = 0, nbatches_used = 0;
    for (int i = 0; i < max_slots; ++i) if (slots[i].busy) nslots_used = i + 1;
    for (int b = 0; b < max_batches; ++b) if (batches[b].live) nbatches_used = b + 1;
    if (nslots_used + nbatches_used > 0) {
      if (descs_dirty) {
        backend::h2d(descs_dev, descs, sizeof(SlotDesc) * ndescs, stream);
        descs_dirty = false;
      }
      // launch over whole tables that are in use (fixed per graph); the
      // read-back covers both phase tables, stale entries of idle slots and
      // lanes are ignored by harvest()
      const int ns = nslots_used ? max_slots : 0, nb = nbatches_used ? max_batches : 0;
      if (graph && (graph_passes != passes || graph_ns != ns || graph_nb != nb)) {
        backend::sync(stream);
        backend::graph_destroy(graph);
        graph = nullptr;
      }
      graph_passes = passes; graph_ns = ns; graph_nb = nb;
      backend::LaunchSet ls{descs_dev, ns, phase_dev, descs_dev + max_slots, nb, phase_dev + max_slots};
      const size_t nphase = (size_t)max_slots + (size_t)max_batches * kBatchLanes;
      backend::launch_ipm_graph(&graph, ls, passes, phase_dev, phase_host, nphase, stream);
      pending = true;
      if (wait) finish_launch();
    }
    busy_time += seconds_since(t0);
    return (int)queue.size() + busy_requests;
  }
};

extern "C" {

int clh_server_create(clh_server** out, int64_t hbm_budget_bytes, int32_t max_slots, const clh_settings* settings,
                      void* stream) {
  CLH_TRY({
    if (!out || max_slots <= 0 || hbm_budget_bytes <= 0) throw std::invalid_argument("invalid server arguments");
    std::unique_ptr<clh_server> S(new clh_server);
    if (settings) S->settings = *settings;
    else default_settings(&S->settings);
    check_settings(S->settings);
    if (S->settings.serve_engine == CLH_ENGINE_GRID)
      throw std::invalid_argument("the grid engine runs one problem on the whole GPU: use clh_solver_create");
    // auto serves with the device-resident loop: measured on the MPC stream it
    // outruns the launch-per-round workgroup engine (5455 vs 4729 solves/s)
    // and halves the median latency at half load
    S->resident = S->settings.serve_engine == CLH_ENGINE_RESIDENT || S->settings.serve_engine == CLH_ENGINE_AUTO;
    S->fp32 = S->settings.direct_solve_method == CLH_METHOD_HIPMIXED;
    S->base = S->settings.schedule;
    if (S->base.use_tuned)  // the cached schedules of every size class of this device and method
      for (int c = 0; c < 40; ++c) {
        clh_schedule t;
        if (tune_cache_lookup(tune_key("serve", S->fp32, c), &t)) {
          S->cls[c] = t;
          S->base.blocks_per_cu = t.blocks_per_cu;
        }
      }
    S->dsettings = to_dev(S->settings);
    if (S->fp32) S->dsettings.static_regularization_constant = kMixedStaticReg;  // as gpu_preprocess
    if (const char* e = std::getenv("CLH_LANE_MAX_FLOPS")) S->lane_max_flops = std::atof(e);
    S->budget = hbm_budget_bytes;
    S->max_slots = max_slots;
    // lane batches: enough for max_slots requests in full batches plus one
    // partly filled batch for each of up to 64 shapes
    S->max_batches = (max_slots + kBatchLanes - 1) / kBatchLanes + 64;
    S->stream = stream ? stream : backend::stream_create();
    S->own_stream = stream == nullptr;
    S->slots.resize(max_slots);
    S->batches.resize(S->max_batches);
    S->ndescs = (size_t)max_slots + S->max_batches;
    S->descs = (SlotDesc*)backend::host_alloc(sizeof(SlotDesc) * S->ndescs);
    for (size_t i = 0; i < S->ndescs; ++i) { std::memset(&S->descs[i], 0, sizeof(SlotDesc)); S->descs[i].n = -1; }
    S->descs_dev = (SlotDesc*)backend::dev_alloc(sizeof(SlotDesc) * S->ndescs);
    S->settings_dev = (DevSettings*)backend::dev_alloc(sizeof(DevSettings));
    backend::h2d(S->settings_dev, &S->dsettings, sizeof(DevSettings), S->stream);
    backend::sync(S->stream);
    const size_t nphase = (size_t)max_slots + (size_t)S->max_batches * kBatchLanes;
    S->phase_dev = (int32_t*)backend::dev_alloc(4 * nphase + 256);
    S->phase_host = (int32_t*)backend::host_alloc(4 * nphase + 256);
    for (size_t i = 0; i < nphase; ++i) S->phase_host[i] = -1;
    if (S->resident) {
      S->rctrl = (clh_server::ResCtrl*)backend::host_alloc_coherent(sizeof(clh_server::ResCtrl));
      std::memset(S->rctrl, 0, sizeof(clh_server::ResCtrl));
      S->rring = (SlotDesc*)backend::host_alloc_coherent(sizeof(SlotDesc) * max_slots);
      S->rdone = (uint64_t*)backend::host_alloc_coherent(sizeof(uint64_t) * max_slots);
      std::memset(S->rdone, 0, sizeof(uint64_t) * max_slots);
      S->rtail = (uint64_t*)backend::dev_alloc(sizeof(uint64_t));
      const uint64_t zero = 0;
      backend::h2d(S->rtail, &zero, sizeof(zero), S->stream);
      backend::sync(S->stream);
      S->rblocks = S->resident_block_count();
      S->upload_stream = backend::stream_create();
    }
    *out = S.release();
    return 0;
  })
}

int clh_server_submit(clh_server* srv, int64_t* ticket, int32_t n, int32_t m, const int32_t* Pp, const int32_t* Pi,
                      const double* Px, const double* q, const int32_t* Ap, const int32_t* Ai, const double* Ax,
                      const double* b, int32_t ncones, const int32_t* cone_type, const int32_t* cone_dim,
                      const double* cone_param) {
  CLH_TRY({
    if (!srv || !ticket || n < 0 || m < 0 || !Pp || !Ap) throw std::invalid_argument("invalid submit arguments");
    auto cones = cone_specs(ncones, cone_type, cone_dim, cone_param);
    Request r;
    try {
      r.st = g_shapes.get(n, m, Pp, Pi, Ap, Ai, cones, srv->fp32, sym_opts(srv->sched_for(n + m)));
    } catch (const std::runtime_error& e) {
      throw std::invalid_argument(e.what());
    }
    r.ticket = srv->next_ticket++;
    r.lanes = srv->use_lanes(*r.st);
    r.Px.assign(Px, Px + Pp[n]);
    r.q.assign(q, q + n);
    r.Ax.assign(Ax, Ax + Ap[n]);
    r.b.assign(b, b + m);
    r.t_submit = Clock::now();
    *ticket = r.ticket;
    srv->queue.push_back(std::move(r));
    srv->submitted++;
    return 0;
  })
}

int clh_server_progress(clh_server* srv, int32_t passes) {
  CLH_TRY({
    if (!srv) throw std::invalid_argument("null server");
    return srv->progress(std::max(1, (int)passes));
  })
}

int clh_server_progress_async(clh_server* srv, int32_t passes) {
  CLH_TRY({
    if (!srv) throw std::invalid_argument("null server");
    return srv->progress(std::max(1, (int)passes), false);
  })
}

int clh_server_cancel(clh_server* srv, int64_t ticket) {
  CLH_TRY({
    if (!srv) throw std::invalid_argument("null server");
    if (ticket <= 0 || ticket >= srv->next_ticket) throw std::invalid_argument("unknown ticket");
    return srv->cancel(ticket);
  })
}

int clh_server_wait(clh_server* srv, int64_t ticket, int32_t passes, double timeout_s) {
  CLH_TRY({
    if (!srv) throw std::invalid_argument("null server");
    if (ticket <= 0 || ticket >= srv->next_ticket) throw std::invalid_argument("unknown ticket");
    const auto t0 = Clock::now();
    for (;;) {
      if (srv->done.count(ticket)) return 1;
      if (seconds_since(t0) >= timeout_s) return 0;
      // a ticket neither finished, queued nor resident was already collected
      if (srv->progress(std::max(1, (int)passes)) == 0 && !srv->done.count(ticket))
        throw std::invalid_argument("ticket already collected");
    }
  })
}

int clh_server_poll(clh_server* srv, int64_t ticket) {
  if (!srv) return -1;
  if (srv->done.count(ticket)) return 1;
  if (ticket <= 0 || ticket >= srv->next_ticket) return -1;
  return 0;
}

int clh_server_result(clh_server* srv, int64_t ticket, double* x, double* s, double* z, clh_info* info,
                      double* latency_s) {
  if (!srv) return -1;
  auto it = srv->done.find(ticket);
  if (it == srv->done.end()) { g_last_error = "result not available"; return -1; }
  Result& r = it->second;
  if (r.info.status < 0) { g_last_error = "request rejected: exceeds the HBM budget"; srv->done.erase(it); return 2; }
  if (x) std::memcpy(x, r.x.data(), 8 * r.x.size());
  if (s) std::memcpy(s, r.s.data(), 8 * r.s.size());
  if (z) std::memcpy(z, r.z.data(), 8 * r.z.size());
  if (info) *info = r.info;
  if (latency_s) *latency_s = r.latency;
  srv->done.erase(it);
  return 0;
}

int clh_server_autotune(clh_server* srv, double max_seconds, clh_tune_report* report) {
  CLH_TRY({
    if (!srv) throw std::invalid_argument("null server");
    const clh_tune_report R = srv->autotune(max_seconds);
    if (report) *report = R;
    return 0;
  })
}

int clh_server_get_schedule(clh_server* srv, int32_t n_plus_m, clh_schedule* out) {
  if (!srv || !out) { g_last_error = "null argument"; return -1; }
  *out = srv->sched_for(n_plus_m);
  out->blocks_per_cu = srv->base.blocks_per_cu;
  return 0;
}

int clh_server_stats_get(clh_server* srv, clh_server_stats* st) {
  if (!srv || !st) return -1;
  std::memset(st, 0, sizeof(*st));
  st->submitted = srv->submitted;
  st->completed = srv->completed;
  st->rejected = srv->rejected;
  st->active_slots = srv->busy_requests;
  st->queued = (int32_t)srv->queue.size();
  st->hbm_in_use = srv->in_use;
  if (!srv->latencies.empty()) {
    std::vector<double> l = srv->latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {
      double idx = p * (l.size() - 1);
      size_t lo = (size_t)std::floor(idx), hi = (size_t)std::ceil(idx);
      return l[lo] + (l[hi] - l[lo]) * (idx - lo);
    };
    st->p50_latency_s = pct(0.5);
    st->p99_latency_s = pct(0.99);
    double sum = 0;
    for (double v : l) sum += v;
    st->mean_latency_s = sum / l.size();
  }
  st->busy_time_s = srv->busy_time;
  st->launches = srv->launches;
  st->kernel_time_s = srv->kernel_time;
  st->cancelled = srv->cancelled;
  return 0;
}

void clh_server_destroy(clh_server* srv) { delete srv; }

// ---------------------------------------------------------------------------
// LDL' plugin
// ---------------------------------------------------------------------------
}  // extern "C"

struct clh_ldl {
  std::shared_ptr<Structure> st;
  std::shared_ptr<DevStructure> dst;
  ArenaPlan plan;
  char* arena = nullptr;
  SlotDesc desc;
  SlotDesc* desc_dev = nullptr;
  int32_t* map_dev = nullptr;
  double* out_dev = nullptr;
  DevSettings ds;
  bool fp32 = false;
  double last[4] = {0, 0, 0, 0};  // ok, regularized pivots, refinement steps, error
  int n_pos = 0, n_neg = 0;
  // a factorization of at least kGridMinFlops runs over a cooperative grid
  // of grid_nb workgroups (the whole GPU), else on one workgroup
  int grid_nb = 0;
  double* gred = nullptr;
  size_t gred_bytes = 0;
  int32_t* abort_word = nullptr;
  ~clh_ldl() {
    if (abort_word) backend::host_free(abort_word);
    if (gred) backend::dev_free(gred);
    if (arena) backend::dev_free(arena);
    if (desc_dev) backend::dev_free(desc_dev);
    if (map_dev) backend::dev_free(map_dev);
    if (out_dev) backend::dev_free(out_dev);
  }
  void run(const backend::LdlArgs& a, void* stream, bool read_out) {
    if (grid_nb > 0) {
      *(volatile int32_t*)abort_word = 0;
      backend::launch_grid_ldl(desc_dev, a, gred, grid_nb, stream);
      if (read_out) backend::d2h(last, out_dev, sizeof(last), stream);
      backend::sync(stream);
      if (*(volatile int32_t*)abort_word) {
        const std::vector<char> zero(gred_bytes, 0);
        backend::h2d(gred, zero.data(), gred_bytes, stream);
        backend::sync(stream);
        *(volatile int32_t*)abort_word = 0;
        throw std::runtime_error("grid launch aborted: a grid barrier waited 30 s for a workgroup");
      }
      return;
    }
    backend::launch_ldl(desc_dev, a, stream);
    if (read_out) backend::d2h(last, out_dev, sizeof(last), stream);
    backend::sync(stream);
  }
};

extern "C" {

int clh_ldl_create(clh_ldl** out, int32_t n, const int32_t* rowptr, const int32_t* colval, int32_t index_base,
                   const int32_t* dsigns, int32_t precision) {
  CLH_TRY({
    if (!out || n <= 0 || !rowptr || !colval || !dsigns) throw std::invalid_argument("invalid arguments to clh_ldl_create");
    if (precision != CLH_METHOD_HIP && precision != CLH_METHOD_HIPMIXED)
      throw std::invalid_argument("precision must be CLH_METHOD_HIP or CLH_METHOD_HIPMIXED");
    std::unique_ptr<clh_ldl> H(new clh_ldl);
    std::vector<int32_t> map;
    try {
      H->st = build_ldl_structure(n, rowptr, colval, index_base, dsigns, map, false);
    } catch (const std::runtime_error& e) {
      throw std::invalid_argument(e.what());
    }
    H->fp32 = precision == CLH_METHOD_HIPMIXED;
    H->dst = upload_structure(H->st, nullptr);
    int threads = 0;
    if (H->st->flops >= kGridMinFlops || std::getenv("CLH_LDL_GRID")) {  // CLH_LDL_GRID: diagnostics, always
      int cap = H->st->flops >= kGridWideFlops ? kGridWideBlocks : kGridMaxBlocks;
      if (const char* e = std::getenv("CLH_GRID_BLOCKS")) cap = std::max(1, std::atoi(e));
      H->grid_nb = backend::grid_ldl_blocks(cap);
      threads = H->grid_nb * backend::grid_threads();
      H->gred_bytes = sizeof(double) * 2 * H->grid_nb + backend::kGridBarBytes;
      H->gred = (double*)backend::dev_alloc(H->gred_bytes);
      const std::vector<char> zero(H->gred_bytes, 0);
      backend::h2d(H->gred, zero.data(), H->gred_bytes, nullptr);
      H->abort_word = (int32_t*)backend::host_alloc_coherent(sizeof(int32_t));
      *(volatile int32_t*)H->abort_word = 0;
    }
    H->plan = plan_arena(*H->st, 0, 1, threads);
    H->arena = (char*)backend::dev_alloc(H->plan.bytes);
    H->desc = make_desc(*H->dst, H->plan, H->arena, H->fp32);
    H->desc.abort_flag = H->abort_word;
    H->desc_dev = (SlotDesc*)backend::dev_alloc(sizeof(SlotDesc));
    H->map_dev = (int32_t*)backend::dev_alloc(4 * map.size() + 8);
    H->out_dev = (double*)backend::dev_alloc(8 * 4);
    clh_settings cs;
    default_settings(&cs);
    H->ds = to_dev(cs);
    // the plugin stands in for the reference's factorization object, whose
    // dynamic regularization sets sign * delta (the caller's settings carry
    // no other rule)
    H->ds.dyn_keep_magnitude = 0;
    backend::h2d(H->desc_dev, &H->desc, sizeof(SlotDesc), nullptr);
    backend::h2d(H->map_dev, map.data(), 4 * map.size(), nullptr);
    backend::sync(nullptr);
    *out = H.release();
    return 0;
  })
}

int clh_ldl_refactor(clh_ldl* h, const double* vals_dev, double static_eps, int32_t dynamic_enable, double dyn_eps,
                     double dyn_delta, void* stream) {
  CLH_TRY({
    if (!h || !vals_dev) throw std::invalid_argument("invalid arguments to clh_ldl_refactor");
    h->ds.dynamic_regularization_enable = dynamic_enable ? 1 : 0;
    h->ds.dynamic_regularization_eps = dyn_eps;
    h->ds.dynamic_regularization_delta = dyn_delta;
    backend::h2d(h->arena + h->plan.st, &h->ds, sizeof(DevSettings), stream);
    backend::LdlArgs a{};
    a.op = 0; a.vals = vals_dev; a.map = h->map_dev; a.static_eps = static_eps; a.out = h->out_dev;
    h->run(a, stream, true);
    // inertia from the pivot signs
    const int N = h->st->N;
    std::vector<double> d(N);
    if (h->fp32) {
      std::vector<float> f(N);
      backend::d2h(f.data(), h->desc.Dv, 4 * (size_t)N, stream);
      backend::sync(stream);
      for (int i = 0; i < N; ++i) d[i] = f[i];
    } else {
      backend::d2h(d.data(), h->desc.Dv, 8 * (size_t)N, stream);
      backend::sync(stream);
    }
    h->n_pos = h->n_neg = 0;
    for (double v : d) { if (v > 0) h->n_pos++; else if (v < 0) h->n_neg++; }
    if (h->last[0] == 0) { g_last_error = "zero or non-finite pivot"; return 1; }
    return 0;
  })
}

int clh_ldl_solve(clh_ldl* h, double* x_dev, const double* b_dev, void* stream) {
  CLH_TRY({
    if (!h || !x_dev || !b_dev) throw std::invalid_argument("invalid arguments to clh_ldl_solve");
    backend::LdlArgs a{};
    a.op = 1; a.x = x_dev; a.b = b_dev; a.out = h->out_dev;
    h->run(a, stream, false);
    return 0;
  })
}

int clh_ldl_solve_refine(clh_ldl* h, const double* vals_dev, double* x_dev, const double* b_dev, double reltol,
                         double abstol, int32_t maxiter, double stopratio, void* stream) {
  CLH_TRY({
    if (!h || !vals_dev || !x_dev || !b_dev || maxiter < 0)
      throw std::invalid_argument("invalid arguments to clh_ldl_solve_refine");
    backend::LdlArgs a{};
    a.op = 2; a.vals = vals_dev; a.map = h->map_dev; a.x = x_dev; a.b = b_dev;
    a.reltol = reltol; a.abstol = abstol; a.maxiter = maxiter; a.stopratio = stopratio; a.out = h->out_dev;
    h->run(a, stream, true);
    if (h->last[0] == 0) { g_last_error = "refinement residual is not finite"; return 1; }
    return 0;
  })
}

int clh_ldl_get_info(clh_ldl* h, clh_ldl_info* info) {
  if (!h || !info) { g_last_error = "invalid arguments to clh_ldl_get_info"; return -1; }
  std::memset(info, 0, sizeof(*info));
  info->nnz_L = h->st->nnzL;
  info->factor_flops = h->st->flops;
  info->n_supernodes = h->st->nsn;
  info->n_levels = h->st->nlev;
  info->n_positive = h->n_pos;
  info->n_negative = h->n_neg;
  info->last_dynreg = (int32_t)h->last[1];
  info->last_ir_steps = (int32_t)h->last[2];
  info->last_refine_error = h->last[3];
  info->grid_blocks = h->grid_nb;
  return 0;
}

void clh_ldl_destroy(clh_ldl* h) { delete h; }

int clh_device_alloc(void** ptr, size_t bytes) {
  CLH_TRY({
    if (!ptr) throw std::invalid_argument("null pointer");
    *ptr = backend::dev_alloc(bytes);
    return 0;
  })
}
int clh_device_free(void* ptr) {
  backend::dev_free(ptr);
  return 0;
}
int clh_memcpy_to_device(void* dst_dev, const void* src, size_t bytes) {
  CLH_TRY({
    backend::h2d(dst_dev, src, bytes, nullptr);
    backend::sync(nullptr);
    return 0;
  })
}
int clh_memcpy_to_host(void* dst, const void* src_dev, size_t bytes) {
  CLH_TRY({
    backend::d2h(dst, src_dev, bytes, nullptr);
    backend::sync(nullptr);
    return 0;
  })
}

// ---------------------------------------------------------------------------
// composite cone operations
// ---------------------------------------------------------------------------
}  // extern "C"

struct clh_cones {
  std::shared_ptr<Structure> st;
  std::shared_ptr<DevStructure> dst;
  ArenaPlan plan;
  char* arena = nullptr;
  SlotDesc desc;
  SlotDesc* desc_dev = nullptr;
  double* partial = nullptr;  // 2 per workgroup
  double* scal = nullptr;     // reduced scalars
  int64_t* hsmap = nullptr;   // hs_row (m), hs_soc, hs_nsym, hs_psd
  int64_t hs_len = 0;
  int m = 0, nblocks = 1, n_zero = 0, n_nonneg = 0;
  ~clh_cones() {
    for (void* p : {(void*)arena, (void*)desc_dev, (void*)partial, (void*)scal, (void*)hsmap})
      if (p) backend::dev_free(p);
  }
  backend::ConeArgs args(int op) const {
    backend::ConeArgs a{};
    a.op = op;
    a.partial = partial;
    a.result = scal;
    const int64_t* hm = hsmap;
    a.hs_row = hm;
    a.hs_soc = hm + m;
    a.hs_nsym = a.hs_soc + st->soc_off.size();
    a.hs_psd = a.hs_nsym + st->exp_off.size() + st->pow_off.size();
    return a;
  }
  void run(const backend::ConeArgs& a, void* stream) const { backend::launch_cones(desc_dev, a, nblocks, stream); }
  // a reduction whose result is copied to the caller's device scalars
  void reduce(const backend::ConeArgs& a, int mode, double* out_dev, int nout, void* stream) const {
    backend::launch_cones_reduce(a, nblocks, mode, stream);
    if (out_dev) backend::d2d(out_dev, scal, sizeof(double) * nout, stream);
  }
};

#define CLH_CONES_CHECK(h) \
  if (!(h)) throw std::invalid_argument("null cone handle")

extern "C" {

int clh_cones_create(clh_cones** out, int32_t ncones, const int32_t* cone_type, const int32_t* cone_dim,
                     const double* cone_param, const clh_settings* settings) {
  CLH_TRY({
    if (!out || ncones < 0 || (ncones > 0 && (!cone_type || !cone_dim))) throw std::invalid_argument("invalid cones");
    std::unique_ptr<clh_cones> H(new clh_cones);
    auto cones = cone_specs(ncones, cone_type, cone_dim, cone_param);
    int m = 0;
    for (auto& c : cones) m += c.dim;
    if (m <= 0) throw std::invalid_argument("the cones have no rows");
    const int32_t zero = 0;
    try {
      H->st = build_structure(0, m, &zero, nullptr, &zero, nullptr, cones, false);
    } catch (const std::runtime_error& e) {
      throw std::invalid_argument(e.what());
    }
    H->m = m;
    clh_settings cs;
    if (settings) cs = *settings;
    else default_settings(&cs);
    DevSettings ds = to_dev(cs);
    H->dst = upload_structure(H->st, nullptr);
    H->plan = plan_arena(*H->st, 0, 1);
    H->arena = (char*)backend::dev_alloc(H->plan.bytes);
    H->desc = make_desc(*H->dst, H->plan, H->arena, false);
    H->desc_dev = (SlotDesc*)backend::dev_alloc(sizeof(SlotDesc));
    const Structure& S = *H->st;
    int work = m;
    H->nblocks = backend::cone_blocks(work);
    H->partial = (double*)backend::dev_alloc(sizeof(double) * 2 * H->nblocks);
    H->scal = (double*)backend::dev_alloc(sizeof(double) * 4);
    // set-up of clh_cones_step_length_noalloc: this device's scratch address
    (void)backend::cone_scratch(nullptr);
    // Hs block positions in cone order
    std::vector<int64_t> hm(m + S.soc_off.size() + S.exp_off.size() + S.pow_off.size() + S.psd_off.size() + 1, -1);
    int64_t* hs_row = hm.data();
    int64_t* hs_soc = hs_row + m;
    int64_t* hs_nsym = hs_soc + S.soc_off.size();
    int64_t* hs_psd = hs_nsym + S.exp_off.size() + S.pow_off.size();
    int64_t pos = 0;
    int off = 0, ksoc = 0, kexp = 0, kpow = 0, kpsd = 0;
    for (auto& c : cones) {
      switch (c.type) {
        case CT_ZERO: H->n_zero++; for (int i = 0; i < c.dim; ++i) hs_row[off + i] = pos++; break;
        case CT_NONNEG: H->n_nonneg++; for (int i = 0; i < c.dim; ++i) hs_row[off + i] = pos++; break;
        case CT_SOC: hs_soc[ksoc++] = pos; pos += c.dim > 5 ? c.dim : (int64_t)c.dim * c.dim; break;
        case CT_EXP: hs_nsym[kexp++] = pos; pos += 9; break;
        case CT_POW: hs_nsym[S.exp_off.size() + kpow++] = pos; pos += 9; break;
        case CT_PSD: hs_psd[kpsd++] = pos; pos += (int64_t)c.dim * c.dim; break;
        default: break;
      }
      off += c.dim;
    }
    H->hs_len = pos;
    H->hsmap = (int64_t*)backend::dev_alloc(sizeof(int64_t) * hm.size());
    backend::h2d(H->hsmap, hm.data(), sizeof(int64_t) * hm.size(), nullptr);
    backend::h2d(H->desc_dev, &H->desc, sizeof(SlotDesc), nullptr);
    backend::h2d(H->arena + H->plan.st, &ds, sizeof(DevSettings), nullptr);
    backend::sync(nullptr);
    *out = H.release();
    return 0;
  })
}

int clh_cones_get_info(clh_cones* h, clh_cones_info* info) {
  if (!h || !info) { g_last_error = "invalid arguments to clh_cones_get_info"; return -1; }
  std::memset(info, 0, sizeof(*info));
  const Structure& S = *h->st;
  info->m = h->m;
  info->degree = S.degree;
  info->n_zero = h->n_zero;
  info->n_nonneg = h->n_nonneg;
  info->n_soc = (int32_t)S.soc_off.size();
  info->n_exp = (int32_t)S.exp_off.size();
  info->n_pow = (int32_t)S.pow_off.size();
  info->n_psd = (int32_t)S.psd_off.size();
  info->symmetric = S.symmetric ? 1 : 0;
  info->hs_len = h->hs_len;
  return 0;
}

int clh_cones_update_scaling(clh_cones* h, const double* s, const double* z, double mu, int32_t scaling_strategy,
                             double* ok_dev, void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    if (!s || !z) throw std::invalid_argument("null vector");
    auto a = h->args(backend::CO_UPDATE_SCALING);
    a.a = s; a.b = z; a.s0 = mu; a.flag = scaling_strategy ? 1 : 0;
    h->run(a, stream);
    h->reduce(a, 0, ok_dev, 1, stream);
    return 0;
  })
}

int clh_cones_set_identity_scaling(clh_cones* h, void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    h->run(h->args(backend::CO_IDENTITY_SCALING), stream);
    return 0;
  })
}

int clh_cones_get_Hs(clh_cones* h, double* Hsblocks, void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    if (!Hsblocks) throw std::invalid_argument("null vector");
    auto a = h->args(backend::CO_GET_HS);
    a.x = Hsblocks;
    h->run(a, stream);
    return 0;
  })
}

int clh_cones_mul_Hs(clh_cones* h, double* y, const double* x, void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    if (!y || !x) throw std::invalid_argument("null vector");
    auto a = h->args(backend::CO_MUL_HS);
    a.x = y; a.a = x;
    h->run(a, stream);
    return 0;
  })
}

int clh_cones_affine_ds(clh_cones* h, double* ds, const double* s, void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    if (!ds || !s) throw std::invalid_argument("null vector");
    auto a = h->args(backend::CO_AFFINE_DS);
    a.x = ds; a.a = s;
    h->run(a, stream);
    return 0;
  })
}

int clh_cones_combined_ds_shift(clh_cones* h, double* shift, double* step_z, double* step_s, double sigma_mu,
                                void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    if (!shift || !step_z || !step_s) throw std::invalid_argument("null vector");
    auto a = h->args(backend::CO_COMBINED_DS_SHIFT);
    a.x = shift; a.y = step_z; a.z = step_s; a.s0 = sigma_mu;
    h->run(a, stream);
    return 0;
  })
}

int clh_cones_ds_from_dz_offset(clh_cones* h, double* out, const double* ds, const double* z, void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    if (!out || !ds || !z) throw std::invalid_argument("null vector");
    auto a = h->args(backend::CO_DS_FROM_DZ_OFFSET);
    a.x = out; a.a = ds; a.b = z;
    h->run(a, stream);
    return 0;
  })
}

int clh_cones_step_length(clh_cones* h, const double* dz, const double* ds, const double* z, const double* s,
                          double alpha_max, double* alpha_dev, void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    if (!dz || !ds || !z || !s || !alpha_dev) throw std::invalid_argument("null vector");
    auto a = h->args(backend::CO_STEP_SYMMETRIC);
    a.a = dz; a.b = ds; a.c = z; a.d = s; a.s0 = alpha_max;
    h->run(a, stream);
    if (h->st->symmetric) {
      h->reduce(a, 0, alpha_dev, 1, stream);
    } else {
      backend::launch_cones_reduce(a, h->nblocks, 0, stream);
      a.op = backend::CO_STEP_NONSYMMETRIC;  // backtracks from min(alpha, max_step_fraction) in scal[0]
      h->run(a, stream);
      h->reduce(a, 0, alpha_dev, 1, stream);
    }
    return 0;
  })
}

// the same operation with its partials and reduced scalars in the library's
// static scratch: no allocation, one composite of at most the scratch's
// workgroups at a time per GPU
int clh_cones_step_length_noalloc(clh_cones* h, const double* dz, const double* ds, const double* z, const double* s,
                                  double alpha_max, double* alpha_dev, void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    if (!dz || !ds || !z || !s || !alpha_dev) throw std::invalid_argument("null vector");
    int max_blocks = 0;
    double* scr = backend::cone_scratch(&max_blocks);
    if (h->nblocks > max_blocks)
      throw std::invalid_argument("clh_cones_step_length_noalloc: this composite cone takes " +
                                  std::to_string(h->nblocks) + " workgroups, the static scratch holds " +
                                  std::to_string(max_blocks) + " (use clh_cones_step_length)");
    auto a = h->args(backend::CO_STEP_SYMMETRIC);
    a.partial = scr;
    a.result = scr + 2 * max_blocks;
    a.a = dz; a.b = ds; a.c = z; a.d = s; a.s0 = alpha_max;
    h->run(a, stream);
    backend::launch_cones_reduce(a, h->nblocks, 0, stream);
    if (!h->st->symmetric) {
      a.op = backend::CO_STEP_NONSYMMETRIC;  // backtracks from min(alpha, max_step_fraction) in result[0]
      h->run(a, stream);
      backend::launch_cones_reduce(a, h->nblocks, 0, stream);
    }
    backend::d2d(alpha_dev, a.result, sizeof(double), stream);
    return 0;
  })
}

int clh_cones_compute_barrier(clh_cones* h, const double* z, const double* s, const double* dz, const double* ds,
                              double alpha, double* barrier_dev, void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    if (!z || !s || !dz || !ds || !barrier_dev) throw std::invalid_argument("null vector");
    auto a = h->args(backend::CO_BARRIER);
    a.a = z; a.b = s; a.c = dz; a.d = ds; a.s0 = alpha;
    h->run(a, stream);
    h->reduce(a, 1, barrier_dev, 1, stream);
    return 0;
  })
}

int clh_cones_margins(clh_cones* h, const double* z, double* out_dev, void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    if (!z || !out_dev) throw std::invalid_argument("null vector");
    auto a = h->args(backend::CO_MARGINS);
    a.a = z;
    h->run(a, stream);
    h->reduce(a, 2, out_dev, 2, stream);
    return 0;
  })
}

int clh_cones_scaled_unit_shift(clh_cones* h, double* z, double alpha, int32_t primal_cone, void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    if (!z) throw std::invalid_argument("null vector");
    auto a = h->args(backend::CO_UNIT_SHIFT);
    a.x = z; a.s0 = alpha; a.flag = primal_cone ? 1 : 0;
    h->run(a, stream);
    return 0;
  })
}

int clh_cones_unit_initialization(clh_cones* h, double* z, double* s, void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    if (!z || !s) throw std::invalid_argument("null vector");
    auto a = h->args(backend::CO_UNIT_INIT);
    a.x = z; a.y = s;
    h->run(a, stream);
    return 0;
  })
}

int clh_cones_rectify_equilibration(clh_cones* h, double* delta, const double* e, int32_t* any_changed,
                                    void* stream) {
  CLH_TRY({
    CLH_CONES_CHECK(h);
    if (!delta || !e) throw std::invalid_argument("null vector");
    auto a = h->args(backend::CO_RECTIFY);
    a.x = delta; a.a = e;
    h->run(a, stream);
    const Structure& S = *h->st;
    if (any_changed)
      *any_changed = (S.soc_off.size() + S.exp_off.size() + S.pow_off.size() + S.psd_off.size()) > 0 ? 1 : 0;
    return 0;
  })
}

void clh_cones_destroy(clh_cones* h) { delete h; }

}  // extern "C"

