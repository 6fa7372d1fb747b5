// Types shared by the host library and the device engine.
//
// A "slot" is one problem resident in HBM.  Everything the device engine
// needs for that problem is reached through one SlotDesc (device pointers
// into the slot's arena, laid out by the host at admission time).  The
// workgroup engine runs one slot per 256-thread workgroup; the lane engine
// runs a batch of up to 64 same-shape slots per 64-lane workgroup, one slot
// per lane.
#pragma once
#include <stdint.h>

namespace clh {

enum ConeCode { CT_ZERO = 0, CT_NONNEG = 1, CT_SOC = 2, CT_EXP = 3, CT_POW = 4, CT_PSD = 5 };

// status codes: numbering of src/statuscodes.jl:25-37
enum StatusCode {
  ST_UNSOLVED = 0, ST_SOLVED, ST_PRIMAL_INFEASIBLE, ST_DUAL_INFEASIBLE, ST_ALMOST_SOLVED,
  ST_ALMOST_PRIMAL_INFEASIBLE, ST_ALMOST_DUAL_INFEASIBLE, ST_MAX_ITERATIONS, ST_MAX_TIME,
  ST_NUMERICAL_ERROR, ST_INSUFFICIENT_PROGRESS
};

// phase of a slot's state machine
enum Phase { PH_SETUP = 0, PH_START = 1, PH_ITER = 2, PH_DONE = 3 };

// Solver settings: the fields of src/settings.jl:70-148 that the hot path
// reads, with the same names and meaning.
struct DevSettings {
  int32_t max_iter;
  int32_t equilibrate_enable;
  int32_t equilibrate_max_iter;
  int32_t static_regularization_enable;
  int32_t dynamic_regularization_enable;
  int32_t iterative_refinement_enable;
  int32_t iterative_refinement_max_iter;
  int32_t grid_rowwise;  // diagnostics: the grid engine updates wide supernodes row by row, not by tiles
  double time_limit;
  double max_step_fraction;
  double tol_gap_abs, tol_gap_rel, tol_feas, tol_infeas_abs, tol_infeas_rel, tol_ktratio;
  double reduced_tol_gap_abs, reduced_tol_gap_rel, reduced_tol_feas;
  double reduced_tol_infeas_abs, reduced_tol_infeas_rel, reduced_tol_ktratio;
  double equilibrate_min_scaling, equilibrate_max_scaling;
  double linesearch_backtrack_step, min_switch_step_length, min_terminate_step_length;
  double static_regularization_constant, static_regularization_proportional;
  double dynamic_regularization_eps, dynamic_regularization_delta;
  double iterative_refinement_reltol, iterative_refinement_abstol, iterative_refinement_stop_ratio;
  // mixed precision: refinement that stalls above mixed_switch_reltol * max(1, |b|_inf)
  // moves the slot to FP64 factors with static regularizer mixed_fp64_static_reg
  double mixed_switch_reltol;
  double mixed_fp64_static_reg;
  // the grid engine's tile schedule: supernodes of at least tile_min_cols
  // columns take their updates as dense tiles, panels of at least
  // dense_min_cols columns are factored block by block over the whole grid
  // (0: the compiled defaults)
  int32_t tile_min_cols;
  int32_t dense_min_cols;
  // a wrongly signed pivot larger than dynamic_regularization_delta keeps
  // its magnitude (0: the reference's rule, sign * delta)
  int32_t dyn_keep_magnitude;
};

// Packed symbolic records read by the factorization and the solves: one
// wide load gives everything a supernode or an update needs.
struct SnDesc {        // one supernode
  int32_t f, ns;       // first column, columns
  int32_t r0, nr;      // offset in sn_rows, rows (diagonal block included)
  int32_t asm0, asm1;  // assembly list range
  int32_t upd0, upd1;  // update list range (descendants updating it)
  int64_t lptr, iptr;  // panel offset, diagonal-block inverse offset
};
struct UpdDesc {       // descendant d updating target s
  int32_t src;         // d
  int32_t roff;        // first row of d's panel inside s's columns
  int32_t ncol;        // rows of d inside s's columns (columns of s touched)
  int32_t fd, nsd;     // d's first column and column count
  int32_t nrd;         // d's row count (panel leading dimension)
  int32_t rel;         // offset in upd_rel of the positions of d's rows roff.. in s's panel
  int32_t pad;
  int64_t lptr;        // d's panel offset
};

struct RowUpd {        // one contribution to one row of a target panel
  int64_t base;        // descendant panel offset + roff (the descendant's rows in this target)
  int32_t a;           // the descendant row (relative to roff) that lands on this target row
  int32_t nsd, nrd;    // descendant columns, leading dimension
  int32_t fd;          // descendant's first column (its pivots)
  int32_t bmax;        // min(a + 1, ncol): target columns updated
  int32_t rel;         // offset in upd_rel of this update's row positions
};

// Per-slot scalar state, persisted in HBM between launches.
struct SlotScalars {
  double tau, kappa;                 // current iterate
  double ptau, pkappa;               // previous iterate (rollback snapshot)
  double dtau, dkappa;               // step (lhs)
  double rtau_rhs, rkappa_rhs;       // step rhs
  double rtau, dot_qx, dot_bz, dot_sz, dot_xPx;
  double mu, sigma, alpha;
  double c, normq, normb;            // equilibration cost scale, unscaled norms
  double cost_primal, cost_dual, res_primal, res_dual, res_primal_inf, res_dual_inf;
  double gap_abs, gap_rel, ktratio;
  double prev_cost_primal, prev_cost_dual, prev_res_primal, prev_res_dual, prev_gap_abs, prev_gap_rel;
  double obj_val, obj_val_dual;
  double static_eps;                 // last static regularizer
  double time_left;                  // seconds the solve may run from its first pass (time_limit
                                     // less setup and queueing; +inf: no limit); < 0 ends it
  uint64_t clock0;                   // device clock (100 MHz ticks) at the solve's start
  int32_t iter;                      // IPM iteration counter (counts KKT updates)
  int32_t info_iterations;
  int32_t status;
  int32_t scaling;                   // 0 PrimalDual, 1 Dual
  int32_t phase;
  int32_t ir_steps;                  // IR refinement solves taken (diagnostic)
  int32_t n_dynreg;                  // dynamically regularized pivots (last factor)
  int32_t kkt_solves;                // KKT solves (diagnostic)
  int32_t trace_len;                 // rows written to SlotDesc::trace
  int32_t reuse_equil;               // setup phase: keep d, e, c of the previous solve (data updates)
  int32_t prec64;                    // mixed precision: 1 once the slot factors in FP64
  int32_t n_fp32_factors;            // factorizations done in FP32 (diagnostic)
};

// Workgroup engine inbox record (host-mapped memory, read by the kernel at a
// freshly admitted slot's first launch): the initial SlotScalars, padded to
// kInboxHeaderDoubles doubles, then the values Px, Ax, Atx, q, b.
constexpr int kInboxHeaderDoubles = (int)((sizeof(SlotScalars) + 63) / 64 * 8);

// Pointer kinds of the per-problem value arrays.  The host and the
// workgroup engine use raw pointers; the lane-batched engine (one problem per
// lane of a 64-lane wavefront, all lanes of one batch sharing one structure)
// reads the same fields as lane-interleaved arrays: element i of lane l lives
// at base[i * kBatchLanes + l].  Both kinds are one pointer wide, so one
// descriptor layout serves both.
constexpr int kBatchLanes = 64;
template <class T> using RawPtr = T*;

// Everything the device engine needs for one slot (workgroup engine) or one
// batch of same-shape slots (lane engine).
template <template <class> class VP>
struct SlotDescT {
  // ---- sizes
  int32_t n, m, p, N;
  int32_t nnzP, nnzA, nnzK, nnzKf;
  int32_t nsoc, nexp, npow, nsparse;
  int32_t degree, symmetric, nsn, nlev;
  int32_t factor_fp32;               // 1: FP32 panels + FP32 solves, FP64 refinement
  int32_t ngroups;                   // supernode lane groups per workgroup
  int32_t nleaf;                     // single-column supernodes leading level 0
  int32_t group;                     // lanes per supernode group (16, 32 or 64; 1 in serial builds)
  int32_t maxns;                     // widest supernode (columns)
  int32_t npsd;                      // PSD cones
  // ---- per-lane strides (elements) of the lane-contiguous arrays; lane engine only
  int32_t inbox_lstride, out_lstride, trace_lstride, pad2;
  // ---- problem data (raw on upload; scaled in place by the setup phase)
  const int32_t *Pp, *Pi;  VP<double> Px;   // full symmetric P, CSR
  const int32_t *Ap, *Ai;  VP<double> Ax;   // A, CSR (rows)
  const int32_t *Atp, *Ati; VP<double> Atx; // A', CSR (= A in CSC)
  VP<double> q, b;
  VP<double> d, dinv, e, einv;              // equilibration
  // lane engine: each lane's uploaded values, contiguous per lane in the
  // order Px, Ax, Atx, q, b (copied into the interleaved arrays at setup).
  // workgroup engine: null, or a freshly admitted slot's inbox record
  const double* inbox;
  // ---- cones
  const int8_t* etype;                       // per row of A: CT_ZERO / CT_NONNEG / other
  const int32_t* lin_kdiag;                  // per row: KKT index of its diagonal (zero/nonneg)
  const int32_t *soc_off, *soc_dim, *soc_kblk, *soc_sp;  // soc_sp: sparse index or -1
  const int32_t* soc_kidx;                   // dense: tri(dim) KKT indices; sparse: dim diag indices
  const int32_t* sp_ku;  const int32_t* sp_kv;  const int32_t* sp_kD;  // sparse SOC maps
  const int32_t* sp_off;                     // offset of each sparse SOC's u/v within sp_ku/sp_kv
  const int32_t *exp_off, *exp_kidx;         // 6 KKT indices per cone (packed triu 3x3)
  const int32_t *pow_off, *pow_kidx;
  const double* pow_alpha;
  const int32_t *psd_off, *psd_n, *psd_kblk, *psd_kidx;  // PSD: rows, side, dense Hs block KKT indices
  const int32_t *psd_soff, *psd_woff;        // PSD: offsets in psd_st / psd_wk
  VP<double> psd_st;                         // per PSD cone: R, Rinv (n^2 each), lambda, lambda^{-1/2} (n each)
  VP<double> psd_wk;                         // per PSD cone: 8 n^2 work
  VP<double> w, lam;                         // per row (nonneg, soc)
  VP<double> soc_eta, soc_d;                 // per soc
  VP<double> sp_u, sp_v;                     // per sparse soc element
  VP<double> nsym;                           // per exp/pow cone: Hd[9] Hs[9] grad[3] zc[3]
  // ---- KKT (triu CSC values in the original ordering)
  VP<double> Kx;
  const int32_t* Kdiag;                      // N
  const int32_t* Pk;                         // per full-P entry (row j, col <= j): KKT index, else -1
  const int32_t* Ak;                         // per A CSR entry: KKT index
  const int8_t* dsign;                       // N
  const int32_t *Kfp, *Kfi, *Kfk;            // full symmetric CSR of KKT -> index into Kx
  // ---- LDL' symbolic (permuted ordering)
  const int32_t* perm;                       // perm[k] = original index of pivot k
  const int32_t* sn_rows;                    // row lists of the supernodes (SnDesc::r0, nr)
  const int32_t* lev_ptr;                    // nlev+1
  const int32_t* lev_group;                  // nlev: lanes per supernode group at each level
  const int32_t* lev_sn;                     // nsn
  const SnDesc* sn;                          // nsn packed supernode records
  const SnDesc* sn_lev;                      // the same in level order (lev_sn)
  const int32_t* rowupd_ptr;                 // per panel row (sn_rows position): range in rowupd
  const RowUpd* rowupd;                      // contributions per panel row, in update order
  const UpdDesc* upd;                        // packed update records, grouped by target
  const int32_t* upd_rel;                    // row positions inside the target panel
  const int32_t *asm_dst, *asm_src;          // assembly lists (SnDesc::asm0, asm1)
  const int8_t* asm_sgn;                     // +1/-1 on diagonal entries (regularization sign), else 0
  void* L;                                   // panels (float or double)
  void* Dv;                                  // pivots, permuted (float or double)
  void* ly;                                  // triangular-solve work vector (float or double), N
  void* Linv;                                // inverses of the unit-lower diagonal blocks
  void* gscr;                                // per-group scratch, maxns entries each
  // ---- iterate and work vectors (FP64)
  VP<double> x, s, z;
  VP<double> dx, ds, dz;                     // step lhs
  VP<double> hx, hs, hz;                     // step rhs
  VP<double> px, ps, pz;                     // previous iterate
  VP<double> rx, rz, rx_inf, rz_inf, Pxv;
  VP<double> x1, z1, x2, z2, workx, workz, wconic;
  VP<double> kb, kx, kw1, kw2;               // N
  // ---- outputs (unscaled solution), written when the slot finishes;
  // lane-contiguous (lane l at + l * out_lstride)
  double *out_x, *out_s, *out_z;
  double* trace;                             // (max_iter+2) x 8 per-iteration record per lane, may be null
  uint64_t* prof;                            // 8 phase timers per lane (device wall-clock ticks), may be null
  SlotScalars* sc;                           // one per lane
  SlotScalars* out_sc;                       // may be null: the final scalars, written when the slot finishes
  const DevSettings* st;
  // the grid engine: a word in host-mapped memory that ends the launch when
  // set (by a grid barrier that waited too long, or by the host); null: none
  int32_t* abort_flag;
  // the device-resident request loop: a word in host-mapped memory the host
  // sets to cancel this request (read at each termination check); null: none
  int32_t* cancel_flag;
};
using SlotDesc = SlotDescT<RawPtr>;

}  // namespace clh
