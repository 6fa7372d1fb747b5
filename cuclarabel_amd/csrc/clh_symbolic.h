// Host-side structural analysis of one problem shape: KKT pattern, value
// maps, fill-reducing ordering and the supernodal symbolic factorization.
// Computed once per sparsity pattern + cone layout and shared by every
// request of that shape.
#pragma once
#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#include "clh_types.h"

namespace clh {

struct ConeSpec {
  int type;    // CT_*
  int dim;     // numel (rows): the dimension for zero/nonneg/soc, 3 for exp/pow, n(n+1)/2 for PSD
  double param;
  int side = 0;  // PSD: the matrix side n
};

// The analysis' schedule knobs (chosen per problem-size class by the
// autotuner): relaxed supernodes -- a column joins its child's panel when the
// panel then holds at most relax_max_cols columns (0: fundamental supernodes
// only) and at most relax_zero_pct percent of explicit zeros -- and the lanes
// per supernode group (0: by panel heights and level widths, else 16, 32 or
// 64 at every level).
struct SymOpts {
  int relax_max_cols = 128;
  int relax_zero_pct = 25;
  int group_lanes = 0;
};

struct Structure {
  int n = 0, m = 0, p = 0, N = 0;
  SymOpts opts;
  // the caller's patterns and cones as given (P and A CSC), so the same shape
  // can be analysed again under another schedule
  std::vector<int32_t> uPp, uPi, uAp, uAi;
  std::vector<ConeSpec> ucones;
  bool fp32_order = false;
  int degree = 0;
  bool symmetric = true;
  uint64_t key = 0;  // hash of everything structural

  // user P (CSC): indices of the upper-triangle entries kept, and their
  // positions in the full symmetric CSR (second position -1 on the diagonal)
  std::vector<int32_t> Puser_keep;   // user nz index of each kept entry
  std::vector<int32_t> Pfull_pos0, Pfull_pos1;
  std::vector<int32_t> Pp, Pi;       // full symmetric CSR
  std::vector<int32_t> Pk;           // per full entry -> KKT index or -1
  // A: CSR and A' CSR (= CSC); csc_to_csr[k] = CSR position of CSC entry k
  std::vector<int32_t> Ap, Ai, Atp, Ati, csc_to_csr;
  std::vector<int32_t> Ak;           // per CSR entry -> KKT index

  // cones
  std::vector<int8_t> etype;
  std::vector<int32_t> lin_kdiag;
  std::vector<int32_t> soc_off, soc_dim, soc_kblk, soc_sp, soc_kidx;
  std::vector<int32_t> sp_ku, sp_kv, sp_kD, sp_off;
  std::vector<int32_t> exp_off, exp_kidx, pow_off, pow_kidx;
  std::vector<double> pow_alpha;
  // PSD cones: row offset, side n, offset of the t(t+1)/2 KKT indices of the
  // dense Hs block (t = n(n+1)/2), offsets of the state (R, Rinv: n^2 each;
  // lambda, lambda^{-1/2}: n each) and of the work area (8 n^2)
  std::vector<int32_t> psd_off, psd_n, psd_kblk, psd_kidx, psd_soff, psd_woff;
  int psd_state = 0, psd_work = 0;   // doubles

  // KKT (triu CSC, reference layout) and its full CSR for refinement products
  int nnzK = 0;
  std::vector<int32_t> Kp, Ki;       // triu CSC pattern
  std::vector<int32_t> Kdiag;
  std::vector<int8_t> dsign;
  std::vector<int8_t> weak;          // pivots that start as the bare regularization: equality rows, variables without a P diagonal
  std::vector<int32_t> Kfp, Kfi, Kfk;

  // supernodal LDL' symbolic (permuted ordering)
  std::vector<int32_t> perm;
  int nsn = 0, nlev = 0;
  std::vector<int32_t> sn_col, sn_rowptr, sn_rows;
  std::vector<int64_t> sn_lptr;
  std::vector<int64_t> sn_iptr;      // offsets of the diagonal-block inverses (ns^2 for ns > 1)
  int maxns = 1;
  int group_hint = 64;               // lanes per supernode group suited to the panel heights
  std::vector<int32_t> lev_ptr, lev_sn;
  std::vector<int32_t> lev_group;    // lanes per supernode group at each level
  std::vector<int32_t> upd_ptr, upd_src, upd_roff, upd_ncol;
  std::vector<SnDesc> sn;            // packed per-supernode records
  std::vector<UpdDesc> upd;          // packed per-update records
  std::vector<int32_t> upd_rel;      // for each update, its rows' positions in the target panel
  std::vector<SnDesc> sn_lev;        // sn in level order
  std::vector<int32_t> rowupd_ptr;   // per panel row (position in sn_rows): range in rowupd
  std::vector<RowUpd> rowupd;        // update contributions per panel row, in update order
  std::vector<int32_t> asm_ptr, asm_dst, asm_src;
  std::vector<int8_t> asm_sgn;
  int64_t nnzL = 0;                  // strictly-lower entries of L
  double flops = 0;                  // sum_k c_k (c_k + 1), c_k = off-diagonal count of column k
  int64_t Lsize = 0;                 // panel entries (incl. diagonal blocks)
  int nleaf = 0;                     // single-column supernodes leading level 0 (one thread each)
  double path_cost = 0;              // estimated critical path of one factorization, in
                                     // sequential wavefront steps (level-synchronous schedule)
};

// Build the structure.  P is CSC n x n (only entries with row <= col are
// used), A is CSC m x n.  Throws std::runtime_error on invalid input.
// fp32_order: the elimination order for FP32 factors, in which no pivot row
// with a negative sign is eliminated before one of its neighbours.
std::shared_ptr<Structure> build_structure(int n, int m, const int32_t* Pp, const int32_t* Pi, const int32_t* Ap,
                                           const int32_t* Ai, const std::vector<ConeSpec>& cones, bool fp32_order,
                                           const SymOpts& opts = SymOpts());

// Structure of a standalone symmetric matrix (the LDL' plugin): a full
// symmetric CSR (index base 0 or 1) with pivot signs +1/-1.  Only the KKT and
// LDL' parts are filled; csr_of_K[k] is the CSR position of triu CSC entry k.
std::shared_ptr<Structure> build_ldl_structure(int N, const int32_t* rowptr, const int32_t* colval, int base,
                                               const int32_t* dsigns, std::vector<int32_t>& csr_of_K,
                                               bool fp32_order, const SymOpts& opts = SymOpts());

// structural hash used by the per-shape cache
uint64_t structure_key(int n, int m, const int32_t* Pp, const int32_t* Pi, const int32_t* Ap, const int32_t* Ai,
                       const std::vector<ConeSpec>& cones, bool fp32_order, const SymOpts& opts = SymOpts());

}  // namespace clh
