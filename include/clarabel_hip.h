/*
 * clarabel_hip.h -- C ABI of libclarabel_hip.so, the MI355X (gfx950) engine
 * behind direct_solve_method = :hip / :hipmixed.
 *
 * Conventions
 *  - 0-based int32 compressed-sparse-column (CSC) matrices, FP64 values,
 *    host pointers unless a name ends in _dev.
 *  - Return value: 0 ok, < 0 usage error (bad argument, unsupported input),
 *    > 0 runtime/device error.  clh_last_error() gives the message of the last
 *    failing call on the calling thread.  No exception crosses the ABI.
 *  - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *  - A handle is not thread-safe; distinct handles are independent.
 *  - Solutions are those of Clarabel's solve!: (x, s, z) of
 *      min 1/2 x'Px + q'x  s.t.  Ax + s = b,  s in K.
 *
 * Reference interfaces each group replaces (paths relative to the
 * CuClarabel repository):
 *  - settings / status / info : src/settings.jl:70-148, src/statuscodes.jl:25-37,
 *                               src/types.jl:227-265, :297-332
 *  - solver                   : Clarabel.Solver(P,q,A,b,cones,settings) + solve!
 *                               (src/solver.jl:5-70, :129-195, :231-422)
 *  - data updates             : update_P!/update_q!/update_A!/update_b!
 *                               (src/data_updating.jl:26-202)
 *  - serving                  : no counterpart in the reference (its GPU path
 *                               solves one problem at a time)
 */
#ifndef CLARABEL_HIP_H
#define CLARABEL_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cone codes, in the order of src/cones/cone_api.jl:24-56 */
enum { CLH_ZERO = 0, CLH_NONNEG = 1, CLH_SOC = 2, CLH_EXP = 3, CLH_POW = 4, CLH_PSD = 5 };

/* status codes, same numbering as src/statuscodes.jl:25-37 */
enum {
  CLH_UNSOLVED = 0, CLH_SOLVED, CLH_PRIMAL_INFEASIBLE, CLH_DUAL_INFEASIBLE, CLH_ALMOST_SOLVED,
  CLH_ALMOST_PRIMAL_INFEASIBLE, CLH_ALMOST_DUAL_INFEASIBLE, CLH_MAX_ITERATIONS, CLH_MAX_TIME,
  CLH_NUMERICAL_ERROR, CLH_INSUFFICIENT_PROGRESS
};

/* direct_solve_method values (src/Clarabel.jl:27 lists :cudss, :cudssmixed) */
enum {
  CLH_METHOD_HIP = 0,      /* FP64 LDL' and solves: the drop-in for :cudss          */
  CLH_METHOD_HIPMIXED = 1  /* FP32 LDL' and solves + FP64 refinement: for :cudssmixed */
};

/* Schedule: the engine's launch and tile choices (no counterpart in the
   reference, whose cuDSS picks its own).  Zero in a field keeps the built-in
   rule.  clh_server_autotune / clh_solver_autotune pick them by timing
   candidates on the device and remember the winners per device and
   problem-size class in a cache file ($CLH_TUNE_CACHE, else
   $HOME/.cache/clarabel_hip/schedules.txt); use_tuned = 1 makes a new solver
   or server take the cached choices where the cache has them. */
typedef struct {
  int32_t relax_max_cols;   /* relaxed supernodes: most columns of a panel (default 128; -1: none) */
  int32_t relax_zero_pct;   /* ... and most explicit zeros in it, percent (default 25)            */
  int32_t group_lanes;      /* lanes per supernode group: 16, 32 or 64 at every level
                               (0: by panel heights and level widths)                             */
  int32_t blocks_per_cu;    /* serving, resident loop: workgroups per CU (0: as many as fit)      */
  int32_t grid_blocks;      /* single problem on the grid engine: workgroups of the launch
                               (0: 512, or 1024 from 1e12 factorization flops)                     */
  int32_t tile_min_cols;    /* grid engine: supernodes this wide take their updates as tiles (128) */
  int32_t dense_min_cols;   /* grid engine: panels this wide are factored over the grid (128)     */
  int32_t use_tuned;        /* 1: take the cached tuned schedule for this device and size class   */
} clh_schedule;

/* What an autotuning run found: the rate of the default schedule and of the
   chosen one on the same work (solves/s for a server, IPM iterations/s for a
   single solver), the trials timed and the seconds they took. */
typedef struct {
  double default_rate;
  double tuned_rate;
  int32_t trials;
  int32_t classes;          /* problem-size classes tuned (servers: one per power of two of n + m) */
  double seconds;
} clh_tune_report;

/* Settings: every field of src/settings.jl:70-148 that the GPU path uses,
   with the same names and defaults (clh_default_settings). */
typedef struct {
  int32_t max_iter;
  double time_limit;
  int32_t verbose;
  double max_step_fraction;
  double tol_gap_abs, tol_gap_rel, tol_feas, tol_infeas_abs, tol_infeas_rel, tol_ktratio;
  double reduced_tol_gap_abs, reduced_tol_gap_rel, reduced_tol_feas;
  double reduced_tol_infeas_abs, reduced_tol_infeas_rel, reduced_tol_ktratio;
  int32_t equilibrate_enable;
  int32_t equilibrate_max_iter;
  double equilibrate_min_scaling, equilibrate_max_scaling;
  double linesearch_backtrack_step, min_switch_step_length, min_terminate_step_length;
  int32_t static_regularization_enable;
  double static_regularization_constant;     /* :hipmixed always uses sqrt(eps(Float32)),
                                                 overriding this field, as gpu_preprocess
                                                 does for :cudssmixed (src/solver.jl:604-615) */
  double static_regularization_proportional;
  int32_t dynamic_regularization_enable;
  double dynamic_regularization_eps, dynamic_regularization_delta;
  int32_t iterative_refinement_enable;
  double iterative_refinement_reltol, iterative_refinement_abstol;
  int32_t iterative_refinement_max_iter;
  double iterative_refinement_stop_ratio;
  int32_t direct_solve_method;               /* CLH_METHOD_* */
  int32_t soc_threshold;                     /* src/settings.jl:147; SOCs above it use the
                                                 sparse expansion (only 5 is supported) */
  /* :hipmixed only (no counterpart in the reference): when refinement with
     the FP32 factors stalls above mixed_switch_reltol * max(1, |rhs|_inf),
     the slot refactors in FP64 (static regularizer mixed_fp64_static_reg)
     for the rest of the solve */
  double mixed_switch_reltol;
  double mixed_fp64_static_reg;
  /* no counterpart in the reference: the engine a problem runs on.
     Serving: CLH_ENGINE_LANES batches up to 64 requests of one shape in a
     64-lane workgroup, one request per lane; CLH_ENGINE_WORKGROUP gives each
     request a 256-thread workgroup, advanced by rounds of launches;
     CLH_ENGINE_RESIDENT (below) runs those workgroups in one persistent
     launch fed by a queue; CLH_ENGINE_AUTO uses CLH_ENGINE_RESIDENT (the
     fastest measured on MI355X for the MPC stream).  Single-problem solver: CLH_ENGINE_GRID runs the problem over a
     cooperative grid of workgroups filling the GPU, CLH_ENGINE_WORKGROUP on
     one workgroup; CLH_ENGINE_AUTO picks the grid when the factorization
     costs at least 1e8 flops. */
  int32_t serve_engine;
  /* no counterpart in the reference: 1 (default) lets a pivot that dynamic
     regularization replaces keep its magnitude when that is above
     dynamic_regularization_delta (sign * max(delta, |d|)); 0 applies the
     reference's rule, sign * delta, always (QDLDL through
     src/kktsolvers/direct-ldl/directldl_qdldl.jl:62-86).  The first keeps
     a pivot that cancellation pushed across zero near the end of a solve
     from overflowing the columns coupled to it. */
  int32_t dynamic_regularization_keep_magnitude;
  clh_schedule schedule;                     /* no counterpart in the reference (above) */
} clh_settings;

enum { CLH_ENGINE_AUTO = 0, CLH_ENGINE_WORKGROUP = 1, CLH_ENGINE_LANES = 2, CLH_ENGINE_GRID = 3,
       CLH_ENGINE_RESIDENT = 4 };
/* CLH_ENGINE_RESIDENT (serving): a persistent launch whose workgroups take
   requests from a queue in host memory as they arrive and solve each one to
   completion (the device-resident request loop; one 256-thread workgroup
   per request, as CLH_ENGINE_WORKGROUP).  A single-problem solver treats it
   as CLH_ENGINE_WORKGROUP. */

/* Information of a finished solve (DefaultInfo / DefaultSolution fields) */
typedef struct {
  int32_t status;
  int32_t iterations;
  double obj_val, obj_val_dual;
  double r_prim, r_dual;
  double solve_time;            /* seconds, device-side solve (excludes setup)      */
  double setup_time;            /* seconds, host structure analysis + upload         */
  double gap_abs, gap_rel, ktratio, mu;
  int32_t ir_steps;             /* refinement solves taken in total                  */
  int32_t kkt_solves;
  int64_t nnz_L;                /* strictly lower entries of the symbolic factor     */
  double factor_flops;          /* sum_k c_k (c_k + 1)                               */
  int32_t fp32_factors;         /* factorizations done with FP32 panels              */
  int32_t switched_fp64;        /* 1 if a :hipmixed solve moved to FP64 factors      */
  int64_t nnz_K;                /* stored (upper-triangle) entries of the KKT matrix */
  int32_t kkt_dim;              /* n + m + p                                         */
  int32_t n_supernodes;         /* supernodes of the symbolic factor                 */
  int32_t n_levels;             /* levels of the supernodal elimination tree         */
  int32_t pad;
  double phase_seconds[8];      /* device time by phase (clh_solve only): setup,
                                   residuals+info, cone scaling, factorization,
                                   triangular solves, refinement products, step
                                   length, whole IPM passes */
} clh_info;

typedef struct clh_solver clh_solver;
typedef struct clh_server clh_server;

const char* clh_last_error(void);
const char* clh_version(void);
int clh_device_info(char* name, size_t len, int32_t* num_cu);

int clh_default_settings(clh_settings* s);

/* Clarabel.Solver(P,q,A,b,cones,settings) -- src/solver.jl:5-17, setup! :129-195.
   P: n x n CSC (entries below the diagonal are ignored, as triu(P) in
   src/problemdata.jl:19-23); A: m x n CSC.  cone_param holds alpha of power
   cones (ignored otherwise).  cone_dims: numel of zero/nonneg/soc cones, 3 for
   exp/pow.  Structure analysis runs here, once. */
int clh_solver_create(clh_solver** out, int32_t n, int32_t m,
                      const int32_t* Pp, const int32_t* Pi, const double* Px,
                      const double* q,
                      const int32_t* Ap, const int32_t* Ai, const double* Ax,
                      const double* b,
                      int32_t ncones, const int32_t* cone_type, const int32_t* cone_dim,
                      const double* cone_param, const clh_settings* settings);

/* solve!(solver) -- src/solver.jl:231-422.  Blocks until the solve ends. */
int clh_solve(clh_solver* s, void* stream);

/* solution and info of the last solve; x[n], s[m], z[m] host buffers (may be NULL);
   trace: (iterations+1) x 8 rows of the verbose table
   (pcost dcost gap pres dres k/t mu step), may be NULL, max_rows bounds it. */
int clh_get_solution(clh_solver* s, double* x, double* sv, double* z, clh_info* info);
int clh_get_trace(clh_solver* s, double* trace, int32_t max_rows, int32_t* rows);

/* update_P! / update_q! / update_A! / update_b! (src/data_updating.jl:26-202):
   new values for the same sparsity pattern (P: values of the entries given at
   creation, in the same order), kept in the equilibration of the first solve.
   update_P follows the CPU path and applies the cost scaling. */
int clh_update_P(clh_solver* s, const double* Px);
int clh_update_q(clh_solver* s, const double* q);
int clh_update_A(clh_solver* s, const double* Ax);
int clh_update_b(clh_solver* s, const double* b);

void clh_solver_destroy(clh_solver* s);

/* Device-pointer variants, the counterpart of the reference's Python bridge
   (src/python/py2jl.jl:1-26 wraps CuPy pointers as CuArrays for the solver;
   src/python/jl2py.py hands CuVectors back as CuPy arrays): the problem's
   arrays (0-based int32 CSC patterns, FP64 values) in device memory; the
   cone description stays on the host.  The patterns are copied to the host
   for the structure analysis; results are equal to the host-array entry
   points' on the same data. */
int clh_solver_create_device(clh_solver** out, int32_t n, int32_t m,
                             const int32_t* Pp_dev, const int32_t* Pi_dev, const double* Px_dev,
                             const double* q_dev,
                             const int32_t* Ap_dev, const int32_t* Ai_dev, const double* Ax_dev,
                             const double* b_dev,
                             int32_t ncones, const int32_t* cone_type, const int32_t* cone_dim,
                             const double* cone_param, const clh_settings* settings);
/* the last solve's x[n], s[m], z[m] copied device to device (NULL skips one) */
int clh_get_solution_device(clh_solver* s, double* x_dev, double* s_dev, double* z_dev, void* stream);
int clh_update_P_device(clh_solver* s, const double* Px_dev);
int clh_update_q_device(clh_solver* s, const double* q_dev);
int clh_update_A_device(clh_solver* s, const double* Ax_dev);
int clh_update_b_device(clh_solver* s, const double* b_dev);

/* ---- serving (no counterpart in the reference) ------------------------------
   A server owns `max_slots` workgroup slots in HBM.  Requests are admitted as
   slots and the HBM budget free up; active slots advance together in
   iteration launches, so solves of different sizes overlap on the device.
   Every result is bitwise the same as the same request solved alone. */
typedef struct {
  int64_t submitted, completed, rejected;
  int32_t active_slots, queued;
  int64_t hbm_in_use;
  double p50_latency_s, p99_latency_s, mean_latency_s;
  double busy_time_s;
  int64_t launches;
  double kernel_time_s;         /* sum of the iteration launches' device durations */
  int64_t cancelled;            /* requests ended by clh_server_cancel             */
} clh_server_stats;

int clh_server_create(clh_server** out, int64_t hbm_budget_bytes, int32_t max_slots,
                      const clh_settings* settings, void* stream);
/* queue one problem (same arguments as clh_solver_create); the data are copied */
int clh_server_submit(clh_server* srv, int64_t* ticket, int32_t n, int32_t m,
                      const int32_t* Pp, const int32_t* Pi, const double* Px,
                      const double* q,
                      const int32_t* Ap, const int32_t* Ai, const double* Ax,
                      const double* b,
                      int32_t ncones, const int32_t* cone_type, const int32_t* cone_dim,
                      const double* cone_param);
/* drive the device: admit, launch `passes` IPM loop passes on every active
   slot, harvest finished ones.  Returns the number still in flight. */
int clh_server_progress(clh_server* srv, int32_t passes);
/* the same without waiting for the launch: it runs on the device while the
   caller continues (submitting more requests, for instance); the next
   progress, wait or cancel call finishes it and harvests its results */
int clh_server_progress_async(clh_server* srv, int32_t passes);
/* 1 if finished (result available), 0 if not, < 0 unknown ticket */
int clh_server_poll(clh_server* srv, int64_t ticket);
/* drive the device (clh_server_progress with `passes`) until the ticket has
   finished or timeout_s seconds have passed: 1 finished, 0 timed out, < 0
   unknown ticket */
int clh_server_wait(clh_server* srv, int64_t ticket, int32_t passes, double timeout_s);
/* cancel a request: a queued one finishes at once with status MAX_TIME and
   no iterations; a resident one is given a negative time limit and finishes
   at its next termination check, with MAX_TIME (or the ALMOST_* status its
   iterate earns under the reduced tolerances, as for any time limit).
   0 cancelled, 1 already finished, < 0 unknown ticket.  Each request's own
   time limit is settings.time_limit counted from its submission. */
int clh_server_cancel(clh_server* srv, int64_t ticket);
/* copy a finished result out and release it */
int clh_server_result(clh_server* srv, int64_t ticket, double* x, double* s, double* z, clh_info* info,
                      double* latency_s);
int clh_server_stats_get(clh_server* srv, clh_server_stats* st);
void clh_server_destroy(clh_server* srv);

/* ---- autotuning (no counterpart in the reference, whose cuDSS schedules
   itself) -------------------------------------------------------------------
   clh_server_autotune tunes on the requests queued in an idle server (submit
   a representative sample first, several times as many requests as the
   server keeps resident; they stay queued and are served by the next
   progress calls under the chosen schedule).  For each size class (n + m in
   [2^k, 2^(k+1))) it times the supernode relaxation and the group lanes by
   coordinate descent, each candidate by serving copies of the whole sample
   at once (results dropped); then the resident workgroups per CU.  clh_solver_autotune tunes the grid engine's workgroups and
   tile thresholds on the solver's own problem, each candidate timed over a
   solve capped at 3 iterations; it must run before the first clh_solve (a
   solver on one workgroup has nothing to tune: 0 trials).  Both keep the
   winners for later admissions / solves and store them in the cache that
   schedule.use_tuned reads.  max_seconds bounds the tuning: candidates not
   yet tried when it is spent are skipped. */
int clh_server_autotune(clh_server* srv, double max_seconds, clh_tune_report* report);
int clh_solver_autotune(clh_solver* s, double max_seconds, clh_tune_report* report);
/* the schedule in use: a server's for a problem with n + m rows and
   columns, a solver's for its problem (grid_blocks: the launch's count) */
int clh_server_get_schedule(clh_server* srv, int32_t n_plus_m, clh_schedule* out);
int clh_solver_get_schedule(clh_solver* s, clh_schedule* out);

/* ---- LDL' plugin: the direct solver of one symmetric quasi-definite matrix
   on the device, at the seam of the reference's LDL solver protocol
   (src/kktsolvers/direct-ldl/directldl_defaults.jl:1-38; the GPU solvers
   src/kktsolvers/gpu/directldl_cudss.jl:5-58 and directldl_mixed_cudss.jl:5-78,
   called by GPULDLKKTSolver at src/kktsolvers/kktsolver_directldl_gpu.jl:90,
   :247, :308, :381).  The matrix is a full symmetric CSR (the reference's GPU
   shape :full, directldl_cudss.jl:31) whose values live in device memory and
   change between refactorizations; `dsigns` are the expected pivot signs. */
typedef struct clh_ldl clh_ldl;

/* ldlsolver_constructor(::Val{:cudss}) / CUDSSDirectLDLSolver(KKT, x, b)
   (directldl_cudss.jl:13-28): the symbolic analysis, once.  precision is
   CLH_METHOD_HIP (FP64 factors) or CLH_METHOD_HIPMIXED (FP32 factors, as
   :cudssmixed); index_base 0 or 1 (the reference passes 1-based Int32). */
int clh_ldl_create(clh_ldl** out, int32_t n, const int32_t* rowptr, const int32_t* colval, int32_t index_base,
                   const int32_t* dsigns, int32_t precision);
/* refactor!(s) (directldl_cudss.jl:33-45): numeric LDL' of the current values
   (device pointer, CSR order).  static_eps is added to the diagonal with the
   sign of dsigns (0 when the caller regularized the values already); with
   dynamic_enable, pivots of the wrong sign or below dyn_eps become
   dyn_delta * sign (QDLDL's rule).  Returns 0, or 1 for a zero or non-finite
   pivot (the reference's Bool failure). */
int clh_ldl_refactor(clh_ldl* h, const double* vals_dev, double static_eps, int32_t dynamic_enable, double dyn_eps,
                     double dyn_delta, void* stream);
/* solve!(s, x, b) (directldl_cudss.jl:50-57): x = K^{-1} b with the factors
   (device pointers, length n). */
int clh_ldl_solve(clh_ldl* h, double* x_dev, const double* b_dev, void* stream);
/* _iterative_refinement (kktsolver_directldl_gpu.jl:349-428): the solve
   refined against the unregularized values vals_dev, until
   |b - K x|_inf <= abstol + reltol |b|_inf, at most maxiter steps, stopping
   when a step improves the error by less than stopratio.  Returns 0, or 1 when
   a residual is not finite. */
int clh_ldl_solve_refine(clh_ldl* h, const double* vals_dev, double* x_dev, const double* b_dev, double reltol,
                         double abstol, int32_t maxiter, double stopratio, void* stream);
typedef struct {
  int64_t nnz_L;                /* strictly lower entries of the factor            */
  double factor_flops;          /* sum_k c_k (c_k + 1)                             */
  int32_t n_supernodes, n_levels;
  int32_t n_positive, n_negative;  /* inertia of the last factorization (pivot signs) */
  int32_t last_dynreg;          /* pivots regularized by the last refactor          */
  int32_t last_ir_steps;        /* refinement steps of the last solve_refine        */
  double last_refine_error;     /* its final |b - K x|_inf                          */
  int32_t grid_blocks;          /* workgroups of the cooperative grid the operations
                                   run on (a factorization of at least 1e8 flops:
                                   the whole GPU), 0: one workgroup               */
  int32_t pad_ldl;
} clh_ldl_info;
int clh_ldl_get_info(clh_ldl* h, clh_ldl_info* info);
void clh_ldl_destroy(clh_ldl* h);

/* ---- Cone operations: the batched cone kernels of the reference's GPU
   composite cone (src/gpucones/coneops_compositecone_gpu.jl; family kernels
   in coneops_{nncone,socone,expcone,powcone,psdtrianglecone}_gpu.jl) for a
   caller that drives its own interior-point loop.  A handle holds one
   composite cone of m rows (the cone list with the cone codes above, PSD dim
   = matrix side) and its scaling state in device memory.  Vectors are device
   pointers of length m; scalar results are written to device memory, as the
   reference keeps them in CuVectors.  Each call is asynchronous on `stream`.
   The operations run over a grid of workgroups: one thread per nonnegative
   row and per cone (second-order cones serially in their thread, PSD cones
   with their dense work in the handle's workspace). */
typedef struct clh_cones clh_cones;
typedef struct {
  int32_t m, degree;                 /* rows; the barrier degree nu            */
  int32_t n_zero, n_nonneg, n_soc, n_exp, n_pow, n_psd;  /* cones per family */
  int32_t symmetric;                 /* no exponential or power cone           */
  int32_t pad;
  int64_t hs_len;                    /* length of the get_Hs output            */
} clh_cones_info;
/* CompositeConeGPU(cones) (src/gpucones/compositecone_type_gpu.jl:9-216);
   settings supplies max_step_fraction, linesearch_backtrack_step and
   min_terminate_step_length (NULL: the defaults) */
int clh_cones_create(clh_cones** out, int32_t ncones, const int32_t* cone_type, const int32_t* cone_dim,
                     const double* cone_param, const clh_settings* settings);
int clh_cones_get_info(clh_cones* h, clh_cones_info* info);
/* update_scaling! (coneops_compositecone_gpu.jl:185-249): scaling_strategy
   0 primal-dual, 1 dual (the asymmetric cones' ScalingStrategy); *ok_dev = 1
   when every cone's scaling succeeded, else 0 */
int clh_cones_update_scaling(clh_cones* h, const double* s, const double* z, double mu, int32_t scaling_strategy,
                             double* ok_dev, void* stream);
/* set_identity_scaling! (:161-183) */
int clh_cones_set_identity_scaling(clh_cones* h, void* stream);
/* get_Hs! (:252-309): the Hs blocks in cone order, d entries (the diagonal)
   for zero, nonnegative and sparse-expanded second-order cones (dim > 5), a
   d x d column-major block otherwise; hs_len doubles */
int clh_cones_get_Hs(clh_cones* h, double* Hsblocks, void* stream);
/* mul_Hs! (:315-406): y = Hs x */
int clh_cones_mul_Hs(clh_cones* h, double* y, const double* x, void* stream);
/* affine_ds! (:409-456): ds = lambda o lambda (symmetric), s (asymmetric) */
int clh_cones_affine_ds(clh_cones* h, double* ds, const double* s, void* stream);
/* combined_ds_shift! (:458-513): shift = W step_z o W^-T step_s - sigma_mu e
   (Mehrotra) or the third-order correction (asymmetric); step_z and step_s
   are updated in place as in the reference */
int clh_cones_combined_ds_shift(clh_cones* h, double* shift, double* step_z, double* step_s, double sigma_mu,
                                void* stream);
/* Δs_from_Δz_offset! (:515-563): out = W^T (lambda \ ds) (symmetric), ds (asymmetric) */
int clh_cones_ds_from_dz_offset(clh_cones* h, double* out, const double* ds, const double* z, void* stream);
/* step_length (:566-648): the largest alpha <= alpha_max keeping z + alpha dz
   and s + alpha ds in the cone; the asymmetric cones backtrack from
   min(alpha, max_step_fraction) */
int clh_cones_step_length(clh_cones* h, const double* dz, const double* ds, const double* z, const double* s,
                          double alpha_max, double* alpha_dev, void* stream);
/* No counterpart in the reference: clh_cones_step_length for callers that
   cannot allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid allocation).  The same kernels and result;
   the per-workgroup reduction partials and the reduced scalars live in a
   static device array of the library (2 x 2048 + 8 doubles per GPU, its
   address looked up when the first cone handle is created on that device),
   so the call makes no allocation of any kind.  A composite needing more
   than 2048 workgroups (about 500 000 rows and cones) is rejected (< 0).
   Every call on one GPU shares that array: calls that overlap on a GPU, on
   any streams, race. */
int clh_cones_step_length_noalloc(clh_cones* h, const double* dz, const double* ds, const double* z,
                                  const double* s, double alpha_max, double* alpha_dev, void* stream);
/* compute_barrier (:650-704): barrier of (z + alpha dz, s + alpha ds) */
int clh_cones_compute_barrier(clh_cones* h, const double* z, const double* s, const double* dz, const double* ds,
                              double alpha, double* barrier_dev, void* stream);
/* margins (:42-79): out_dev[0] the least
   margin, out_dev[1] the sum of the positive margins */
int clh_cones_margins(clh_cones* h, const double* z, double* out_dev, void* stream);
/* scaled_unit_shift! (:81-113): z += alpha e; primal_cone != 0 also zeroes
   the zero cone's rows (PrimalCone) */
int clh_cones_scaled_unit_shift(clh_cones* h, double* z, double alpha, int32_t primal_cone, void* stream);
/* unit_initialization! (:115-159) */
int clh_cones_unit_initialization(clh_cones* h, double* z, double* s, void* stream);
/* rectify_equilibration! (:21-39): delta = 1, and mean(e)/e on each
   non-elementwise cone; *any_changed (host, may be NULL) as the reference's
   return value */
int clh_cones_rectify_equilibration(clh_cones* h, double* delta, const double* e, int32_t* any_changed,
                                    void* stream);
void clh_cones_destroy(clh_cones* h);

/* device memory for callers without a GPU runtime of their own (tests, the
   LDL' plugin's vectors); no counterpart in the reference (CuArrays there) */
int clh_device_alloc(void** ptr, size_t bytes);
int clh_device_free(void* ptr);
int clh_memcpy_to_device(void* dst_dev, const void* src, size_t bytes);
int clh_memcpy_to_host(void* dst, const void* src_dev, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
