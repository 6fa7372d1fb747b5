/*
 * CPU oracle for the interior-point hot path: a from-scratch C++ restatement
 * of Clarabel.jl v0.10.0's CPU path (direct_solve_method = :qdldl) with
 * presolve and chordal decomposition disabled, which is what the GPU path
 * runs (src/problemdata.jl:87-100).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the cuclarabel_amd package links,
 * loads or calls this library; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker and as the CPU
 * baseline.  It is never the thing measured for the GPU figures.
 *
 * Layout of the C interface: compressed-sparse-column (CSC) matrices with
 * 0-based int32 indices, cones as parallel arrays (type, dim, parameter).
 */
#ifndef CLARABEL_ORACLE_H
#define CLARABEL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cone type codes (src/cones/cone_api.jl:24-56) */
enum { ORC_ZERO = 0, ORC_NONNEG = 1, ORC_SOC = 2, ORC_EXP = 3, ORC_POW = 4, ORC_PSD = 5 };

/* status codes (src/statuscodes.jl:25-37), same numbering */
enum {
  ORC_UNSOLVED = 0, ORC_SOLVED, ORC_PRIMAL_INFEASIBLE, ORC_DUAL_INFEASIBLE,
  ORC_ALMOST_SOLVED, ORC_ALMOST_PRIMAL_INFEASIBLE, ORC_ALMOST_DUAL_INFEASIBLE,
  ORC_MAX_ITERATIONS, ORC_MAX_TIME, ORC_NUMERICAL_ERROR, ORC_INSUFFICIENT_PROGRESS
};

/* Settings, field-for-field the defaults of src/settings.jl:70-148 */
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
  double static_regularization_constant, static_regularization_proportional;
  int32_t dynamic_regularization_enable;
  double dynamic_regularization_eps, dynamic_regularization_delta;
  int32_t iterative_refinement_enable;
  double iterative_refinement_reltol, iterative_refinement_abstol;
  int32_t iterative_refinement_max_iter;
  double iterative_refinement_stop_ratio;
  /* 1 = GPU data-update semantics of update_P! (no cost scaling c,
     src/data_updating.jl:87-104), 0 = CPU semantics (:71-85) */
  int32_t update_P_without_cscale;
  /* 0 = the reference's dynamic regularization (QDLDL: a pivot with the
     wrong sign becomes sign * delta); 1 = the rule the GPU engine takes by
     default, sign * max(delta, |d|): no counterpart in the reference, a
     switch so that tests can compare the two rules on one implementation */
  int32_t dynamic_regularization_keep_magnitude;
} orc_settings;

typedef struct {
  int32_t status;
  int32_t iterations;
  double obj_val, obj_val_dual;
  double r_prim, r_dual;
  double solve_time, setup_time;
  /* per-iteration trace (iterations+1 rows), columns as the verbose
     printout (src/info_print.jl:122-137): pcost dcost gap pres dres k/t mu step */
  int32_t trace_len;
  /* diagnostics of the last KKT factorization */
  int64_t nnz_L;
  double flops_factor;
} orc_result;

void orc_default_settings(orc_settings* s);

/* Solve min 1/2 x'Px + q'x  s.t. Ax + s = b, s in K.
   P: n x n CSC (upper triangle is used; lower entries ignored)
   A: m x n CSC
   x[n], s[m], z[m] out; trace[8*(max_iter+2)] out (may be NULL). */
int orc_solve(int32_t n, int32_t m,
              const int32_t* Pp, const int32_t* Pi, const double* Px,
              const double* q,
              const int32_t* Ap, const int32_t* Ai, const double* Ax,
              const double* b,
              int32_t ncones, const int32_t* cone_type, const int32_t* cone_dim,
              const double* cone_param,
              const orc_settings* settings,
              double* x, double* s, double* z,
              orc_result* res, double* trace);

/* Solve, then replace P values, b, q and/or A values (same patterns) keeping
   the first solve's equilibration, and solve again (src/data_updating.jl).
   Pass NULL for an update that is not made.  Outputs refer to the second
   solve. */
int orc_solve_update(int32_t n, int32_t m,
                     const int32_t* Pp, const int32_t* Pi, const double* Px,
                     const double* q,
                     const int32_t* Ap, const int32_t* Ai, const double* Ax,
                     const double* b,
                     int32_t ncones, const int32_t* cone_type, const int32_t* cone_dim,
                     const double* cone_param,
                     const orc_settings* settings,
                     const double* Px_new, const double* b_new,
                     const double* q_new, const double* Ax_new,
                     double* x, double* s, double* z,
                     orc_result* res, double* trace);

/* ---- component entry points used by the unit tests ---- */

/* Wright omega function w(z): w + log(w) = z (src/cones/coneops_expcone.jl:463-519) */
double orc_wright_omega(double z);

/* Exp cone primal gradient g(s) (coneops_expcone.jl:284-297) */
void orc_exp_gradient_primal(const double* s, double* g);
/* Exp cone dual barrier gradient and Hessian (coneops_expcone.jl:389-418) */
void orc_exp_dual_grad_H(const double* z, double* grad, double* H);
/* Power cone primal gradient (coneops_powcone.jl:288-316) */
void orc_pow_gradient_primal(const double* s, double alpha, double* g);
void orc_pow_dual_grad_H(const double* z, double alpha, double* grad, double* H);

/* SOC NT scaling for one cone of dimension dim: outputs w[dim], lambda[dim], eta
   and, for the sparse representation, d, u[dim], v[dim]
   (src/cones/coneops_socone.jl:75-154).  Returns 1 on success. */
int orc_soc_update_scaling(int32_t dim, const double* s, const double* z,
                           double* w, double* lambda, double* eta,
                           double* d, double* u, double* v);
/* step length for one SOC pair (coneops_socone.jl:448-517) */
double orc_soc_step_length_component(int32_t dim, const double* x, const double* y, double amax);

/* PSD NT scaling of one cone (side n) from svec s, z: R, Rinv (n x n
   column-major), lambda[n], Hs (t x t column-major, t = n(n+1)/2).
   (src/cones/coneops_psdtrianglecone.jl:79-142).  Returns 1 on success. */
int orc_psd_update_scaling(int32_t n, const double* s, const double* z,
                           double* R, double* Rinv, double* lambda, double* Hs);
/* Exponential (type 3) or power (type 4, parameter alpha) cone at the dual
   point z: the third-order correction eta(ds, v) of combined_ds_shift!
   (coneops_expcone.jl:338-385, coneops_powcone.jl:329-404), and the scaling
   matrix Hs (3x3 column-major) of update_Hs (strategy 0 primal-dual, 1 dual;
   coneops_nonsymmetric_common.jl:50-164) */
void orc_nonsym_higher_correction(int32_t type, double alpha, const double* z, const double* ds, const double* v,
                                  double* eta);
void orc_nonsym_update_Hs(int32_t type, double alpha, const double* s, const double* z, double mu, int32_t strategy,
                          double* Hs);
/* symmetric eigenvalues (ascending) of an n x n column-major matrix */
void orc_sym_eigvals(int32_t n, const double* M, double* evals);

/* Sparse LDL' of a symmetric quasi-definite matrix given as upper-triangle
   CSC, with the QDLDL sign-based dynamic regularization (keep_magnitude:
   the rule of orc_settings.dynamic_regularization_keep_magnitude); solves
   K x = b.  Returns 0 on success.  Used to pin the factorization by backward
   error. */
int orc_ldl_solve(int32_t n, const int32_t* Kp, const int32_t* Ki, const double* Kx,
                  const int32_t* dsigns, double reg_eps, double reg_delta, int32_t keep_magnitude,
                  const double* b, double* x, int64_t* nnz_L);

#ifdef __cplusplus
}
#endif
#endif
