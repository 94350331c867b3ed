/*
 * lblg.h -- C-ABI of the MI355X-native LBL (supernodal LDL^T with
 * Bunch-Kaufman pivoting) engine.  Plain pointers and sizes only; every entry
 * point returns an int status (0 = success, negative = error class) and never
 * exits the process.  One handle owns one factorization and one HIP stream;
 * handles are independent; a handle is not safe for concurrent use.
 *
 * Each entry point cites the reference interface it replaces.  The reference
 * is the C++ struct nasoq::SolverSettings
 * (include/nasoq/QP/linear_solver_wrapper.h:53-306,
 *  src/QP/linear_solver_wrapper.cpp).
 */
#ifndef LBLG_H
#define LBLG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define LBLG_OK 0
#define LBLG_ERR_ARG (-1)      /* invalid argument */
#define LBLG_ERR_STATE (-2)    /* call out of order (e.g. solve before factor) */
#define LBLG_ERR_DEVICE (-3)   /* HIP runtime error or no usable gfx950 device */
#define LBLG_ERR_NOMEM (-4)    /* host or device allocation failed */
#define LBLG_ERR_NUMERIC (-5)  /* zero pivot / non-finite values */

/* factorization precisions */
#define LBLG_F64 0
#define LBLG_F32 1
#define LBLG_BF16 2

/* orderings */
#define LBLG_ORDER_ND 0      /* this engine's nested dissection (reference: METIS_NodeND) */
#define LBLG_ORDER_GIVEN 1   /* caller-supplied permutation (reference: #elif GIVEN) */
#define LBLG_ORDER_NATURAL 2 /* identity (reference: #ifdef NATURAL) */

typedef struct lblg_handle lblg_handle;

/* Solver settings.  Field names and defaults follow
 * SolverSettings::default_setting (src/QP/linear_solver_wrapper.cpp:205-244). */
typedef struct lblg_settings {
  int ldl_variant;        /* 4: level-scheduled supernodal BK (the reference's ldl_left_sn_parallel_02) */
  int req_ref_iter;       /* GMRES outer iterations (itr_max), default 2 */
  int max_inner_iter;     /* GMRES restart length (mr), default 2 */
  double tol_abs;         /* GMRES absolute tolerance, default 1e-15 */
  double tol_rel;         /* GMRES relative tolerance, default 1e-15 */
  double reg_diag;        /* diagonal perturbation for |K_jj| < reg_diag, default 1e-8 */
  int regularization;     /* 1: replace tiny diagonals by reg_diag (find_perturbation) */
  int cost_param;         /* level-schedule partitions (reference: num_thread) */
  int level_param;        /* default -3 */
  int final_seq_node;     /* default 50 */
  int n_relax[3];         /* default {4,16,48} */
  double z_relax[3];      /* default {0.8,0.1,0.05} */
  int ordering;           /* LBLG_ORDER_* */
  int precision;          /* LBLG_F64 / LBLG_F32 / LBLG_BF16 factor storage and MFMA type */
  double target_bwd_err;  /* >0: keep refining until the backward error is <= target
                             (at most max_refine_rounds GMRES cycles); 0: the reference's
                             fixed GMRES budget */
  int max_refine_rounds;  /* cap for target-driven refinement, default 20 */
} lblg_settings;

/* Defaults of SolverSettings::default_setting (linear_solver_wrapper.cpp:205-244). */
void lblg_default_settings(lblg_settings* s);

/* SolverSettings(CSC* A, double* rhs) (linear_solver_wrapper.cpp:81-98).
 * A is n x n symmetric, lower-triangular CSC (stype = -1), sorted rows, the
 * diagonal entry present and first in every column.  Arrays are copied. */
int lblg_create(lblg_handle** h, int n, const int* Ap, const int* Ai, const double* Ax);

/* SolverSettings::~SolverSettings (linear_solver_wrapper.cpp:134-203). */
int lblg_destroy(lblg_handle* h);

/* SolverSettings::symbolic_analysis (linear_solver_wrapper.cpp:396-536) ->
 * symbolic_analysis_lin_solve (src/symbolic/symbolic_phase.cpp:167-890).
 * perm (length n, perm[new] = old) is read when s->ordering == LBLG_ORDER_GIVEN;
 * extra_cols (length n, may be NULL) flags columns kept as single-column
 * supernodes (the reference's extra_cols, SYM_REMOV). */
int lblg_analyze(lblg_handle* h, const lblg_settings* s, const int* perm, const int* extra_cols);

/* Update the numeric settings of an analysed handle (refinement budget and
 * tolerances, reg_diag, precision) -- the reference's public fields set
 * between calls.  Structural fields (ordering, relaxation, level schedule)
 * must match the analysis, else LBLG_ERR_ARG. */
int lblg_set_settings(lblg_handle* h, const lblg_settings* s);

/* Nested-dissection ordering alone (replaces METIS_NodeND,
 * src/symbolic/symbolic_phase.cpp:448-527). perm[new] = old. */
int lblg_order_nd(int n, const int* Ap, const int* Ai, int* perm);

/* Introspection of the symbolic result (the reference's BCSC fields,
 * include/nasoq/common/def.h:131-220, and the level schedule).
 * scalar ids: 0 n, 1 nsuper, 2 n_level, 3 n_par, 4 max_sup_wid, 5 max_col,
 *             6 xsize, 7 ssize.  Returns the value, or negative on error. */
int64_t lblg_symbolic_scalar(const lblg_handle* h, int which);
/* int arrays: 0 Perm, 1 IPerm, 2 Parent, 3 ColCount, 4 super, 5 col2Sup,
 * 6 sParent, 7 s (row indices), 8 level_ptr, 9 par_ptr, 10 par_set, 11 extra_cols.
 * Returns the length; copies into dst when dst != NULL. */
int64_t lblg_symbolic_ints(const lblg_handle* h, int which, int* dst);
/* 64-bit arrays: 0 pi, 1 p, 2 i_ptr.  Returns the length. */
int64_t lblg_symbolic_longs(const lblg_handle* h, int which, int64_t* dst);
/* sum of squared column counts (flop measure) and nnz(L) of the unrelaxed factor. */
int lblg_symbolic_flops(const lblg_handle* h, double* fl, double* lnz);


/* ---------------------------------------------------------------------------
 * Numeric phase (device).  All numeric work runs on the first visible gfx950
 * device; there is no host fallback.
 * ------------------------------------------------------------------------- */

/* Number of visible HIP devices (0 when none). */
int lblg_device_count(int* count);

/* Message of the last failed call on this thread. */
const char* lblg_last_error(void);

/* SolverSettings::numerical_factorization (linear_solver_wrapper.cpp:602-768),
 * variant 4 (ldl_left_sn_parallel_02, src/ldl/parallel_blocked_ldlt_02.cpp:17-366)
 * followed by reorder_matrix (:1432-1460).  Uploads the handle's values and
 * factors P K~ P^T = L D L^T on the device, K~ = K with |K_jj| < reg_diag
 * replaced by reg_diag (find_perturbation, :352-363). */
int lblg_factor(lblg_handle* h);

/* Re-factor from the values already resident in device memory (same
 * pattern, same values); the work unit the benchmark times. */
int lblg_refactor_resident(lblg_handle* h);

/* Replace the values of K (same pattern; the reference edits A->x in place). */
int lblg_set_values(lblg_handle* h, const double* Ax);

/* SolverSettings::solve_only(n_rhs, rhs) (linear_solver_wrapper.cpp:1300-1404):
 * solve_phase_ldl_blocked_parallel_permuted (src/linear_solver/solve_phase.cpp:113-128)
 * then pmgmres_ldlt_auto refinement (src/gmres/mgmres.cpp:1015-1211) on the
 * unperturbed K.  B and X are column-major n x nrhs host arrays (1 <= nrhs <= 8);
 * the reference solves one right-hand side per call, this entry point solves
 * nrhs together.  bwd_err (nrhs doubles, may be NULL) receives the backward
 * error of each column (SolverSettings::backward_error, :1462-1498). */
int lblg_solve(lblg_handle* h, int nrhs, const double* B, int64_t ldb, double* X, int64_t ldx,
               double* bwd_err);

/* As lblg_solve with B and X in device memory (no host transfer). */
int lblg_solve_device(lblg_handle* h, int nrhs, const double* dB, int64_t ldb, double* dX, int64_t ldx,
                      double* bwd_err);

/* GMRES cycles run by the last solve (the reference's num_ref_iter). */
int lblg_num_ref_iter(const lblg_handle* h);

/* Wait for the handle's stream. */
int lblg_synchronize(lblg_handle* h);

/* The handle's HIP stream (hipStream_t), for callers that order their own work. */
void* lblg_stream(lblg_handle* h);

/* Copy the factor to host memory (any pointer may be NULL): valL (xsize),
 * d_val (2n: D diagonal, then 2x2 sub-diagonals), rows (ssize row labels in
 * the pivoted ordering), perm_final (n: Perm composed with the pivot
 * permutation, as L->Perm after reorder_matrix). */
int lblg_get_factor(lblg_handle* h, double* valL, double* d_val, int* rows, int* perm_final);

/* Inertia (positive, negative, zero eigenvalue counts) of K~ from D.
 * The reference does not compute it; the count follows Sylvester's law. */
int lblg_inertia(lblg_handle* h, int64_t* pos, int64_t* neg, int64_t* zero);

/* zero_pivot: 1 if a zero pivot occurred (the reference ignores dsytrf's info);
 * k0_norm_inf: ||K0||_inf of the unperturbed matrix; factor_seconds: wall time
 * of the last lblg_factor. */
int lblg_factor_info(const lblg_handle* h, int* zero_pivot, double* k0_norm_inf, double* factor_seconds);

/* Pivoting statistics of the wide (cooperative) panels of the last factor:
 * Bunch-Kaufman steps that had to inspect a second column
 * (|a_kk| < alpha * colmax) and row/column interchanges.  Engine
 * diagnostics; the reference has no counterpart. */
int lblg_pivot_stats(const lblg_handle* h, int64_t* second_column_steps, int64_t* interchanges);

/* Per-kernel timing with HIP events recorded on the handle's stream (engine
 * diagnostics; no reference counterpart).  lblg_kernel_time returns the
 * accumulated milliseconds and launch count of a kernel family since timing
 * was enabled: bk_small, bk_panel, bk_panel_group, trailing, offdiag, schur,
 * fwd_tri, fwd_gemv, bwd_gemv, bwd_tri, spmv. */
int lblg_set_kernel_timing(lblg_handle* h, int on);
int lblg_kernel_time(lblg_handle* h, const char* name, double* total_ms, int64_t* launches);

/* Device self-test of the FP64 MFMA operand layout: C(16x16) = A(16x4) B(4x16),
 * all row-major.  No reference counterpart (engine diagnostics). */
int lblg_selftest_mfma(const double* A, const double* B, double* C);

#ifdef __cplusplus
}
#endif
#endif /* LBLG_H */
