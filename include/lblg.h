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
  int precision;          /* factor storage and MFMA type: LBLG_F64 (plain, the factor
                             of K~); LBLG_F32 (FP32 factor and sweeps of S K S, S
                             the diagonal scaling s_i = 1/sqrt(max_j |K_ij|) of
                             every supernode's rows and columns, diagonals below
                             2^-12 raised to it; residual, GMRES and the solution
                             update stay FP64 on K); LBLG_BF16 (as LBLG_F32, with
                             the large Schur and trailing updates on BF16 MFMA
                             operands and FP32 accumulation) */
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

/* Change some values of K in place: Ax[idx[k]] = vals[k], idx being positions
 * in the lower-CSC arrays given to lblg_create (add_del_matrix_qp,
 * linear_solver_wrapper.cpp:1020-1137, edits A_ord->x the same way).  The
 * next lblg_factor is then the SoMod update (update_factorization,
 * linear_solver_wrapper.cpp:1151-1290, with update_ldl_left_sn_02_v2,
 * src/ldl/Serial_update_ldl.cpp:17-277): it uploads only the changed values
 * and refactors only the supernodes on the elimination-tree paths from the
 * changed entries to the root, as long as the handle holds a factor of the
 * previous values with the same precision, reg_diag and regularization
 * (otherwise it factors everything).  The FP32 / BF16 factors then keep the
 * diagonal scaling of the last full factorization.  Indices out of range:
 * LBLG_ERR_ARG, nothing changed. */
int lblg_update_values(lblg_handle* h, int64_t count, const int64_t* idx, const double* vals);
/* Supernodes the last lblg_factor refactored (all of them after a full
 * factorization; engine extension, the reference's modified_sns count). */
int64_t lblg_refactored_supernodes(const lblg_handle* h);

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

/* SolverSettings::compute_norms + backward_error (linear_solver_wrapper.cpp:1462-1498,
 * norm_dense / norm_sparse_int of src/common/Norm.cpp) for caller-given
 * solutions X of B, recomputed on the device with the unperturbed K0 of the
 * last factorization (set_diags(0)).  B and X are column-major host arrays in
 * the caller's ordering, 1 <= nrhs <= 8.  norms[4 c + 0..3] receive res_l1,
 * x_l1, rhs_l1 and bwd_err = res_l1 / (A_l1 x_l1 + rhs_l1) of column c; A_l1
 * is lblg_factor_info's k0_norm_inf.  Not for batch handles (see
 * lblg_batch_backward_errors). */
int lblg_residual_norms(lblg_handle* h, int nrhs, const double* B, int64_t ldb, const double* X, int64_t ldx,
                        double* norms);

/* The same backward errors with B, X and d_err (nrhs doubles) in device
 * memory, enqueued on the handle's stream without host synchronisation
 * (lblg_synchronize or the stream orders the result).  Partial sums go to a
 * buffer the handle allocates on the first call. */
int lblg_backward_error_device(lblg_handle* h, int nrhs, const double* dB, int64_t ldb, const double* dX,
                               int64_t ldx, double* d_err);

/* lblg_backward_error_device that allocates nothing (no hipMalloc, hipFree or
 * allocating library call), for callers that cannot allocate: code under
 * stream capture, serving loops, harnesses that forbid allocation.  No
 * counterpart in the reference.  Its partial sums go to a fixed 192 KiB
 * device array of the library (one per GPU, kept until the process ends),
 * so it needs no per-handle buffer; limit: 1 <= nrhs <= 8 (LBLG_ERR_ARG
 * beyond).  Every call on a GPU shares that array: calls that overlap on one
 * GPU (other streams, other threads) race and must be serialised by the
 * caller. */
int lblg_backward_error_device_noalloc(lblg_handle* h, int nrhs, const double* dB, int64_t ldb,
                                       const double* dX, int64_t ldx, double* d_err);

/* Device schedule of a handle (engine extension, no reference counterpart:
 * the reference's schedule is fixed by its cost model).  Defaults come from
 * LBLG_GROUP_W, LBLG_CHAIN_W, LBLG_SMALL_SOLVE_NNZ and LBLG_SCHUR_BIG_M. */
typedef struct lblg_tuning {
  int group_w;          /* widest supernode with a single-workgroup panel (wider: cooperative panel) */
  int chain_w;          /* narrowest supernode swept by one cooperative launch */
  int small_solve_nnz;  /* largest factor (entries of L) swept in one workgroup; 0 = never */
  int schur_big_m;      /* off-diagonal rows from which Schur updates use 128-row MFMA tiles */
  int wide_sub;         /* cooperative panels per blocked trailing update (rank wide_sub * 64), 1..16 */
  double factor_ms;     /* lblg_autotune: factor time of the chosen schedule */
  double solve_ms;      /* lblg_autotune: solve time of the chosen schedule (settings' refinement) */
  int schedules_tried;  /* lblg_autotune: schedules timed */
} lblg_tuning;
int lblg_get_tuning(lblg_handle* h, lblg_tuning* t);
/* takes effect at the next lblg_factor (the device schedule is rebuilt) */
int lblg_set_tuning(lblg_handle* h, const lblg_tuning* t);
/* Autotuner: times candidate schedules on the handle's current values (panel
 * group width, Schur tile threshold and the rank of the wide panels'
 * trailing updates by the factor; chain width and the
 * one-workgroup sweep limit by a solve of nrhs fixed right-hand sides with
 * the handle's settings), keeps the fastest and leaves the handle factored
 * with it.  Reuses the symbolic analysis.  reps: timed repetitions per
 * schedule (best of). */
int lblg_autotune(lblg_handle* h, int nrhs, int reps, lblg_tuning* t);

/* Launch sequences replayed from captured HIP graphs (engine extension, no
 * reference counterpart): the factor, the solve with its fixed GMRES cycles,
 * and each GMRES cycle of target-driven refinement are captured on their
 * second call with the same configuration and replayed from then on, except
 * where cooperative wide panels or chain sweeps take part (their launches
 * carry per-launch epochs).  LBLG_GRAPH=0 turns
 * capture off.  Counts of captures and of graph launches so far. */
int lblg_graph_stats(const lblg_handle* h, int64_t* captures, int64_t* launches);

/* Wait for the handle's stream. */
int lblg_synchronize(lblg_handle* h);

/* The handle's HIP stream (hipStream_t), for callers that order their own work. */
void* lblg_stream(lblg_handle* h);

/* Copy the factor to host memory (any pointer may be NULL): valL (xsize),
 * d_val (2n: D diagonal, then 2x2 sub-diagonals), rows (ssize row labels in
 * the pivoted ordering), perm_final (n: Perm composed with the pivot
 * permutation, as L->Perm after reorder_matrix). */
int lblg_get_factor(lblg_handle* h, double* valL, double* d_val, int* rows, int* perm_final);

/* The diagonal scaling s (n, original ordering) of an FP32 or BF16 factor,
 * which factors S K S with S = diag(s); all ones for FP64.
 * Engine extension (the reference factors K~ itself). */
int lblg_get_scaling(lblg_handle* h, double* s);

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
/* Cooperative-panel counts of the last factorization (engine diagnostics; no
 * reference counterpart): panels factored, and panels in which some column
 * failed the 1x1 no-interchange test so that the step-by-step pivot search
 * ran from that column on. */
int lblg_panel_stats(const lblg_handle* h, int64_t* panels, int64_t* panels_pivoted);

/* Per-kernel timing with HIP events recorded on the handle's stream (engine
 * diagnostics; no reference counterpart).  lblg_kernel_time returns the
 * accumulated milliseconds and launch count of a kernel family since timing
 * was enabled: bk_small, bk_panel, bk_panel_group, trailing, offdiag, schur
 * (64-row tiles), scale_ld, schur_big (128-row tiles), fwd_tri, fwd_gemv,
 * bwd_gemv, bwd_tri, spmv.  lblg_set_kernel_timing_only times the families
 * named in a comma-separated list alone (two events per launch of them, none
 * for the others), so a measured run keeps its launch rate;
 * lblg_set_kernel_timing(h, 0) ends either mode.
 * Captured launch sequences (HIP graphs) are not used while timing is on. */
int lblg_set_kernel_timing(lblg_handle* h, int on);
int lblg_set_kernel_timing_only(lblg_handle* h, const char* name);
int lblg_kernel_time(lblg_handle* h, const char* name, double* total_ms, int64_t* launches);

/* Device self-test of the 16x16x4 MFMA lane layout the engine assumes, for
 * the FP64 (precision LBLG_F64) or FP32 (LBLG_F32) instruction:
 * C(16x16) = A(16x4) B(4x16), all row-major (inputs rounded to the type).
 * No reference counterpart (engine diagnostics). */
int lblg_selftest_mfma(const double* A, const double* B, double* C, int precision);

/* ---------------------------------------------------------------------------
 * Batched systems (engine extension; the reference factors one matrix per
 * SolverSettings): `count` independent symmetric matrices of uneven size, each
 * lower CSC with the diagonal first, held by one handle as a block-diagonal
 * matrix.  lblg_batch_analyze orders every block by its own nested dissection
 * (or naturally), so each block's symbolic result is the one it would have
 * alone; lblg_factor / lblg_refactor_resident then factor all blocks in the
 * same level-scheduled launches, and lblg_solve / lblg_solve_device solve
 * all blocks at once (right-hand sides concatenated in block order).  For a
 * batch, the backward error the solve calls return and that target-driven
 * refinement checks is the worst system's.
 * ------------------------------------------------------------------------- */
int lblg_batch_create(lblg_handle** h, int count, const int* n, const int* const* Ap, const int* const* Ai,
                      const double* const* Ax);
int lblg_batch_analyze(lblg_handle* h, const lblg_settings* s);
/* count, and offsets[count + 1]: first row of every block (may be NULL) */
int lblg_batch_info(const lblg_handle* h, int* count, int64_t* offsets);
/* per block and column, the backward error of SolverSettings::backward_error
 * (linear_solver_wrapper.cpp:1462-1498) on that block: err[c * count + k] */
int lblg_batch_backward_errors(const lblg_handle* h, int nrhs, const double* B, int64_t ldb, const double* X,
                               int64_t ldx, double* err);

/* ---------------------------------------------------------------------------
 * NASOQ QP solver over the engine:  min 1/2 x'Hx + q'x  s.t.  Ax = a, Bx <= b
 * (nasoq::Nasoq, include/nasoq/nasoq.h:61-493, src/nasoq.cpp).  The active-set
 * loop runs on the host; every KKT factorization and solve is a device call
 * on the super-KKT [H A' B'; A 0 0; B 0 0] (build_super_matrix,
 * src/QP/linear_solver_wrapper.cpp:246-350).
 * ------------------------------------------------------------------------- */
typedef struct lblg_qp lblg_qp;

/* variants (nasoq_mode, nasoq.h:44-46) and statuses (nasoq_status, :47-49) */
#define LBLG_QP_FIXED 0
#define LBLG_QP_AUTO 1
#define LBLG_QP_TUNED 2
#define LBLG_QP_PREDET 3
#define LBLG_QP_INFEASIBLE 0
#define LBLG_QP_OPTIMAL 1
#define LBLG_QP_INACCURATE 2
#define LBLG_QP_NOT_CONVERGED 3

/* Public settings of Nasoq (Nasoq::default_setting, nasoq.cpp:186-199). */
typedef struct lblg_qp_settings {
  double diag_perturb;  /* default 1e-9 */
  double eps_abs;       /* default 1e-3 */
  double stop_tol;      /* default 1e-15 */
  int max_iter;         /* refinement iterations of the PREDET variant, default 0 */
  int max_iter_nas;     /* active-set iteration cap, default 4000 */
  int variant;          /* LBLG_QP_*, default LBLG_QP_FIXED */
  int ordering;         /* LBLG_ORDER_ND (default) or LBLG_ORDER_NATURAL */
  int precision;        /* factor of the KKT systems: LBLG_F64 (default; the
                           reference's), LBLG_F32 or LBLG_BF16 with FP64
                           refinement (engine extension) */
  int separate_rows;    /* 1 (default, the reference's): every inequality row
                           of the super-KKT is its own supernode (extra_cols,
                           supernode_detection.cpp:221-226); 0 (engine
                           extension): rows join supernodes like any column.
                           Inactive rows are decoupled either way, so both
                           factor the same matrix; with 0 the separators full
                           of coupled rows become dense supernodes instead of
                           chains of one-column supernodes (fewer levels) */
} lblg_qp_settings;

/* Results and KKT checks of the last solve (the reference's public fields). */
typedef struct lblg_qp_info {
  int status;            /* LBLG_QP_* status (check_solve_status, nasoq.cpp:712-728) */
  int num_iter;          /* active-set iterations */
  int n_active;          /* active inequalities */
  int inner_iter_ref;    /* refinement budget used (after the variant's choice) */
  int outer_iter_ref;
  double diag_perturb;   /* perturbation used */
  double stop_tol;
  double primal_obj, dual_obj, objective;
  double lag_res;               /* ||Hx + q + A'y + B'z||_inf */
  double cons_sat_norm;         /* max(||max(Bx - b, 0)||_inf, ||Ax - a||_inf) */
  double non_negativity_infn;   /* ||min(z, 0)||_inf */
  double complementarity_infn;  /* ||z .* (Bx - b)||_inf */
  int64_t factorizations;       /* device KKT factorizations */
  int64_t kkt_solves;           /* device KKT solves */
  /* engine extension (no reference counterpart): host wall time of the KKT
   * factorizations (value upload included) and of the refined KKT solves,
   * and the supernodes those factorizations refactored, all summed since the
   * handle was created (as factorizations and kkt_solves are); the number of
   * supernodes of the super-KKT */
  double factor_seconds;
  double solve_seconds;
  int64_t refactored_supernodes;
  int64_t supernodes;
} lblg_qp_info;

void lblg_qp_default_settings(lblg_qp_settings* s);

/* Nasoq(H_size, Hp, Hi, Hx, q, A_size1, A_size2, Ap, Ai, Ax, a, B_size1, B_size2,
 * Bp, Bi, Bx, b) (nasoq.cpp:36-135): H is n x n lower CSC; A is me x n and B
 * is mi x n, both CSC with n columns (me or mi may be 0, pointers NULL then).
 * All arrays are copied. */
int lblg_qp_create(lblg_qp** qp, int n, const int* Hp, const int* Hi, const double* Hx, const double* q, int me,
                   const int* Ap, const int* Ai, const double* Ax, const double* a, int mi, const int* Bp,
                   const int* Bi, const double* Bx, const double* b);
int lblg_qp_destroy(lblg_qp* qp);
int lblg_qp_set_settings(lblg_qp* qp, const lblg_qp_settings* s);

/* Nasoq::symbolic_QP (nasoq.cpp:201-278) alone: the variant's refinement
 * settings, the super-KKT and its symbolic analysis (host only). */
int lblg_qp_analyze(lblg_qp* qp);

/* Device schedule of the QP's KKT handle chosen by the autotuner
 * (lblg_autotune) on the super-KKT with every inequality inactive; kept for
 * the solves that follow (engine extension, no reference counterpart). */
int lblg_qp_autotune(lblg_qp* qp, int reps, lblg_tuning* t);

/* Nasoq::solve() (nasoq.cpp:656-710); *status receives the LBLG_QP_* status. */
int lblg_qp_solve(lblg_qp* qp, int* status);

/* Nasoq::solve(primal, dual, active_set) (nasoq.cpp:740-760): warm start from
 * x0 (n), z0 (mi), the active inequalities (n_active) and y0 (me, may be NULL).
 * *converged receives 1 when the active-set loop converged. */
int lblg_qp_solve_warm(lblg_qp* qp, const double* x0, const double* z0, const int* active, int n_active,
                       const double* y0, int* converged);

/* Warm start from a predicted active set (engine extension for the learned
 * warm start; no reference counterpart): the predicted constraints are made
 * equalities, those whose multipliers come out negative are dropped until
 * the KKT point is dual feasible, then the dual active-set loop runs from
 * there (as after set_warm_start, nasoq.cpp:413-436) to the same optimum as
 * lblg_qp_solve.  *status: LBLG_QP_*; *prune_solves: KKT solves spent on the
 * predicted set (the loop's own iterations are in lblg_qp_info.num_iter). */
int lblg_qp_solve_predicted(lblg_qp* qp, const int* active, int n_active, int* status, int* prune_solves);

/* New linear terms q (n), a (me), b (mi) for the next solve; a NULL pointer
 * keeps that vector.  The super-KKT pattern and its analysis are kept (engine
 * extension: the reference builds a new Nasoq for new data, nasoq.h:125-134). */
int lblg_qp_set_linear(lblg_qp* qp, const double* q, const double* a, const double* b);

/* primal (n), dual_ineq (mi), dual_eq (me), active set (n_active ints); any may be NULL. */
int lblg_qp_get(const lblg_qp* qp, double* primal, double* dual_ineq, double* dual_eq, int* active);
int lblg_qp_get_info(const lblg_qp* qp, lblg_qp_info* info);

/* The super-KKT of the last solve, lower CSC of order n + me + mi, with the
 * current values (active inequalities carry B, the others zero).  Returns the
 * number of entries; copies when the pointers are not NULL. */
int64_t lblg_qp_super_kkt(const lblg_qp* qp, int* p, int* i, double* x);

/* ---------------------------------------------------------------------------
 * Mode-1 LBL (the super-KKT with inequality rows added and dropped, the
 * SolverSettings expert mode: constructors SolverSettings(H, rhs, B[, BT, C,
 * CT]), linear_solver_wrapper.cpp:100-132; update_somod / downdate_somod
 * :995-1018; add_del_matrix_qp :1020-1137; update_factorization :1151-1290).
 * The handle is an lblg_qp created by lblg_qp_create (q, a, b unused here).
 * ------------------------------------------------------------------------- */

/* Build the super-KKT and analyse it with the LBL settings s (reg_diag,
 * refinement budget, ordering); every inequality row starts inactive. */
int lblg_kkt_analyze(lblg_qp* qp, const lblg_settings* s);
/* Order of the super-KKT (n + me + mi), or negative before lblg_kkt_analyze. */
int64_t lblg_kkt_size(const lblg_qp* qp);
/* The engine handle of the analysed super-KKT (NULL before lblg_kkt_analyze),
 * owned by qp: the symbolic introspection, lblg_get_factor, lblg_inertia and
 * lblg_residual_norms apply to it (the reference's L, valL, d_val of mode 1). */
lblg_handle* lblg_kkt_handle(lblg_qp* qp);
/* Activate (add = 1) or deactivate (add = 0) inequality rows (add_del_matrix_qp). */
int lblg_kkt_edit(lblg_qp* qp, int add, int count, const int* rows);
/* Factor the super-KKT with the current active rows (update_factorization /
 * numerical_factorization): inactive rows are decoupled, which equals the
 * reference's factor with their supernodes hidden. */
int lblg_kkt_factor(lblg_qp* qp);
/* Solve the super-KKT (solve_only, mode 1): rhs and x have n + me + mi entries. */
int lblg_kkt_solve(lblg_qp* qp, const double* rhs, double* x, double* bwd_err);

/* ---------------------------------------------------------------------------
 * Active-set predictor of the learned warm start (engine extension; no
 * reference counterpart): an MLP in -> h1 -> h2 -> out with ReLU hidden
 * layers and sigmoid outputs, BF16 weights and FP32 accumulation on the
 * GPU's BF16 MFMA, trained with binary cross-entropy and Adam.  Parameters
 * are one FP32 vector: W1 (h1 x in, row-major), b1, W2 (h2 x h1), b2,
 * W3 (out x h2), b3.  X rows hold `in` features, Y / P rows `out` labels /
 * probabilities; all buffers are host memory.
 * ------------------------------------------------------------------------- */
typedef struct lblg_mlp lblg_mlp;
int lblg_mlp_create(lblg_mlp** m, int in, int h1, int h2, int out, uint64_t seed, int max_batch);
int lblg_mlp_destroy(lblg_mlp* m);
int64_t lblg_mlp_num_params(const lblg_mlp* m);
int lblg_mlp_get_params(lblg_mlp* m, float* params);
/* also resets the Adam state */
int lblg_mlp_set_params(lblg_mlp* m, const float* params);
/* gradient of the BCE averaged over all batch x out entries (no update), and that loss */
int lblg_mlp_gradients(lblg_mlp* m, const float* X, const float* Y, int batch, float* grad, float* loss);
/* `epochs` passes of Adam over shuffled full minibatches; *last_loss: mean loss of the last pass */
int lblg_mlp_train(lblg_mlp* m, const float* X, const float* Y, int nsamples, int batch, int epochs, float lr,
                   uint64_t seed, float* last_loss);
int lblg_mlp_predict(lblg_mlp* m, const float* X, int nsamples, float* P);
/* one 16x16x16 BF16 MFMA, C = A B (row-major, inputs rounded to BF16): lane layout check */
int lblg_selftest_mfma_bf16(const float* A, const float* B, float* C);

#ifdef __cplusplus
}
#endif
#endif /* LBLG_H */
