/*
 * cpu_lbl.c -- TEST INFRASTRUCTURE ONLY (the CPU baseline leg of bench.py).
 *
 * The reference's CPU numeric path restated with BLAS-3, multi-threaded: a
 * CPU restatement, not the reference binary (METIS and the reference's
 * OpenBLAS build are not in the image).  Same algorithm and data layout as
 * oracle_factor in lbl_oracle.c -- supernodal left-looking LDL^T,
 * ldl_left_sn_parallel_02 (src/ldl/parallel_blocked_ldlt_02.cpp:17-366) --
 * with the dense steps on the same LAPACK / BLAS calls the reference makes:
 *   descendant updates     DGEMM (parallel_blocked_ldlt_02.cpp:130-136) after
 *                          the D scaling of blocked_2by2_mult (:128)
 *   diagonal block         DSYTRF (:155), reorder_after_sytrf (:156)
 *   rows below it          the column permutation of DLAPMT (:164-165), then
 *                          DTRSM with the unit lower block (:180-186) and the
 *                          D^-1 of blocked_2by2_solver (:188)
 *   solves                 blocked L / D / L^T sweeps with DTRSM and DGEMM
 *                          (src/triangularSolve/Triangular_BCSC.cpp:378-446,
 *                          534-600) over all right-hand sides at once
 * Parallelism follows the reference's scheme: independent supernodes of one
 * tree level run in parallel with single-threaded BLAS (its level-set
 * partitions, :37-206); levels with fewer supernodes than threads run one
 * supernode at a time with threaded BLAS (its final sequential level,
 * :209-343).
 *
 * BLAS / LAPACK come from a shared library given at run time (MKL's
 * libmkl_rt with its GNU OpenMP threading layer, or any library exporting
 * the Fortran symbols dgemm_, dtrsm_, dsytrf_).
 */
#include <dlfcn.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* shared with lbl_oracle.c */
int oracle_reorder_after_sytrf(int n, double* a, int lda, const int* ipiv, int* perm, double* D, int64_t lda_D);
void oracle_permute_pinv(int n, const int* Ap, const int* Ai, const double* Ax, const int* pinv, int** Bp_o,
                         int** Bi_o, double** Bx_o);

typedef void (*dgemm_fn)(const char*, const char*, const int*, const int*, const int*, const double*,
                         const double*, const int*, const double*, const int*, const double*, double*, const int*);
typedef void (*dtrsm_fn)(const char*, const char*, const char*, const char*, const int*, const int*, const double*,
                         const double*, const int*, double*, const int*);
typedef void (*dsytrf_fn)(const char*, const int*, double*, const int*, int*, double*, const int*, int*);

static dgemm_fn p_dgemm;
static dtrsm_fn p_dtrsm;
static dsytrf_fn p_dsytrf;
static int (*p_set_local)(int);   /* MKL_Set_Num_Threads_Local: BLAS threads of the calling thread */
static char g_blas_name[256] = "";

/* Load BLAS / LAPACK from `path`; 0 on success. */
int cpu_lbl_init(const char* path) {
  void* h = dlopen(path, RTLD_NOW | RTLD_GLOBAL);
  if (!h) return -1;
  int (*set_layer)(int) = (int (*)(int))dlsym(h, "MKL_Set_Threading_Layer");
  if (set_layer) set_layer(3); /* MKL_THREADING_GNU: the same OpenMP runtime as this file */
  p_dgemm = (dgemm_fn)dlsym(h, "dgemm_");
  p_dtrsm = (dtrsm_fn)dlsym(h, "dtrsm_");
  p_dsytrf = (dsytrf_fn)dlsym(h, "dsytrf_");
  p_set_local = (int (*)(int))dlsym(h, "MKL_Set_Num_Threads_Local");
  if (!p_dgemm || !p_dtrsm || !p_dsytrf) return -2;
  void (*ver)(char*, int) = (void (*)(char*, int))dlsym(h, "MKL_Get_Version_String");
  if (ver) ver(g_blas_name, (int)sizeof(g_blas_name) - 1);
  else snprintf(g_blas_name, sizeof(g_blas_name), "%s", path);
  return 0;
}

const char* cpu_lbl_blas_name(void) { return g_blas_name; }

static void blas_threads(int t) {
  if (p_set_local) p_set_local(t);
}

/* per-thread scratch, grown on demand */
typedef struct {
  double *T, *C, *work;
  int64_t nT, nC, nwork;
  int *map, *ipiv, *lperm;
  int nmap, nipiv;
} Scratch;

static double* grow(double** p, int64_t* cap, int64_t need) {
  if (need > *cap) {
    free(*p);
    *p = (double*)malloc(sizeof(double) * (size_t)need);
    *cap = need;
  }
  return *p;
}

/* one supernode: assemble, pull the descendants' updates, DSYTRF, DTRSM */
static void factor_one(int t, int n, const int* super, const int64_t* pi, const int* s_in, const int64_t* p,
                       const int* Bp, const int* Bi, const double* Bx, const int* tcount, const int* dlist,
                       double* valL, double* d_val, int* perm_piv, Scratch* S, int* npiv) {
  const int c0 = super[t], c1 = super[t + 1], w = c1 - c0;
  const int r = (int)(pi[t + 1] - pi[t]);
  double* cur = valL + p[c0];
  int* map = S->map;
  for (int64_t q = pi[t], k = 0; q < pi[t + 1]; ++q, ++k) map[s_in[q]] = (int)k;
  for (int c = c0; c < c1; ++c)
    for (int q = Bp[c]; q < Bp[c + 1]; ++q) cur[(int64_t)(c - c0) * r + map[Bi[q]]] = Bx[q];
  const double one = 1.0, zero = 0.0, mone = -1.0;
  for (int e = tcount[t]; e < tcount[t + 1]; ++e) {
    const int d = dlist[e];
    const int dc0 = super[d], dw = super[d + 1] - dc0;
    const int dr = (int)(pi[d + 1] - pi[d]);
    const int* drow = s_in + pi[d];
    int lb = -1, ub = -1;
    for (int q = 0; q < dr; ++q) {
      if (drow[q] >= c0 && lb < 0) lb = q;
      if (lb >= 0 && drow[q] < c1) ub = q;
    }
    const int m = dr - lb, k1 = ub - lb + 1;
    const double* src = valL + p[dc0] + lb;
    const double* Dd = d_val + dc0;
    double* T = grow(&S->T, &S->nT, (int64_t)m * dw);
    double* C = grow(&S->C, &S->nC, (int64_t)m * k1);
    for (int l = 0; l < dw;) { /* T = src D_d (blocked_2by2_mult) */
      if (Dd[l + n] == 0.0) {
        for (int i = 0; i < m; ++i) T[(int64_t)l * m + i] = Dd[l] * src[(int64_t)l * dr + i];
        l += 1;
      } else {
        const double d1 = Dd[l], d2 = Dd[l + 1], od = Dd[l + n];
        for (int i = 0; i < m; ++i) {
          const double x0 = src[(int64_t)l * dr + i], x1 = src[(int64_t)(l + 1) * dr + i];
          T[(int64_t)l * m + i] = d1 * x0 + od * x1;
          T[(int64_t)(l + 1) * m + i] = od * x0 + d2 * x1;
        }
        l += 2;
      }
    }
    /* C (m x k1) = T (m x dw) src(0:k1, :)^T */
    p_dgemm("N", "T", &m, &k1, &dw, &one, T, &m, src, &dr, &zero, C, &m);
    for (int j = 0; j < k1; ++j) {
      const int col = map[drow[lb + j]];
      for (int i = j; i < m; ++i) cur[(int64_t)col * r + map[drow[lb + i]]] -= C[(int64_t)j * m + i];
    }
  }
  /* Bunch-Kaufman LDL^T of the diagonal block (LAPACK dsytrf, 'L') */
  int info = 0, lwork = -1;
  double wq = 0.0;
  p_dsytrf("L", &w, cur, &r, S->ipiv, &wq, &lwork, &info);
  lwork = (int)wq > w ? (int)wq : w;
  double* work = grow(&S->work, &S->nwork, lwork);
  p_dsytrf("L", &w, cur, &r, S->ipiv, work, &lwork, &info);
  const int is_perm = oracle_reorder_after_sytrf(w, cur, r, S->ipiv, S->lperm, d_val + c0, n);
  const int rows = r - w;
  if (is_perm && rows > 0) { /* dlapmt: new column k <- old column lperm[k] of the rows below */
    double* tmp = grow(&S->C, &S->nC, (int64_t)rows * w);
    for (int k = 0; k < w; ++k)
      memcpy(tmp + (int64_t)k * rows, cur + (int64_t)S->lperm[k] * r + w, sizeof(double) * rows);
    for (int k = 0; k < w; ++k) memcpy(cur + (int64_t)k * r + w, tmp + (int64_t)k * rows, sizeof(double) * rows);
  }
  if (is_perm) __atomic_fetch_add(npiv, 1, __ATOMIC_RELAXED);
  for (int k = 0; k < w; ++k) perm_piv[c0 + k] = c0 + S->lperm[k];
  for (int l = 0; l < w; ++l) {
    d_val[c0 + l] = cur[l + (int64_t)l * r];
    cur[l + (int64_t)l * r] = 1.0;
  }
  if (rows > 0) {
    /* B := B L^-T (unit lower diagonal block), then B := B D^-1 */
    p_dtrsm("R", "L", "T", "U", &rows, &w, &one, cur, &r, cur + w, &r);
    for (int l = 0; l < w; ++l) {
      double* b0 = cur + (int64_t)l * r + w;
      if (d_val[c0 + l + n] == 0.0) {
        const double inv = 1.0 / d_val[c0 + l];
        for (int i = 0; i < rows; ++i) b0[i] *= inv;
      } else {
        double* b1 = cur + (int64_t)(l + 1) * r + w;
        const double d1 = d_val[c0 + l], d2 = d_val[c0 + l + 1], od = d_val[c0 + l + n];
        const double inv = 1.0 / (d1 * d2 - od * od);
        for (int i = 0; i < rows; ++i) {
          const double x1 = b0[i], x2 = b1[i];
          b0[i] = (x1 * d2 - x2 * od) * inv;
          b1[i] = (x2 * d1 - x1 * od) * inv;
        }
        ++l;
      }
    }
  }
  (void)mone;
}

/*
 * Inputs and outputs as oracle_factor (lbl_oracle.c), plus sparent (the
 * assembly tree, for the level schedule) and the thread count.  Returns the
 * number of supernodes that pivoted.
 */
int cpu_lbl_factor(int n, const int* Ap, const int* Ai, const double* Ax, double reg_diag, const int* perm,
                   int nsuper, const int* super, const int64_t* pi, const int* s_in, const int64_t* p,
                   const int* sparent, double* valL, double* d_val, int* s_lab, int* perm_final, int nthreads) {
  double* Kx = (double*)malloc(sizeof(double) * (Ap[n] ? Ap[n] : 1));
  memcpy(Kx, Ax, sizeof(double) * Ap[n]);
  for (int j = 0; j < n; ++j)
    if (reg_diag > 0 && fabs(Kx[Ap[j]]) < reg_diag) Kx[Ap[j]] = reg_diag; /* find_perturbation */
  int* pinv = (int*)malloc(sizeof(int) * n);
  for (int k = 0; k < n; ++k) pinv[perm[k]] = k;
  int *Bp, *Bi;
  double* Bx;
  oracle_permute_pinv(n, Ap, Ai, Kx, pinv, &Bp, &Bi, &Bx);
  free(Kx);
  int* col2sup = (int*)malloc(sizeof(int) * n);
  for (int t = 0; t < nsuper; ++t)
    for (int c = super[t]; c < super[t + 1]; ++c) col2sup[c] = t;
  int* tcount = (int*)calloc(nsuper + 1, sizeof(int));
  for (int d = 0; d < nsuper; ++d) {
    int w = super[d + 1] - super[d], last = -1;
    for (int64_t q = pi[d] + w; q < pi[d + 1]; ++q) {
      int t = col2sup[s_in[q]];
      if (t != last) {
        tcount[t + 1]++;
        last = t;
      }
    }
  }
  for (int t = 0; t < nsuper; ++t) tcount[t + 1] += tcount[t];
  int* dlist = (int*)malloc(sizeof(int) * (tcount[nsuper] ? tcount[nsuper] : 1));
  int* dfill = (int*)malloc(sizeof(int) * (nsuper + 1));
  memcpy(dfill, tcount, sizeof(int) * (nsuper + 1));
  for (int d = 0; d < nsuper; ++d) {
    int w = super[d + 1] - super[d], last = -1;
    for (int64_t q = pi[d] + w; q < pi[d + 1]; ++q) {
      int t = col2sup[s_in[q]];
      if (t != last) {
        dlist[dfill[t]++] = d;
        last = t;
      }
    }
  }
  free(dfill);
  /* levels by height in the assembly tree */
  int* lev = (int*)calloc(nsuper, sizeof(int));
  int nlev = 0;
  for (int t = 0; t < nsuper; ++t) {
    if (lev[t] + 1 > nlev) nlev = lev[t] + 1;
    if (sparent[t] >= 0 && lev[sparent[t]] < lev[t] + 1) lev[sparent[t]] = lev[t] + 1;
  }
  int* lptr = (int*)calloc(nlev + 1, sizeof(int));
  for (int t = 0; t < nsuper; ++t) lptr[lev[t] + 1]++;
  for (int l = 0; l < nlev; ++l) lptr[l + 1] += lptr[l];
  int* lset = (int*)malloc(sizeof(int) * nsuper);
  int* lfill = (int*)malloc(sizeof(int) * (nlev + 1));
  memcpy(lfill, lptr, sizeof(int) * (nlev + 1));
  for (int t = 0; t < nsuper; ++t) lset[lfill[lev[t]]++] = t;
  free(lfill);
  int maxw = 1;
  for (int t = 0; t < nsuper; ++t)
    if (super[t + 1] - super[t] > maxw) maxw = super[t + 1] - super[t];

  memset(valL, 0, sizeof(double) * (size_t)p[n]);
  memset(d_val, 0, sizeof(double) * 2 * (size_t)n);
  int* perm_piv = (int*)malloc(sizeof(int) * n);
  int npiv = 0;
  if (nthreads < 1) nthreads = 1;
  Scratch* S = (Scratch*)calloc(nthreads, sizeof(Scratch));
  for (int k = 0; k < nthreads; ++k) {
    S[k].map = (int*)malloc(sizeof(int) * n);
    S[k].ipiv = (int*)malloc(sizeof(int) * maxw);
    S[k].lperm = (int*)malloc(sizeof(int) * maxw);
  }
  for (int l = 0; l < nlev; ++l) {
    const int cnt = lptr[l + 1] - lptr[l];
    if (cnt >= nthreads && nthreads > 1) {
#pragma omp parallel num_threads(nthreads)
      {
        blas_threads(1);
        Scratch* me = S + omp_get_thread_num();
#pragma omp for schedule(dynamic, 1)
        for (int q = lptr[l]; q < lptr[l + 1]; ++q)
          factor_one(lset[q], n, super, pi, s_in, p, Bp, Bi, Bx, tcount, dlist, valL, d_val, perm_piv, me, &npiv);
      }
    } else {
      blas_threads(nthreads);
      for (int q = lptr[l]; q < lptr[l + 1]; ++q)
        factor_one(lset[q], n, super, pi, s_in, p, Bp, Bi, Bx, tcount, dlist, valL, d_val, perm_piv, S, &npiv);
    }
  }
  int* newpos = (int*)malloc(sizeof(int) * n);
  for (int k = 0; k < n; ++k) newpos[perm_piv[k]] = k;
  for (int64_t q = 0; q < pi[nsuper]; ++q) s_lab[q] = newpos[s_in[q]];
  for (int t = 0; t < nsuper; ++t)
    for (int k = 0; k < super[t + 1] - super[t]; ++k) s_lab[pi[t] + k] = super[t] + k;
  for (int k = 0; k < n; ++k) perm_final[k] = perm[perm_piv[k]];
  for (int k = 0; k < nthreads; ++k) {
    free(S[k].T);
    free(S[k].C);
    free(S[k].work);
    free(S[k].map);
    free(S[k].ipiv);
    free(S[k].lperm);
  }
  free(S);
  free(newpos);
  free(perm_piv);
  free(lset);
  free(lptr);
  free(lev);
  free(dlist);
  free(tcount);
  free(col2sup);
  free(pinv);
  free(Bp);
  free(Bi);
  free(Bx);
  return npiv;
}

/* forward step of supernode t: X_t := L_tt^-1 X_t, X_rows -= L_rt X_t
 * (atomic scatter when supernodes of one level run concurrently) */
static void fwd_one(int t, int n, int nrhs, const int* super, const int64_t* pi, const int* s_lab, const int64_t* p,
                    const double* valL, double* X, double* Y, int atomic) {
  const double one = 1.0, zero = 0.0;
  const int c0 = super[t], w = super[t + 1] - c0, r = (int)(pi[t + 1] - pi[t]), rows = r - w;
  const double* L = valL + p[c0];
  p_dtrsm("L", "L", "N", "U", &w, &nrhs, &one, L, &r, X + c0, &n);
  if (rows <= 0) return;
  p_dgemm("N", "N", &rows, &nrhs, &w, &one, L + w, &r, X + c0, &n, &zero, Y, &rows);
  for (int c = 0; c < nrhs; ++c)
    for (int i = 0; i < rows; ++i) {
      double* dst = X + (int64_t)c * n + s_lab[pi[t] + w + i];
      const double v = Y[(int64_t)c * rows + i];
      if (atomic) {
#pragma omp atomic
        *dst -= v;
      } else {
        *dst -= v;
      }
    }
}

/* backward step of supernode t: X_t -= L_rt^T X_rows, X_t := L_tt^-T X_t
 * (reads ancestors' final values only; no conflicts within a level) */
static void bwd_one(int t, int n, int nrhs, const int* super, const int64_t* pi, const int* s_lab, const int64_t* p,
                    const double* valL, double* X, double* Y) {
  const double one = 1.0, mone = -1.0;
  const int c0 = super[t], w = super[t + 1] - c0, r = (int)(pi[t + 1] - pi[t]), rows = r - w;
  const double* L = valL + p[c0];
  if (rows > 0) {
    for (int c = 0; c < nrhs; ++c)
      for (int i = 0; i < rows; ++i) Y[(int64_t)c * rows + i] = X[(int64_t)c * n + s_lab[pi[t] + w + i]];
    p_dgemm("T", "N", &w, &nrhs, &rows, &mone, L + w, &r, Y, &rows, &one, X + c0, &n);
  }
  p_dtrsm("L", "L", "T", "U", &w, &nrhs, &one, L, &r, X + c0, &n);
}

/*
 * X (n x nrhs, column-major, leading dimension n, pivoted ordering) :=
 * L^-T D^-1 L^-1 X with the factor of cpu_lbl_factor: per supernode DTRSM on
 * the diagonal block and DGEMM for the rows below, all columns at once.
 * Supernodes of one tree level (height above the leaves; the parent of a
 * supernode is the supernode of its first row below the diagonal block) run
 * in parallel with single-threaded BLAS, as in the factor; levels with fewer
 * supernodes than threads run one at a time with threaded BLAS.
 */
void cpu_lbl_solve(int n, int nrhs, int nsuper, const int* super, const int64_t* pi, const int* s_lab,
                   const int64_t* p, const double* valL, const double* d_val, double* X, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  int maxr = 1;
  for (int t = 0; t < nsuper; ++t)
    if (pi[t + 1] - pi[t] > maxr) maxr = (int)(pi[t + 1] - pi[t]);
  int* col2sup = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int t = 0; t < nsuper; ++t)
    for (int c = super[t]; c < super[t + 1]; ++c) col2sup[c] = t;
  int* lev = (int*)calloc((size_t)(nsuper > 0 ? nsuper : 1), sizeof(int));
  int nlev = 0;
  for (int t = 0; t < nsuper; ++t) {
    if (lev[t] + 1 > nlev) nlev = lev[t] + 1;
    const int w = super[t + 1] - super[t];
    if (pi[t + 1] - pi[t] > w) {
      const int par = col2sup[s_lab[pi[t] + w]];
      if (lev[par] < lev[t] + 1) lev[par] = lev[t] + 1;
    }
  }
  int* lptr = (int*)calloc(nlev + 1, sizeof(int));
  for (int t = 0; t < nsuper; ++t) lptr[lev[t] + 1]++;
  for (int l = 0; l < nlev; ++l) lptr[l + 1] += lptr[l];
  int* lset = (int*)malloc(sizeof(int) * (size_t)(nsuper > 0 ? nsuper : 1));
  int* lfill = (int*)malloc(sizeof(int) * (nlev + 1));
  memcpy(lfill, lptr, sizeof(int) * (nlev + 1));
  for (int t = 0; t < nsuper; ++t) lset[lfill[lev[t]]++] = t;
  free(lfill);
  double* Yall = (double*)malloc(sizeof(double) * (size_t)maxr * nrhs * nthreads);
  for (int l = 0; l < nlev; ++l) {
    const int cnt = lptr[l + 1] - lptr[l];
    if (cnt >= nthreads && nthreads > 1) {
#pragma omp parallel num_threads(nthreads)
      {
        blas_threads(1);
        double* Y = Yall + (size_t)maxr * nrhs * omp_get_thread_num();
#pragma omp for schedule(dynamic, 16)
        for (int q = lptr[l]; q < lptr[l + 1]; ++q) fwd_one(lset[q], n, nrhs, super, pi, s_lab, p, valL, X, Y, 1);
      }
    } else {
      blas_threads(nthreads);
      for (int q = lptr[l]; q < lptr[l + 1]; ++q) fwd_one(lset[q], n, nrhs, super, pi, s_lab, p, valL, X, Yall, 0);
    }
  }
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int c = 0; c < nrhs; ++c) {
    double* x = X + (int64_t)c * n;
    for (int i = 0; i < n; ++i) {
      if (d_val[i + n] == 0.0) {
        x[i] /= d_val[i];
      } else {
        const double d1 = d_val[i], d2 = d_val[i + 1], od = d_val[i + n];
        const double x1 = x[i], x2 = x[i + 1];
        const double inv = 1.0 / (d1 * d2 - od * od);
        x[i] = (x1 * d2 - x2 * od) * inv;
        x[i + 1] = (x2 * d1 - x1 * od) * inv;
        ++i;
      }
    }
  }
  for (int l = nlev - 1; l >= 0; --l) {
    const int cnt = lptr[l + 1] - lptr[l];
    if (cnt >= nthreads && nthreads > 1) {
#pragma omp parallel num_threads(nthreads)
      {
        blas_threads(1);
        double* Y = Yall + (size_t)maxr * nrhs * omp_get_thread_num();
#pragma omp for schedule(dynamic, 16)
        for (int q = lptr[l]; q < lptr[l + 1]; ++q) bwd_one(lset[q], n, nrhs, super, pi, s_lab, p, valL, X, Y);
      }
    } else {
      blas_threads(nthreads);
      for (int q = lptr[l]; q < lptr[l + 1]; ++q) bwd_one(lset[q], n, nrhs, super, pi, s_lab, p, valL, X, Yall);
    }
  }
  free(Yall);
  free(lset);
  free(lptr);
  free(lev);
  free(col2sup);
}
