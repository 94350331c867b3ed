/*
 * lbl_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the reference's numeric LBL path, used as the
 * checker for the HIP engine (tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py only; the product never links or loads it).
 *
 *   oracle_sytf2         unblocked Bunch-Kaufman LDL^T of a dense block,
 *                        LAPACK dsytf2 ('L') pivot rules, as called through
 *                        LAPACKE_dsytrf at src/ldl/parallel_blocked_ldlt_02.cpp:155
 *   oracle_factor        supernodal left-looking LDL^T:
 *                        ldl_left_sn_02_v2 / ldl_left_sn_parallel_02
 *                        (src/ldl/Serial_blocked_ldl_02_2.cpp:18-183,
 *                         src/ldl/parallel_blocked_ldlt_02.cpp:17-366) with
 *                        reorder_after_sytrf, blocked_2by2_mult,
 *                        blocked_2by2_solver (src/common/Sym_BLAS.cpp:128-196,
 *                        492-529, 573-599), LAPACKE_dlapmt and the unit DTRSM;
 *                        find_perturbation (linear_solver_wrapper.cpp:352-363)
 *                        and reorder_matrix (:1432-1460).  Pivot row moves of
 *                        row_reordering (Sym_BLAS.cpp:327-490) are represented
 *                        by relabelling row indices (same matrix L).
 *   oracle_solve         L, D, L^T sweeps: solve_phase_ldl_blocked_parallel
 *                        (src/linear_solver/solve_phase.cpp:88-128,
 *                         src/triangularSolve/Triangular_BCSC.cpp:378-446, 534-600)
 *   oracle_gmres         pmgmres_ldlt_auto (src/gmres/mgmres.cpp:1015-1211)
 *   oracle_spmv          spmv_csc_sym_one_int (src/matrixVector/spmv_CSC.cpp:26-103)
 *   oracle_backward_error SolverSettings::backward_error
 *                        (linear_solver_wrapper.cpp:1462-1498, Norm.cpp:15-54,115-185)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int oracle_version(void) { return 2; }

/* ------------------------------------------------------------------------ */
/* dense Bunch-Kaufman (LAPACK dsytf2, uplo = 'L'); column-major, 0-based   */
/* ipiv follows LAPACK's 1-based convention: ipiv[k] = kp+1 for a 1x1 pivot,*/
/* ipiv[k] = ipiv[k+1] = -(kp+1) for a 2x2 pivot.                           */
/* ------------------------------------------------------------------------ */
#define A_(i, j) a[(int64_t)(j) * lda + (i)]

static int iamax_col(int len, const double* x) {
  int best = 0;
  double bv = -1.0;
  for (int t = 0; t < len; ++t) {
    double v = fabs(x[t]);
    if (v > bv) {
      bv = v;
      best = t;
    }
  }
  return best;
}

int oracle_sytf2(int n, double* a, int lda, int* ipiv) {
  const double alpha = (1.0 + sqrt(17.0)) / 8.0;
  const double sfmin = 2.2250738585072014e-308;
  int info = 0;
  int k = 0;
  while (k < n) {
    int kstep = 1, kp;
    double absakk = fabs(A_(k, k));
    int imax = k;
    double colmax = 0.0;
    if (k < n - 1) {
      imax = k + 1 + iamax_col(n - k - 1, &A_(k + 1, k));
      colmax = fabs(A_(imax, k));
    }
    if ((absakk > colmax ? absakk : colmax) == 0.0 || isnan(absakk)) {
      if (info == 0) info = k + 1;
      kp = k;
    } else {
      if (absakk >= alpha * colmax) {
        kp = k;
      } else {
        /* largest off-diagonal entry in row imax of the trailing matrix */
        double rowmax = 0.0;
        for (int j = k; j < imax; ++j) {
          double v = fabs(A_(imax, j));
          if (v > rowmax) rowmax = v;
        }
        if (imax < n - 1) {
          int jmax = imax + 1 + iamax_col(n - imax - 1, &A_(imax + 1, imax));
          double v = fabs(A_(jmax, imax));
          if (v > rowmax) rowmax = v;
        }
        if (absakk >= alpha * colmax * (colmax / rowmax)) {
          kp = k;
        } else if (fabs(A_(imax, imax)) >= alpha * rowmax) {
          kp = imax;
        } else {
          kp = imax;
          kstep = 2;
        }
      }
      int kk = k + kstep - 1;
      if (kp != kk) {
        /* symmetric interchange of kk and kp in the trailing matrix */
        for (int i = kp + 1; i < n; ++i) {
          double t = A_(i, kk);
          A_(i, kk) = A_(i, kp);
          A_(i, kp) = t;
        }
        for (int j = kk + 1; j < kp; ++j) {
          double t = A_(j, kk);
          A_(j, kk) = A_(kp, j);
          A_(kp, j) = t;
        }
        double t = A_(kk, kk);
        A_(kk, kk) = A_(kp, kp);
        A_(kp, kp) = t;
        if (kstep == 2) {
          t = A_(k + 1, k);
          A_(k + 1, k) = A_(kp, k);
          A_(kp, k) = t;
        }
      }
      if (kstep == 1) {
        if (k < n - 1) {
          double akk = A_(k, k);
          if (fabs(akk) >= sfmin) {
            double d11 = 1.0 / akk;
            for (int j = k + 1; j < n; ++j) {
              double xj = A_(j, k);
              if (xj != 0.0) {
                double f = -d11 * xj;
                for (int i = j; i < n; ++i) A_(i, j) += A_(i, k) * f;
              }
            }
            for (int i = k + 1; i < n; ++i) A_(i, k) *= d11;
          } else {
            for (int i = k + 1; i < n; ++i) A_(i, k) /= akk;
            for (int j = k + 1; j < n; ++j) {
              double xj = A_(j, k);
              if (xj != 0.0) {
                double f = -akk * xj;
                for (int i = j; i < n; ++i) A_(i, j) += A_(i, k) * f;
              }
            }
          }
        }
      } else {
        if (k < n - 2) {
          double d21 = A_(k + 1, k);
          double d11 = A_(k + 1, k + 1) / d21;
          double d22 = A_(k, k) / d21;
          double t = 1.0 / (d11 * d22 - 1.0);
          d21 = t / d21;
          for (int j = k + 2; j < n; ++j) {
            double wk = d21 * (d11 * A_(j, k) - A_(j, k + 1));
            double wkp1 = d21 * (d22 * A_(j, k + 1) - A_(j, k));
            for (int i = j; i < n; ++i) A_(i, j) -= A_(i, k) * wk + A_(i, k + 1) * wkp1;
            A_(j, k) = wk;
            A_(j, k + 1) = wkp1;
          }
        }
      }
    }
    if (kstep == 1) {
      ipiv[k] = kp + 1;
    } else {
      ipiv[k] = -(kp + 1);
      ipiv[k + 1] = -(kp + 1);
    }
    k += kstep;
  }
  return info;
}

/* ------------------------------------------------------------------------ */
/* reorder_after_sytrf: apply LAPACK's interchanges to the earlier columns of */
/* L, split out the 2x2 sub-diagonals into D[lda_D + i], return the composed */
/* local permutation perm[new] = old.                                         */
/* ------------------------------------------------------------------------ */
int oracle_reorder_after_sytrf(int n, double* a, int lda, const int* ipiv, int* perm, double* D,
                               int64_t lda_D) {
  int* swapv = (int*)malloc(sizeof(int) * (n + 1));
  int* kind = (int*)malloc(sizeof(int) * (n + 1));
  int permuted = 0;
  for (int j = 0; j < n; ++j) {
    swapv[j] = j;
    perm[j] = j;
  }
  for (int i = 0; i < n; ++i) {
    int v = ipiv[i];
    if (v > 0) {
      if (v != i + 1) {
        swapv[i] = v - 1;
        permuted = 1;
      }
      kind[i] = 1;
    } else {
      if (-v != i + 2) {
        swapv[i + 1] = -v - 1;
        permuted = 1;
      }
      kind[i] = 2;
      kind[i + 1] = 0;
      D[i + lda_D] = A_(i + 1, i);
      A_(i + 1, i) = 0.0;
      ++i;
    }
  }
  if (permuted) {
    for (int k = 0; k < n; ++k) {
      int c = swapv[k];
      if (c == k) continue;
      int len = kind[k] == 0 ? k - 1 : k;
      for (int j = 0; j < len; ++j) {
        double t = A_(k, j);
        A_(k, j) = A_(c, j);
        A_(c, j) = t;
      }
      int t = perm[k];
      perm[k] = perm[c];
      perm[c] = t;
    }
  }
  free(swapv);
  free(kind);
  return permuted;
}
#undef A_

/* ------------------------------------------------------------------------ */
/* supernodal factorization                                                  */
/* ------------------------------------------------------------------------ */
/* Lower CSC of P K P^T (rows sorted) with P given as pinv (old -> new). */
void oracle_permute_pinv(int n, const int* Ap, const int* Ai, const double* Ax, const int* pinv,
                          int** Bp_o, int** Bi_o, double** Bx_o) {
  int nnz = Ap[n];
  int* cnt = (int*)calloc(n + 1, sizeof(int));
  for (int j = 0; j < n; ++j)
    for (int q = Ap[j]; q < Ap[j + 1]; ++q) {
      int i = Ai[q];
      if (i < j) continue;
      int a = pinv[i], b = pinv[j];
      cnt[(a < b ? a : b) + 1]++;
    }
  for (int j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
  int* Bp = (int*)malloc(sizeof(int) * (n + 1));
  memcpy(Bp, cnt, sizeof(int) * (n + 1));
  int* Bi = (int*)malloc(sizeof(int) * (nnz ? nnz : 1));
  double* Bx = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
  int* fill = (int*)malloc(sizeof(int) * (n + 1));
  memcpy(fill, Bp, sizeof(int) * (n + 1));
  for (int j = 0; j < n; ++j)
    for (int q = Ap[j]; q < Ap[j + 1]; ++q) {
      int i = Ai[q];
      if (i < j) continue;
      int a = pinv[i], b = pinv[j];
      int c = a < b ? a : b, r = a < b ? b : a;
      int d = fill[c]++;
      Bi[d] = r;
      Bx[d] = Ax[q];
    }
  /* sort rows in each column (insertion sort; columns are short) */
  for (int j = 0; j < n; ++j)
    for (int q = Bp[j] + 1; q < Bp[j + 1]; ++q) {
      int r = Bi[q];
      double v = Bx[q];
      int t = q - 1;
      while (t >= Bp[j] && Bi[t] > r) {
        Bi[t + 1] = Bi[t];
        Bx[t + 1] = Bx[t];
        --t;
      }
      Bi[t + 1] = r;
      Bx[t + 1] = v;
    }
  free(cnt);
  free(fill);
  *Bp_o = Bp;
  *Bi_o = Bi;
  *Bx_o = Bx;
}

/*
 * Inputs: original lower CSC K (unperturbed), reg_diag (0 disables
 * perturbation), symbolic ordering perm (perm[new] = old, before pivoting)
 * and supernodal structure (super, pi, s, p).
 * Outputs: valL (xsize; unit diagonal blocks, 2x2 sub-diagonals zeroed),
 * d_val (2n: diagonal in [0,n), 2x2 sub-diagonals in [n,2n)), s_lab (row
 * indices relabelled to the pivoted ordering), perm_final (Perm o perm_piv),
 * perturbed (n flags).  Returns the number of supernodes that pivoted.
 */
int oracle_factor(int n, const int* Ap, const int* Ai, const double* Ax, double reg_diag,
                  const int* perm, int nsuper, const int* super, const int64_t* pi, const int* s_in,
                  const int64_t* p, double* valL, double* d_val, int* s_lab, int* perm_final,
                  int* perturbed) {
  double* Kx = (double*)malloc(sizeof(double) * (Ap[n] ? Ap[n] : 1));
  memcpy(Kx, Ax, sizeof(double) * Ap[n]);
  for (int j = 0; j < n; ++j) {
    perturbed[j] = 0;
    if (reg_diag > 0 && fabs(Kx[Ap[j]]) < reg_diag) {
      Kx[Ap[j]] = reg_diag;
      perturbed[j] = 1;
    }
  }
  int* pinv = (int*)malloc(sizeof(int) * n);
  for (int k = 0; k < n; ++k) pinv[perm[k]] = k;
  int *Bp, *Bi;
  double* Bx;
  oracle_permute_pinv(n, Ap, Ai, Kx, pinv, &Bp, &Bi, &Bx);
  free(Kx);

  int* col2sup = (int*)malloc(sizeof(int) * n);
  for (int t = 0; t < nsuper; ++t)
    for (int c = super[t]; c < super[t + 1]; ++c) col2sup[c] = t;
  /* update lists: descendants d of each target t (distinct targets of d's rows) */
  int64_t ssz = pi[nsuper];
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

  memset(valL, 0, sizeof(double) * (size_t)p[n]);
  memset(d_val, 0, sizeof(double) * 2 * (size_t)n);
  int* map = (int*)malloc(sizeof(int) * n);
  int* perm_piv = (int*)malloc(sizeof(int) * n);
  int maxr = 1, maxw = 1;
  for (int t = 0; t < nsuper; ++t) {
    int r = (int)(pi[t + 1] - pi[t]), w = super[t + 1] - super[t];
    if (r > maxr) maxr = r;
    if (w > maxw) maxw = w;
  }
  double* T = (double*)malloc(sizeof(double) * (size_t)maxr * maxw);
  double* C = (double*)malloc(sizeof(double) * (size_t)maxr * maxw);
  int* ipiv = (int*)malloc(sizeof(int) * maxw);
  int* lperm = (int*)malloc(sizeof(int) * maxw);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)maxr * maxw);
  int npiv = 0;

  for (int t = 0; t < nsuper; ++t) {
    const int c0 = super[t], c1 = super[t + 1], w = c1 - c0;
    const int r = (int)(pi[t + 1] - pi[t]);
    double* cur = valL + p[c0];
    for (int64_t q = pi[t], k = 0; q < pi[t + 1]; ++q, ++k) map[s_in[q]] = (int)k;
    /* assemble */
    for (int c = c0; c < c1; ++c)
      for (int q = Bp[c]; q < Bp[c + 1]; ++q) cur[(int64_t)(c - c0) * r + map[Bi[q]]] = Bx[q];
    /* gather updates from descendants */
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
      /* T = src * D_d  (1x1 and 2x2 blocks), m x dw */
      for (int l = 0; l < dw;) {
        if (Dd[l + n] == 0.0) {
          for (int i = 0; i < m; ++i) T[(int64_t)l * m + i] = Dd[l] * src[(int64_t)l * dr + i];
          l += 1;
        } else {
          double d1 = Dd[l], d2 = Dd[l + 1], od = Dd[l + n];
          for (int i = 0; i < m; ++i) {
            double x0 = src[(int64_t)l * dr + i], x1 = src[(int64_t)(l + 1) * dr + i];
            T[(int64_t)l * m + i] = d1 * x0 + od * x1;
            T[(int64_t)(l + 1) * m + i] = od * x0 + d2 * x1;
          }
          l += 2;
        }
      }
      /* C = T * src[0:k1, :]^T  (m x k1) */
      for (int j = 0; j < k1; ++j) {
        double* cj = C + (int64_t)j * m;
        for (int i = 0; i < m; ++i) cj[i] = 0.0;
        for (int l = 0; l < dw; ++l) {
          double b = src[(int64_t)l * dr + j];
          if (b == 0.0) continue;
          const double* tl = T + (int64_t)l * m;
          for (int i = j; i < m; ++i) cj[i] += tl[i] * b;
        }
      }
      /* scatter-subtract the lower part into the target panel */
      for (int j = 0; j < k1; ++j) {
        int col = map[drow[lb + j]];
        for (int i = j; i < m; ++i) cur[(int64_t)col * r + map[drow[lb + i]]] -= C[(int64_t)j * m + i];
      }
    }
    /* Bunch-Kaufman on the diagonal block */
    oracle_sytf2(w, cur, r, ipiv);
    int is_perm = oracle_reorder_after_sytrf(w, cur, r, ipiv, lperm, d_val + c0, n);
    const int rows = r - w;
    if (is_perm) {
      /* column permutation of the off-diagonal rows: new col k <- old col lperm[k] */
      for (int k = 0; k < w; ++k)
        memcpy(tmp + (int64_t)k * rows, cur + (int64_t)lperm[k] * r + w, sizeof(double) * rows);
      for (int k = 0; k < w; ++k)
        memcpy(cur + (int64_t)k * r + w, tmp + (int64_t)k * rows, sizeof(double) * rows);
      ++npiv;
    }
    for (int k = 0; k < w; ++k) perm_piv[c0 + k] = c0 + lperm[k];
    for (int l = 0; l < w; ++l) {
      d_val[c0 + l] = cur[l + (int64_t)l * r];
      cur[l + (int64_t)l * r] = 1.0;
    }
    /* B := B * L^{-T}, unit lower L (diag block), B = rows x w */
    for (int j = 0; j < w; ++j) {
      double* bj = cur + (int64_t)j * r + w;
      for (int l = 0; l < j; ++l) {
        double f = cur[j + (int64_t)l * r];
        if (f == 0.0) continue;
        const double* bl = cur + (int64_t)l * r + w;
        for (int i = 0; i < rows; ++i) bj[i] -= f * bl[i];
      }
    }
    /* B := B * D^{-1} (blocked_2by2_solver) */
    for (int l = 0; l < w; ++l) {
      double* b0 = cur + (int64_t)l * r + w;
      if (d_val[c0 + l + n] == 0.0) {
        double inv = 1.0 / d_val[c0 + l];
        for (int i = 0; i < rows; ++i) b0[i] *= inv;
      } else {
        double* b1 = cur + (int64_t)(l + 1) * r + w;
        double d1 = d_val[c0 + l], d2 = d_val[c0 + l + 1], od = d_val[c0 + l + n];
        double inv = 1.0 / (d1 * d2 - od * od);
        for (int i = 0; i < rows; ++i) {
          double x1 = b0[i], x2 = b1[i];
          b0[i] = (x1 * d2 - x2 * od) * inv;
          b1[i] = (x2 * d1 - x1 * od) * inv;
        }
        ++l;
      }
    }
  }
  /* relabel row indices into the pivoted ordering (row_reordering) */
  int* newpos = (int*)malloc(sizeof(int) * n);
  for (int k = 0; k < n; ++k) newpos[perm_piv[k]] = k;
  for (int64_t q = 0; q < ssz; ++q) s_lab[q] = newpos[s_in[q]];
  for (int t = 0; t < nsuper; ++t) /* diagonal-block rows are positional */
    for (int k = 0; k < super[t + 1] - super[t]; ++k) s_lab[pi[t] + k] = super[t] + k;
  for (int k = 0; k < n; ++k) perm_final[k] = perm[perm_piv[k]];

  free(newpos);
  free(tmp);
  free(lperm);
  free(ipiv);
  free(C);
  free(T);
  free(perm_piv);
  free(map);
  free(dlist);
  free(tcount);
  free(col2sup);
  free(pinv);
  free(Bp);
  free(Bi);
  free(Bx);
  return npiv;
}

/* ------------------------------------------------------------------------ */
/* solves in the pivoted ordering: x := L^{-T} D^{-1} L^{-1} x              */
/* ------------------------------------------------------------------------ */
void oracle_solve_ordered(int n, int nsuper, const int* super, const int64_t* pi, const int* s_lab,
                          const int64_t* p, const double* valL, const double* d_val, double* x) {
  for (int t = 0; t < nsuper; ++t) {
    const int c0 = super[t], w = super[t + 1] - c0, r = (int)(pi[t + 1] - pi[t]);
    const double* L = valL + p[c0];
    for (int j = 0; j < w; ++j) {
      double xj = x[c0 + j];
      for (int i = j + 1; i < w; ++i) x[c0 + i] -= L[i + (int64_t)j * r] * xj;
    }
    for (int i = w; i < r; ++i) {
      double acc = 0.0;
      for (int j = 0; j < w; ++j) acc += L[i + (int64_t)j * r] * x[c0 + j];
      x[s_lab[pi[t] + i]] -= acc;
    }
  }
  for (int i = 0; i < n; ++i) {
    if (d_val[i + n] == 0.0) {
      x[i] /= d_val[i];
    } else {
      double d1 = d_val[i], d2 = d_val[i + 1], od = d_val[i + n];
      double x1 = x[i], x2 = x[i + 1];
      double inv = 1.0 / (d1 * d2 - od * od);
      x[i] = (x1 * d2 - x2 * od) * inv;
      x[i + 1] = (x2 * d1 - x1 * od) * inv;
      ++i;
    }
  }
  for (int t = nsuper - 1; t >= 0; --t) {
    const int c0 = super[t], w = super[t + 1] - c0, r = (int)(pi[t + 1] - pi[t]);
    const double* L = valL + p[c0];
    for (int j = 0; j < w; ++j) {
      double acc = 0.0;
      for (int i = w; i < r; ++i) acc += L[i + (int64_t)j * r] * x[s_lab[pi[t] + i]];
      x[c0 + j] -= acc;
    }
    for (int j = w - 1; j >= 0; --j) {
      double acc = 0.0;
      for (int i = j + 1; i < w; ++i) acc += L[i + (int64_t)j * r] * x[c0 + i];
      x[c0 + j] -= acc;
    }
  }
}

/* y = K x for lower CSC K (diagonal once, off-diagonals twice) */
void oracle_spmv(int n, const int* Ap, const int* Ai, const double* Ax, const double* x, double* y) {
  for (int i = 0; i < n; ++i) y[i] = 0.0;
  for (int j = 0; j < n; ++j) {
    double yj = 0.0, xj = x[j];
    for (int q = Ap[j]; q < Ap[j + 1]; ++q) {
      int i = Ai[q];
      if (i == j) {
        y[i] += Ax[q] * xj;
      } else if (i > j) {
        y[i] += Ax[q] * xj;
        yj += Ax[q] * x[i];
      }
    }
    y[j] += yj;
  }
}

/* ||b - K x||_1 / (||K||_inf ||x||_1 + ||b||_1), vector norms as the
 * reference's norm_dense(1, n, v, 0) calls (sum of absolute values) */
double oracle_backward_error(int n, const int* Ap, const int* Ai, const double* Ax, const double* x,
                             const double* b) {
  double* r = (double*)malloc(sizeof(double) * n);
  double* W = (double*)calloc(n, sizeof(double));
  oracle_spmv(n, Ap, Ai, Ax, x, r);
  double res = 0, xn = 0, bn = 0, an = 0;
  for (int i = 0; i < n; ++i) {
    res += fabs(b[i] - r[i]);
    xn += fabs(x[i]);
    bn += fabs(b[i]);
  }
  for (int j = 0; j < n; ++j)
    for (int q = Ap[j]; q < Ap[j + 1]; ++q) {
      int i = Ai[q];
      double v = fabs(Ax[q]);
      if (i == j) W[i] += v;
      else if (i > j) {
        W[i] += v;
        W[j] += v;
      }
    }
  for (int i = 0; i < n; ++i)
    if (W[i] > an) an = W[i];
  free(r);
  free(W);
  double den = an * xn + bn;
  return den > 0 ? res / den : 0.0;
}

/* ------------------------------------------------------------------------ */
/* GMRES(mr) left-preconditioned by the LDL^T solve (pmgmres_ldlt_auto),     */
/* in the pivoted ordering; K_ord is the lower CSC of the unperturbed matrix */
/* (perturbed diagonals set to zero, set_diags(0)).  Returns inner steps.    */
/* ------------------------------------------------------------------------ */
static double dot(int n, const double* a, const double* b) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

int oracle_gmres(int n, const int* Kp, const int* Ki, const double* Kx, int nsuper, const int* super,
                 const int64_t* pi, const int* s_lab, const int64_t* p, const double* valL,
                 const double* d_val, double* x, const double* rhs, int itr_max, int mr, double tol_abs,
                 double tol_rel) {
  const double delta = 1.0e-3;
  double* c = (double*)calloc(mr + 1, sizeof(double));
  double* g = (double*)calloc(mr + 1, sizeof(double));
  double* s = (double*)calloc(mr + 1, sizeof(double));
  double* y = (double*)calloc(mr + 1, sizeof(double));
  double* h = (double*)calloc((size_t)(mr + 1) * (mr > 0 ? mr : 1), sizeof(double));
  double* v = (double*)calloc((size_t)(mr + 1) * n, sizeof(double));
  double* r = (double*)calloc(n, sizeof(double));
#define H(i, j) h[(i) * (mr > 0 ? mr : 1) + (j)]
#define V(k) (v + (size_t)(k) * n)
  int used = 0;
  double rho_tol = 0.0, rho;
  for (int itr = 0; itr < itr_max; ++itr) {
    oracle_spmv(n, Kp, Ki, Kx, x, r);
    for (int i = 0; i < n; ++i) r[i] = rhs[i] - r[i];
    oracle_solve_ordered(n, nsuper, super, pi, s_lab, p, valL, d_val, r);
    rho = sqrt(dot(n, r, r));
    if (rho < tol_abs) break;
    if (itr == 0) rho_tol = rho * tol_rel;
    for (int i = 0; i < n; ++i) V(0)[i] = r[i] / rho;
    g[0] = rho;
    for (int i = 1; i <= mr; ++i) g[i] = 0.0;
    for (int i = 0; i <= mr; ++i)
      for (int j = 0; j < mr; ++j) H(i, j) = 0.0;
    int k, kc = 0;
    for (k = 0; k < mr; ++k) {
      kc = k;
      oracle_spmv(n, Kp, Ki, Kx, V(k), V(k + 1));
      oracle_solve_ordered(n, nsuper, super, pi, s_lab, p, valL, d_val, V(k + 1));
      double av = sqrt(dot(n, V(k + 1), V(k + 1)));
      for (int j = 0; j <= k; ++j) {
        H(j, k) = dot(n, V(k + 1), V(j));
        for (int i = 0; i < n; ++i) V(k + 1)[i] -= H(j, k) * V(j)[i];
      }
      H(k + 1, k) = sqrt(dot(n, V(k + 1), V(k + 1)));
      if (av + delta * H(k + 1, k) == av) {
        for (int j = 0; j <= k; ++j) {
          double ht = dot(n, V(k + 1), V(j));
          H(j, k) += ht;
          for (int i = 0; i < n; ++i) V(k + 1)[i] -= ht * V(j)[i];
        }
        H(k + 1, k) = sqrt(dot(n, V(k + 1), V(k + 1)));
      }
      if (H(k + 1, k) != 0.0)
        for (int i = 0; i < n; ++i) V(k + 1)[i] /= H(k + 1, k);
      if (k > 0) {
        for (int i = 0; i < k + 2; ++i) y[i] = H(i, k);
        for (int j = 0; j < k; ++j) {
          double g1 = c[j] * y[j] - s[j] * y[j + 1], g2 = s[j] * y[j] + c[j] * y[j + 1];
          y[j] = g1;
          y[j + 1] = g2;
        }
        for (int i = 0; i < k + 2; ++i) H(i, k) = y[i];
      }
      double mu = sqrt(H(k, k) * H(k, k) + H(k + 1, k) * H(k + 1, k));
      if (mu == 0.0) goto done;
      c[k] = H(k, k) / mu;
      s[k] = -H(k + 1, k) / mu;
      H(k, k) = c[k] * H(k, k) - s[k] * H(k + 1, k);
      H(k + 1, k) = 0.0;
      {
        double g1 = c[k] * g[k] - s[k] * g[k + 1], g2 = s[k] * g[k] + c[k] * g[k + 1];
        g[k] = g1;
        g[k + 1] = g2;
      }
      rho = fabs(g[k + 1]);
      ++used;
      if (rho <= rho_tol && rho <= tol_abs) break;
    }
    k = kc;
    if (H(k, k) == 0.0) goto done;
    y[k] = g[k] / H(k, k);
    for (int i = k - 1; i >= 0; --i) {
      y[i] = g[i];
      for (int j = i + 1; j < k + 1; ++j) y[i] -= H(i, j) * y[j];
      y[i] /= H(i, i);
    }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < k + 1; ++j) x[i] += V(j)[i] * y[j];
    if (rho <= rho_tol && rho <= tol_abs) break;
  }
done:
#undef H
#undef V
  free(c);
  free(g);
  free(s);
  free(y);
  free(h);
  free(v);
  free(r);
  return used;
}

/* Permute lower CSC K into the ordering perm (perm[new] = old): helper for
 * GMRES on the pivoted ordering (reorder_matrix). */
void oracle_permute_lower(int n, const int* Ap, const int* Ai, const double* Ax, const int* perm,
                          int* Bp, int* Bi, double* Bx) {
  int* pinv = (int*)malloc(sizeof(int) * n);
  for (int k = 0; k < n; ++k) pinv[perm[k]] = k;
  int *tp, *ti;
  double* tx;
  oracle_permute_pinv(n, Ap, Ai, Ax, pinv, &tp, &ti, &tx);
  memcpy(Bp, tp, sizeof(int) * (n + 1));
  memcpy(Bi, ti, sizeof(int) * Ap[n]);
  memcpy(Bx, tx, sizeof(double) * Ap[n]);
  free(tp);
  free(ti);
  free(tx);
  free(pinv);
}

/* inertia from D: 1x1 by sign, 2x2 by determinant and trace */
void oracle_inertia(int n, const double* d_val, int64_t* pos, int64_t* neg, int64_t* zero) {
  int64_t a = 0, b = 0, z = 0;
  for (int i = 0; i < n; ++i) {
    if (d_val[i + n] == 0.0) {
      if (d_val[i] > 0) ++a;
      else if (d_val[i] < 0) ++b;
      else ++z;
    } else {
      double d1 = d_val[i], d2 = d_val[i + 1], od = d_val[i + n];
      double det = d1 * d2 - od * od;
      if (det < 0) {
        ++a;
        ++b;
      } else if (det > 0) {
        if (d1 + d2 > 0) a += 2;
        else b += 2;
      } else {
        ++z;
        if (d1 + d2 > 0) ++a;
        else if (d1 + d2 < 0) ++b;
        else ++z;
      }
      ++i;
    }
  }
  *pos = a;
  *neg = b;
  *zero = z;
}
