// Readable state of the drop-in SolverSettings (the reference's L, valL,
// d_val, A_ord, AT_ord, perm_piv, level sets, compute_norms, backward_error,
// reorder_matrix), checked for consistency on a small KKT matrix
// [H A'; A 0] (H: 1-D Laplacian + 1e-2 I, A: 12 rows of 3 entries):
//   L D L' (rebuilt from L->super, L->pi, L->s, valL, d_val) equals the
//   perturbed A_ord (P K~ P' with the final L->Perm) entry by entry;
//   AT_ord is A_ord's transpose; perm_piv composes the symbolic Perm into L->Perm;
//   backward_error() recomputes the residual of the current x and rhs.
// Prints "ok" and exits 0 on success, a non-zero code naming the failed check otherwise.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "nasoq/QP/linear_solver_wrapper.h"

int main() {
  const int n = 60, m = 12, N = n + m;
  // lower CSC of K, diagonal first, sorted rows
  std::vector<std::vector<std::pair<int, double>>> cols(N);
  for (int j = 0; j < n; ++j) {
    cols[j].push_back({j, 2.01});
    if (j + 1 < n) cols[j].push_back({j + 1, -1.0});
  }
  uint64_t seed = 12345;
  auto rnd = [&]() {
    seed = seed * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(seed >> 11) / 9007199254740992.0;
  };
  for (int r = 0; r < m; ++r) {
    for (int k = 0; k < 3; ++k) {
      const int c = (r * 5 + k * 17) % n;
      cols[c].push_back({n + r, 2.0 * rnd() - 1.0});
    }
  }
  for (int r = 0; r < m; ++r) cols[n + r].push_back({n + r, 0.0});
  std::vector<int> Ap(N + 1, 0), Ai;
  std::vector<double> Ax;
  for (int j = 0; j < N; ++j) {
    auto& c = cols[j];
    std::sort(c.begin(), c.end());
    for (auto& e : c) {
      Ai.push_back(e.first);
      Ax.push_back(e.second);
    }
    Ap[j + 1] = (int)Ai.size();
  }
  nasoq::CSC K{};
  K.nrow = K.ncol = N;
  K.nzmax = Ax.size();
  K.p = Ap.data();
  K.i = Ai.data();
  K.x = Ax.data();
  K.stype = -1;
  K.packed = 1;
  K.sorted = 1;
  std::vector<double> rhs(N);
  for (double& v : rhs) v = 2.0 * rnd() - 1.0;
  nasoq::SolverSettings lbl(&K, rhs.data());
  lbl.reg_diag = 1e-9;
  if (lbl.symbolic_analysis() != 1) return 2;
  nasoq::BCSC* L = lbl.L;
  if (!L || (int)L->n != N || L->super[L->nsuper] != N || L->p[N] != L->xsize) return 3;
  if (lbl.n_level < 1 || lbl.level_ptr[lbl.n_level] != lbl.n_par || lbl.max_sup_wid < 1) return 4;
  std::vector<int> perm_sym(L->Perm, L->Perm + N);
  if (lbl.numerical_factorization() != 1) return 5;
  L = lbl.L;
  if (!lbl.valL || !lbl.d_val || !lbl.A_ord || !lbl.AT_ord) return 6;
  for (int k = 0; k < N; ++k)
    if (L->Perm[k] != perm_sym[lbl.perm_piv[k]] || L->IPerm[L->Perm[k]] != k) return 7;
  // L D L' from the fields
  std::vector<double> Ld((size_t)N * N, 0.0), Dd((size_t)N * N, 0.0), M((size_t)N * N, 0.0);
  for (size_t t = 0; t < L->nsuper; ++t) {
    const int c0 = L->super[t], c1 = L->super[t + 1], w = c1 - c0;
    const int r = (int)(L->pi[t + 1] - L->pi[t]);
    for (int j = 0; j < w; ++j)
      for (int q = 0; q < r; ++q) {
        const int row = q < w ? c0 + q : L->s[L->pi[t] + q];
        if (row >= c0 + j) Ld[(size_t)row * N + c0 + j] = lbl.valL[L->p[c0 + j] + q];
      }
  }
  for (int i = 0; i < N; ++i) {
    Dd[(size_t)i * N + i] = lbl.d_val[i];
    if (i + 1 < N && lbl.d_val[N + i] != 0.0) Dd[(size_t)(i + 1) * N + i] = Dd[(size_t)i * N + i + 1] = lbl.d_val[N + i];
  }
  std::vector<double> LD((size_t)N * N, 0.0);
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < N; ++k)
      if (Ld[(size_t)i * N + k] != 0.0)
        for (int j = 0; j < N; ++j) LD[(size_t)i * N + j] += Ld[(size_t)i * N + k] * Dd[(size_t)k * N + j];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      double v = 0.0;
      for (int k = 0; k < N; ++k) v += LD[(size_t)i * N + k] * Ld[(size_t)j * N + k];
      M[(size_t)i * N + j] = v;
    }
  double err = 0.0, big = 0.0;
  const nasoq::CSC* Ao = lbl.A_ord;
  std::vector<double> Kd((size_t)N * N, 0.0);
  for (int j = 0; j < N; ++j)
    for (int q = Ao->p[j]; q < Ao->p[j + 1]; ++q) {
      const int i = Ao->i[q];
      if (i < j) return 8;  // lower triangle
      double v = Ao->x[q];
      if (i == j && std::fabs(v) < lbl.reg_diag) v = lbl.reg_diag;  // find_perturbation
      Kd[(size_t)i * N + j] = Kd[(size_t)j * N + i] = v;
    }
  for (size_t e = 0; e < Kd.size(); ++e) {
    err = std::max(err, std::fabs(M[e] - Kd[e]));
    big = std::max(big, std::fabs(Kd[e]));
  }
  if (!(err <= 1e-10 * big)) {
    std::printf("reconstruction %g of %g\n", err, big);
    return 9;
  }
  // AT_ord is the transpose of A_ord
  const nasoq::CSC* At = lbl.AT_ord;
  if (At->p[N] != Ao->p[N]) return 10;
  for (int j = 0; j < N; ++j)
    for (int q = At->p[j]; q < At->p[j + 1]; ++q)
      if (At->i[q] > j || (Kd[(size_t)At->i[q] * N + j] == 0.0 && At->x[q] != 0.0)) return 11;
  // backward error of the solution, then of an edited x (recomputed, not cached)
  double* x = lbl.solve_only();
  if (!x) return 12;
  const double be = lbl.backward_error();
  if (!(be <= 1e-12)) return 13;
  std::vector<double> r(rhs);
  double rn = 0, xn = 0, bn = 0, an = 0;
  std::vector<double> rowabs(N, 0.0);
  for (int j = 0; j < N; ++j)
    for (int q = Ap[j]; q < Ap[j + 1]; ++q) {
      const int i = Ai[q];
      const double a = Ax[q];  // the zero diagonals of K stay zero in K0
      r[i] -= a * x[j];
      rowabs[i] += std::fabs(a);
      if (i != j) {
        r[j] -= a * x[i];
        rowabs[j] += std::fabs(a);
      }
    }
  x[0] += 1.0;  // edit x: the backward error must follow
  for (int i = 0; i < N; ++i) {
    an = std::max(an, rowabs[i]);
  }
  for (int j = 0; j < N; ++j)
    for (int q = Ap[j]; q < Ap[j + 1]; ++q) {
      const int i = Ai[q];
      if (i == 0) r[j] -= Ax[q] * 1.0;
      if (j == 0 && i != 0) r[i] -= Ax[q] * 1.0;
    }
  for (int i = 0; i < N; ++i) {
    rn += std::fabs(r[i]);
    xn += std::fabs(x[i]);
    bn += std::fabs(rhs[i]);
  }
  const double want = rn / (an * xn + bn);
  const double got = lbl.backward_error();
  if (!(std::fabs(got - want) <= 1e-12 * want) || std::fabs(lbl.A_l1 - an) > 1e-12 * an) {
    std::printf("backward error %.17g want %.17g, A_l1 %.17g want %.17g\n", got, want, lbl.A_l1, an);
    return 14;
  }
  delete[] x;
  std::printf("ok\n");
  return 0;
}
