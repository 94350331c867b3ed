// TEST INFRASTRUCTURE ONLY (oracle).  Linked into oracle/_ref/libnasoq_ref.so
// together with the reference's own symbolic sources, compiled where they lie
// under the reference tree (see oracle/Makefile).  Nothing in the product
// links or loads this file.
//
// The reference's driver symbolic_analysis_lin_solve
// (src/symbolic/symbolic_phase.cpp:167-890) cannot be compiled here because
// its header pulls in metis.h.  This driver repeats its glue for the
// "given permutation" case (the #elif GIVEN branch, symbolic_phase.cpp:420-426)
// and calls the reference's own routines for every step:
//   ptranspose (Transpose.cpp), etreeC (Etree.cpp), postOrderC (PostOrder.cpp),
//   rowcolcounts (ColumnCount.cpp), super_node_detection
//   (supernode_detection.cpp), computeCostperBlock (performanceModel.cpp),
//   getCoarseLevelSet_6 (InspectionLevel_06.cpp),
// plus spmv_csc_sym_one_int (spmv_CSC.cpp) and norm_dense / norm_sparse_int
// (Norm.cpp) for the residual definition.
#include <cstring>
#include <vector>

#include "nasoq/common/Etree.h"
#include "nasoq/common/Norm.h"
#include "nasoq/common/PostOrder.h"
#include "nasoq/common/Transpose.h"
#include "nasoq/matrixVector/spmv_CSC.h"
#include "nasoq/symbolic/ColumnCount.h"
#include "nasoq/symbolic/InspectionLevel_06.h"
#include "nasoq/symbolic/performanceModel.h"
#include "nasoq/symbolic/supernode_detection.h"

using namespace nasoq;

namespace {

struct RefSym {
  int n = 0, nsuper = 0, n_level = 0, n_par = 0, max_sup_wid = 0, max_col = 0;
  long long xsize = 0, ssize = 0;
  std::vector<int> perm, iperm, parent, colcount, super, col2sup, sparent, s, level_ptr, par_ptr,
      par_set, extra;
  std::vector<long long> pi, p, i_ptr;
  std::vector<double> node_cost;
};

CSC* wrap(int n, const int* Ap, const int* Ai, double* Ax) {
  CSC* A = new CSC;
  A->nrow = A->ncol = n;
  A->nzmax = Ap[n];
  A->p = const_cast<int*>(Ap);
  A->i = const_cast<int*>(Ai);
  A->x = Ax;
  A->nz = nullptr;
  A->stype = -1;
  A->xtype = CHOLMOD_PATTERN;
  A->packed = 1;
  A->sorted = 1;
  return A;
}

void drop(CSC* M) {
  if (!M) return;
  delete[] M->p;
  delete[] M->i;
  delete[] M->x;
  delete M;
}

}  // namespace

extern "C" {

void* ref_symbolic(int n, const int* Ap, const int* Ai, const int* order, const int* extra_in,
                   int cost_param, int level_param, int final_seq, const int* nrelax_in,
                   const double* zrelax_in) {
  RefSym* R = new RefSym;
  R->n = n;
  int status = 0;
  std::vector<double> dummy(Ap[n], 0.0);
  CSC* A = wrap(n, Ap, Ai, dummy.data());
  int nrelax[3] = {nrelax_in[0], nrelax_in[1], nrelax_in[2]};
  double zrelax[3] = {zrelax_in[0], zrelax_in[1], zrelax_in[2]};
  std::vector<int> extra(n, 0);
  if (extra_in) extra.assign(extra_in, extra_in + n);

  std::vector<int> Lperm(order, order + n), Parent(n), Post(n), ColCount(n), First(n), Level(n);
  // analyze_ordering (symbolic_phase.cpp:105-165): permute_matrices for
  // stype < 0 gives S = A2 = ptranspose(A, Perm) and F = A1 = ptranspose(A2).
  CSC* A2 = ptranspose(A, 0, Lperm.data(), NULL, 0, status);
  CSC* A1 = ptranspose(A2, 0, NULL, NULL, 0, status);
  etreeC(A2, Parent.data(), status);
  postOrderC(Parent.data(), n, NULL, Post.data(), status);
  int fl = 0, aatfl = 0, lnz = 0;
  rowcolcounts(A1, NULL, 0, Parent.data(), Post.data(), NULL, ColCount.data(), First.data(),
               Level.data(), fl, aatfl, lnz, status);
  drop(A1);
  drop(A2);
  // weighted postorder and its combination (symbolic_phase.cpp:593-646)
  if (postOrderC(Parent.data(), n, ColCount.data(), Post.data(), status) == n) {
    std::vector<int> Wi(n), InvPost(n);
    for (int k = 0; k < n; k++) Wi[k] = Lperm[Post[k]];
    for (int k = 0; k < n; k++) Lperm[k] = Wi[k];
    for (int k = 0; k < n; k++) Wi[k] = ColCount[Post[k]];
    for (int k = 0; k < n; k++) ColCount[k] = Wi[k];
    for (int k = 0; k < n; k++) InvPost[Post[k]] = k;
    for (int k = 0; k < n; ++k) Wi[k] = extra[Lperm[k]];
    for (int k = 0; k < n; ++k) extra[k] = Wi[k];
    for (int nc = 0; nc < n; nc++) {
      int oc = Post[nc];
      int op = Parent[oc];
      Wi[nc] = (op == EMPTY) ? EMPTY : InvPost[op];
    }
    for (int k = 0; k < n; k++) Parent[k] = Wi[k];
  }
  R->perm = Lperm;
  R->iperm.assign(n, 0);
  for (int k = 0; k < n; k++) R->iperm[Lperm[k]] = k;
  R->parent = Parent;
  R->colcount = ColCount;
  R->extra = extra;

  // supernodes (symbolic_phase.cpp:658-673)
  BCSC* L = new BCSC;
  std::memset(L, 0, sizeof(BCSC));
  L->n = n;
  L->ColCount = ColCount.data();
  std::vector<int> Sparent(n);
  A2 = ptranspose(A, 0, Lperm.data(), NULL, 0, status);
  A1 = ptranspose(A2, 0, NULL, NULL, 0, status);
  super_node_detection(1, A2, A1, Parent.data(), extra.data(), L, nrelax, zrelax, Sparent.data(),
                       status);
  const int ns = (int)L->nsuper;
  R->nsuper = ns;
  R->xsize = (long long)L->xsize;
  R->ssize = (long long)L->ssize;
  R->super.assign(L->super, L->super + ns + 1);
  R->sparent.assign(Sparent.begin(), Sparent.begin() + ns);
  R->pi.assign(L->pi, L->pi + ns + 1);
  R->s.assign(L->s, L->s + L->pi[ns]);
  // col2Sup and BCSC pointers (symbolic_phase.cpp:680-711)
  std::vector<int> col2Sup(n);
  int maxSupWid = 0;
  for (int i = 0; i < ns; ++i) {
    int k1 = L->super[i], k2 = L->super[i + 1];
    if (maxSupWid < (k2 - k1)) maxSupWid = (k2 - k1);
    for (int j = k1; j < k2; ++j) col2Sup[j] = i;
  }
  int colLength = 0, maxCol = 0;
  for (int j = 0; j < ns; ++j) {
    int curCol = L->super[j];
    int nxtCol = L->super[j + 1];
    colLength = L->pi[j + 1] - L->pi[j];
    if (colLength > maxCol) maxCol = colLength;
    for (int i = curCol + 1; i < nxtCol + 1; ++i) {
      L->i_ptr[i - 1] = L->pi[j];
      L->p[i] = L->p[curCol] + (i - curCol) * colLength;
    }
  }
  L->i_ptr[n] = L->pi[ns];
  L->p[n] = L->p[n - 1] + colLength;
  R->max_sup_wid = maxSupWid;
  R->max_col = maxCol;
  R->col2sup = col2Sup;
  R->p.assign(L->p, L->p + n + 1);
  R->i_ptr.assign(L->i_ptr, L->i_ptr + n + 1);
  // node cost (symbolic_phase.cpp:715-725)
  R->node_cost.assign(ns, 0.0);
  std::vector<int> xi(2 * A1->ncol, 0);
  for (int s = 1; s <= ns; ++s) {
    int curCol = L->super[s - 1];
    int nxtCol = L->super[s];
    R->node_cost[s - 1] = computeCostperBlock(A1->ncol, curCol, nxtCol, Parent.data(), A1->p, A1->i,
                                              col2Sup.data(), L->super, L->s, L->i_ptr, xi.data());
  }
  int levelNo = 0, parNo = 0;
  int *levelPtr = nullptr, *levelSet = nullptr, *parPtr = nullptr, *partition = nullptr;
  getCoarseLevelSet_6(ns, Sparent.data(), L->super, levelNo, levelPtr, levelSet, parNo, parPtr,
                      partition, cost_param, level_param, final_seq, R->node_cost.data());
  R->n_level = levelNo;
  R->n_par = parNo;
  R->level_ptr.assign(levelPtr, levelPtr + levelNo + 1);
  R->par_ptr.assign(parPtr, parPtr + parNo + 1);
  R->par_set.assign(partition, partition + parPtr[parNo]);
  delete[] levelPtr;
  delete[] parPtr;
  delete[] partition;
  drop(A1);
  drop(A2);
  delete[] L->super;
  delete[] L->p;
  delete[] L->pi;
  delete[] L->i_ptr;
  delete[] L->i;
  delete[] L->s;
  delete L;
  delete A;
  return R;
}

// scalar fields: 0 n, 1 nsuper, 2 n_level, 3 n_par, 4 max_sup_wid, 5 max_col, 6 xsize, 7 ssize
long long ref_symbolic_scalar(void* h, int which) {
  RefSym* R = (RefSym*)h;
  switch (which) {
    case 0: return R->n;
    case 1: return R->nsuper;
    case 2: return R->n_level;
    case 3: return R->n_par;
    case 4: return R->max_sup_wid;
    case 5: return R->max_col;
    case 6: return R->xsize;
    case 7: return R->ssize;
  }
  return -1;
}

// int arrays: 0 perm, 1 iperm, 2 parent, 3 colcount, 4 super, 5 col2sup, 6 sparent, 7 s,
// 8 level_ptr, 9 par_ptr, 10 par_set, 11 extra ; returns length, copies when dst != NULL
long long ref_symbolic_ints(void* h, int which, int* dst) {
  RefSym* R = (RefSym*)h;
  const std::vector<int>* v = nullptr;
  switch (which) {
    case 0: v = &R->perm; break;
    case 1: v = &R->iperm; break;
    case 2: v = &R->parent; break;
    case 3: v = &R->colcount; break;
    case 4: v = &R->super; break;
    case 5: v = &R->col2sup; break;
    case 6: v = &R->sparent; break;
    case 7: v = &R->s; break;
    case 8: v = &R->level_ptr; break;
    case 9: v = &R->par_ptr; break;
    case 10: v = &R->par_set; break;
    case 11: v = &R->extra; break;
    default: return -1;
  }
  if (dst) std::memcpy(dst, v->data(), v->size() * sizeof(int));
  return (long long)v->size();
}

// 64-bit arrays: 0 pi, 1 p, 2 i_ptr
long long ref_symbolic_longs(void* h, int which, long long* dst) {
  RefSym* R = (RefSym*)h;
  const std::vector<long long>* v = which == 0 ? &R->pi : which == 1 ? &R->p : &R->i_ptr;
  if (dst) std::memcpy(dst, v->data(), v->size() * sizeof(long long));
  return (long long)v->size();
}

long long ref_symbolic_cost(void* h, double* dst) {
  RefSym* R = (RefSym*)h;
  if (dst) std::memcpy(dst, R->node_cost.data(), R->node_cost.size() * sizeof(double));
  return (long long)R->node_cost.size();
}

void ref_symbolic_free(void* h) { delete (RefSym*)h; }

// y = K x for symmetric lower CSC K (spmv_csc_sym_one_int, alpha=1, beta=0)
void ref_spmv_sym(int n, const int* Ap, const int* Ai, const double* Ax, const double* x, double* y) {
  double alp[2] = {1.0, 0}, bet[2] = {0.0, 0};
  spmv_csc_sym_one_int(n, const_cast<int*>(Ap), const_cast<int*>(Ai), const_cast<double*>(Ax), -1,
                       alp, bet, 1, const_cast<double*>(x), y);
}

// backward error ||b - Kx|| / (||K|| ||x|| + ||b||) with the reference's norm
// calls (SolverSettings::compute_norms / backward_error,
// src/QP/linear_solver_wrapper.cpp:1462-1498)
double ref_backward_error(int n, const int* Ap, const int* Ai, const double* Ax, const double* x,
                          const double* b) {
  std::vector<double> r(n);
  ref_spmv_sym(n, Ap, Ai, Ax, x, r.data());
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  double x_l1 = norm_dense(1, n, const_cast<double*>(x), 0);
  double rhs_l1 = norm_dense(1, n, const_cast<double*>(b), 0);
  double res_l1 = norm_dense(1, n, r.data(), 0);
  double A_l1 = norm_sparse_int(n, const_cast<int*>(Ap), const_cast<int*>(Ai), const_cast<double*>(Ax),
                                -1, 0);
  if (A_l1 * x_l1 + rhs_l1 > 0) return res_l1 / (A_l1 * x_l1 + rhs_l1);
  return 0;
}

}  // extern "C"
