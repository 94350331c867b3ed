// Host-side symbolic analysis for the supernodal LBL (LDL^T + Bunch-Kaufman) engine.
//
// Restates, in this project's own code, the symbolic phase of the reference
// solver so that its outputs are bit-identical for a given fill-reducing
// permutation:
//   symbolic_analysis_lin_solve   src/symbolic/symbolic_phase.cpp:167-890
//   etreeC                        src/common/Etree.cpp:32
//   postOrderC (plain + weighted) src/common/PostOrder.cpp:13-145
//   rowcolcounts                  src/symbolic/ColumnCount.cpp:73-383
//   super_node_detection          src/symbolic/supernode_detection.cpp:43-746
//   computeCostperBlock           src/symbolic/performanceModel.cpp:74-116
//   getCoarseLevelSet_6           src/symbolic/InspectionLevel_06.cpp:15-402
// The reference's METIS_NodeND call (symbolic_phase.cpp:448-527) is replaced
// by nested_dissection() (ordering.cpp) or a caller-supplied permutation.
// All offsets into the factor are 64-bit (the reference uses int products
// that overflow once a supernode holds more than 2^31 entries).
#pragma once
#include <cstdint>
#include <vector>

namespace lblg {

// Lower-triangular CSC pattern (stype = -1), diagonal entry first in each
// column, as the reference's CSC (include/nasoq/common/def.h:73-122).
struct CscView {
  int n;
  const int* p;
  const int* i;
};

struct LevelSchedule {
  int n_level = 0;               // number of coarse levels
  std::vector<int> level_ptr;    // n_level+1, indexes par_ptr
  std::vector<int> par_ptr;      // n_par+1, indexes par_set
  std::vector<int> par_set;      // supernode ids, partition by partition
};

struct Symbolic {
  int n = 0;
  std::vector<int> perm, iperm;  // L->Perm / L->IPerm (postordered)
  std::vector<int> parent;       // column elimination tree (postordered)
  std::vector<int> colcount;     // column counts of L (postordered)
  std::vector<int> extra_cols;   // permuted extra-column flags
  int nsuper = 0;
  std::vector<int> super;        // nsuper+1, first column of each supernode
  std::vector<int> col2sup;      // n
  std::vector<int> sparent;      // nsuper, supernodal assembly tree
  std::vector<int64_t> pi;       // nsuper+1, offsets into s
  std::vector<int> s;            // row indices (sorted), supernode by supernode
  std::vector<int64_t> p;        // n+1, offset of column j's values
  std::vector<int64_t> i_ptr;    // n+1, offset of column j's row list (= pi of its supernode)
  int64_t xsize = 0, ssize = 0;
  int max_sup_wid = 0, max_col = 0;
  int64_t maxcsize = 0, maxesize = 0;
  double lnz = 0, fl = 0;        // sum c_j and sum c_j^2 (double, never overflows)
  std::vector<double> node_cost; // per supernode, computeCostperBlock
  LevelSchedule lbc;             // getCoarseLevelSet_6 output
};

struct SymbolicOptions {
  int nrelax[3] = {4, 16, 48};
  double zrelax[3] = {0.8, 0.1, 0.05};
  int cost_param = 1;            // innerParts of getCoarseLevelSet_6 (reference: num_thread)
  int level_param = -3;          // minLevelDist
  int final_seq_node = 50;       // divRate
};

// Full symbolic analysis of lower-CSC pattern A with the fill-reducing
// permutation `order` (perm[new] = old), then weighted postorder, relaxed
// supernodes, BCSC pointers and the coarse level schedule.  `extra_cols`
// (nullable) flags columns forced into single-column supernodes.
bool analyze(const CscView& A, const int* order, const int* extra_cols,
             const SymbolicOptions& opt, Symbolic& out);

// Symmetric permuted transpose helpers (ptranspose, src/common/Transpose.cpp:399):
// upper_of_permuted: pattern of triu(P A P^T) as CSC (rows unsorted order irrelevant)
// lower_sorted: lower-triangular CSC with sorted rows from an upper CSC.
void upper_of_permuted(const CscView& A, const int* pinv, const double* Ax,
                       std::vector<int>& Up, std::vector<int>& Ui, std::vector<double>* Ux);
void transpose_pattern(int n, const int* Ap, const int* Ai, const double* Ax,
                       std::vector<int>& Tp, std::vector<int>& Ti, std::vector<double>* Tx);

// Elimination tree of an upper-triangular CSC pattern (etreeC, stype > 0).
void etree_upper(int n, const int* Up, const int* Ui, std::vector<int>& parent);
// Postorder of a forest; weight==nullptr gives children in ascending order,
// otherwise children in ascending weight (ties by index), as postOrderC.
int postorder(int n, const int* parent, const int* weight, std::vector<int>& post);
// Column counts of L for lower CSC pattern F with etree parent and postorder post.
void column_counts(int n, const int* Fp, const int* Fi, const int* parent, const int* post,
                   std::vector<int>& colcount);
// Coarse level set of the assembly tree (getCoarseLevelSet_6).
void coarse_level_set(int nsuper, const int* sparent, const double* node_cost,
                      int inner_parts, int min_level_dist, int div_rate, LevelSchedule& out);

// Fill-reducing nested-dissection ordering of the symmetric graph of lower CSC
// A (diagonal ignored).  perm[new] = old.  Deterministic for a given seed.
void nested_dissection(const CscView& A, std::vector<int>& perm, uint64_t seed = 12345);

}  // namespace lblg
