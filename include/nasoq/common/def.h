// Drop-in CSC type of the reference (include/nasoq/common/def.h:73-122 there):
// same field names and layout for the fields a caller sets.
#ifndef NASOQ_AMD_DEF_H
#define NASOQ_AMD_DEF_H
#include <cstddef>

namespace nasoq {
#define CHOLMOD_PATTERN 0
#define CHOLMOD_REAL 1
#define EMPTY -1

typedef struct sympiler_CSC {
  size_t nrow;   // the matrix is nrow-by-ncol
  size_t ncol;
  size_t nzmax;  // entries stored
  int* p;        // column pointers [ncol+1]
  int* i;        // row indices [nzmax]
  int* nz;       // unused (packed matrices only)
  double* x;     // values [nzmax]
  double* z;
  int stype;     // -1: symmetric, lower triangle stored
  int itype;
  int xtype;
  int dtype;
  int sorted;
  int packed;
} CSC;

// Drop-in supernodal factor type of the reference (include/nasoq/common/def.h:131-220
// there): the fields a caller reads, with the reference's names and index
// types.  Column j of supernode s starts at x[p[j]] with leading dimension
// pi[s+1] - pi[s]; the row labels of supernode s are s[pi[s] .. pi[s+1]).
typedef struct Sympiler_BCSC {
  size_t n;          // L is n-by-n
  size_t minor;      // n: factorization did not fail
  int* Perm;         // fill-reducing ordering composed with the pivot permutation (after the factor)
  int* ColCount;     // column counts of the unrelaxed factor
  int* IPerm;        // inverse of Perm
  size_t nzmax;
  size_t* p;         // value offset of every column (n + 1)
  int* i;            // unused (supernodal factor)
  double* x;         // the factor values (valL) when present on the host
  size_t nsuper;     // number of supernodes
  size_t ssize;      // length of s
  size_t xsize;      // length of x
  size_t maxcsize;
  size_t maxesize;
  int* super;        // first column of each supernode (nsuper + 1)
  int* col2Sup;      // supernode of each column
  size_t* pi;        // row-index pointer of each supernode (nsuper + 1)
  size_t* i_ptr;     // row-index pointer per column as the reference stores it
  int* px;           // unused
  int* s;            // row labels of every supernode (pivoted order after the factor)
  int* sParent;      // assembly tree
  int* Parent;       // elimination tree
  int ordering;
  int is_ll;         // 0: L D L'
  int is_super;      // 1
} BCSC;
}  // namespace nasoq
#endif
