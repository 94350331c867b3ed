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
}  // namespace nasoq
#endif
