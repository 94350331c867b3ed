// Engine handle: host copy of the KKT matrix, settings, symbolic analysis and
// (once factored) the device-resident factor.
#pragma once
#include <memory>
#include <vector>

#include "host/symbolic.hpp"
#include "lblg.h"

namespace lblg {

struct DeviceFactor;  // engine/device.hpp

struct Handle {
  int n = 0;
  std::vector<int> Ap, Ai;
  std::vector<double> Ax;          // values as given (unperturbed)
  lblg_settings set{};
  Symbolic sym;
  bool analyzed = false;
  bool factored = false;
  std::vector<int> perturbed;      // columns whose diagonal was replaced by reg_diag
  std::vector<unsigned char> isdiag;
  std::vector<int64_t> csr_src;
  std::vector<int> csr_rp;
  double k0_norm_inf = 0.0;
  double max_abs_entry = 0.0;      // max |K_ij| of the values last uploaded
  int num_ref_iter = 0;
  int zero_pivot = 0;
  int64_t second_col_steps = 0;  // wide panels: pivot steps that inspected a second column
  int64_t interchanges = 0;      // wide panels: row/column interchanges
  int64_t wide_panels = 0;       // wide panels factored
  int64_t wide_panels_pivoted = 0;  // wide panels whose 1x1 no-interchange pass stopped early
  double factor_seconds = 0.0;
  // entries changed by lblg_update_values since the values were last
  // uploaded (positions in Ax), and whether all of Ax must be uploaded
  std::vector<int64_t> changed;
  std::vector<unsigned char> changed_flag;
  bool values_dirty = true;
  int64_t refactored = 0;  // supernodes refactored by the last factorization
  std::vector<int64_t> batch_off;  // block-diagonal batch: first column of each system (count + 1)
  // device schedule knobs (-1: the environment's or built-in default at setup)
  struct Tuning {
    int group_w = -1;    // widest supernode with a single-workgroup panel
    int chain_w = -1;    // narrowest supernode swept by one cooperative launch
    int small_nnz = -1;  // largest factor (entries of L) swept by one workgroup; 0 = never
    int big_m = -1;      // off-diagonal rows from which Schur updates take 128-row tiles
    int wide_sub = -1;   // cooperative panels per trailing update (rank wide_sub * 64)
  } tune;
  std::unique_ptr<DeviceFactor> dev;
  Handle();
  ~Handle();
};

}  // namespace lblg
