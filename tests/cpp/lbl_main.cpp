// The reference's examples/lbl_main.cpp flow against the drop-in header:
// H = 2I, q = (-4,-4) -> x = (-2,-2); q = (8,8) -> x = (4,4).
#include <cmath>
#include <cstdio>

#include "nasoq/QP/linear_solver_wrapper.h"

int main() {
  int sizeH = 2, nnzH = 2;
  double q[2] = {-4, -4};
  int Hp[3] = {0, 1, 2}, Hi[2] = {0, 1};
  double Hx[2] = {2, 2};
  nasoq::CSC H{};
  H.nzmax = nnzH;
  H.ncol = H.nrow = sizeH;
  H.p = Hp;
  H.i = Hi;
  H.x = Hx;
  H.stype = -1;
  H.packed = 1;
  nasoq::SolverSettings* lbl = new nasoq::SolverSettings(&H, q);
  lbl->ldl_variant = 4;
  lbl->req_ref_iter = 2;
  lbl->solver_mode = 0;
  lbl->reg_diag = std::pow(10, -9);
  if (lbl->symbolic_analysis() != 1) return 2;
  if (lbl->numerical_factorization() != 1) return 3;
  double* x = lbl->solve_only();
  if (!x || std::fabs(x[0] + 2) > 1e-12 || std::fabs(x[1] + 2) > 1e-12) return 4;
  double new_q[2] = {8, 8};
  lbl->solve_only(1, new_q);
  if (std::fabs(x[0] - 4) > 1e-12 || std::fabs(x[1] - 4) > 1e-12) return 5;
  std::printf("Solution: %g,%g\n", x[0], x[1]);
  delete lbl;
  delete[] x;
  return 0;
}
