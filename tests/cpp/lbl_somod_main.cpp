// The reference's examples/lbl_somod_main.cpp flow against the drop-in
// SolverSettings in mode 1: H = 2I, q = (-4,-4), rows x0 and x1 added and
// dropped with update_somod / downdate_somod.  Prints every solution; the
// caller checks them (argument "zero": the rows' right-hand sides are zero,
// which gives the vectors listed in the reference example's comments).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "nasoq/QP/linear_solver_wrapper.h"

int main(int argc, char** argv) {
  const bool zero = argc > 1 && std::strcmp(argv[1], "zero") == 0;
  size_t sizeH = 2;
  double q[2] = {-4, -4};
  int Hp[3] = {0, 1, 2}, Hi[2] = {0, 1};
  double Hx[2] = {2, 2};
  nasoq::CSC H{};
  H.nzmax = 2;
  H.ncol = H.nrow = sizeH;
  H.p = Hp;
  H.i = Hi;
  H.x = Hx;
  H.stype = -1;
  H.packed = 1;
  double d[2] = {0, 0};
  int Cp[3] = {0, 1, 2}, Ci[2] = {0, 1};
  double Cx[2] = {1, 1};
  nasoq::CSC C{};
  C.nzmax = 2;
  C.ncol = sizeH;
  C.nrow = 2;
  C.p = Cp;
  C.i = Ci;
  C.x = Cx;
  C.stype = 0;
  C.packed = 1;
  nasoq::SolverSettings* lbl = new nasoq::SolverSettings(&H, q, &C, d);
  lbl->ldl_variant = 2;
  lbl->req_ref_iter = 2;
  lbl->solver_mode = 1;
  lbl->reg_diag = std::pow(10, -9);
  if (lbl->symbolic_analysis() != 1) return 2;
  if (lbl->numerical_factorization() != 1) return 3;
  double* x = lbl->solve_only();
  std::printf("%.15g %.15g\n", x[0], x[1]);
  std::vector<int> inds{0};
  if (!zero) {
    d[0] = 2;
    d[1] = 1;
  }
  lbl->update_somod(inds, 1, d);
  lbl->update_factorization();
  x = lbl->solve_only();
  std::printf("%.15g %.15g %.15g %.15g\n", x[0], x[1], x[2], x[3]);
  inds = {1};
  lbl->update_somod(inds, 1, d);
  lbl->update_factorization();
  x = lbl->solve_only();
  std::printf("%.15g %.15g %.15g %.15g\n", x[0], x[1], x[2], x[3]);
  inds = {0};
  lbl->downdate_somod(inds, 1);
  lbl->update_factorization();
  x = lbl->solve_only();
  std::printf("%.15g %.15g %.15g %.15g\n", x[0], x[1], x[2], x[3]);
  delete lbl;
  return 0;
}
