// The reference's examples/nasoq_main.cpp flow against the drop-in
// nasoq::Nasoq: min 1/2 x'Hx + q'x s.t. Cx <= d, expected x = {0.4, 1.2},
// z = {1.6, 0, 0, 0}.
#include <cmath>
#include <cstdio>

#include "nasoq/nasoq.h"

int main() {
  size_t sizeH = 2, CRows = 4, CCols = 2;
  double q[2] = {-4, -4};
  int Hp[3] = {0, 1, 2}, Hi[2] = {0, 1};
  double Hx[2] = {2, 2};
  int Cp[3] = {0, 4, 8}, Ci[8] = {0, 1, 2, 3, 0, 1, 2, 3};
  double Cx[8] = {2, 1, -1, -2, 1, -1, -1, 1};
  double d[4] = {2, 1, 1, 2};
  nasoq::Nasoq* qm = new nasoq::Nasoq(sizeH, Hp, Hi, Hx, q, CRows, CCols, Cp, Ci, Cx, d);
  qm->diag_perturb = std::pow(10, -9);
  qm->eps_abs = std::pow(10, -3);
  qm->max_iter = 0;
  qm->variant = nasoq::PREDET;
  int converged = qm->solve();
  if (converged != nasoq::Optimal) return 2;
  const double ex[2] = {0.4, 1.2}, ez[4] = {1.6, 0, 0, 0};
  for (int i = 0; i < 2; ++i)
    if (std::fabs(qm->primal_vars[i] - ex[i]) > 1e-12) return 3;
  for (int i = 0; i < 4; ++i)
    if (std::fabs(qm->dual_vars[i] - ez[i]) > 1e-12) return 4;
  std::printf("Primal variables: %g,%g\nDual variables: %g,%g,%g,%g\n", qm->primal_vars[0], qm->primal_vars[1],
              qm->dual_vars[0], qm->dual_vars[1], qm->dual_vars[2], qm->dual_vars[3]);
  delete qm;
  return 0;
}
