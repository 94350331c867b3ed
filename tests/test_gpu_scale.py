"""Parity at configuration scale (need a gfx950 GPU).

- The plain-LBL configuration itself (64^3 Laplacian + 1e-3 I, 16,384
  constraint rows, N = 278,528, 4 right-hand sides): the inertia equals the
  closed form (n, m, 0) (H SPD and A of full row rank, the perturbed zero
  block keeps it), the refined backward error recomputed by the oracle is
  <= 1e-10, and a second factorization of the same values gives the same
  1x1 / 2x2 block structure, pivot order and inertia (the Schur scatter-adds
  are atomic, so values may differ in the last bits between runs).
- 32^3 (N = 34,816) with a 256-column group width, so the cooperative panels
  run with tens of workgroups: valL, d_val, the pivoted row labels and the
  final permutation against the CPU checker entry by entry (the BLAS-3
  restatement oracle/cpu_lbl.c, itself pinned to the oracle's factor in
  tests/test_oracle.py; the oracle itself when no BLAS library is present).
  Pivots, 2x2 blocks, permutation and row labels must be identical; values
  agree to 8 u rho max(1, |entry|) (u = 2^-53, rho = max|D| / max|K| the
  growth of the factorization): constraint rows ordered before their
  couplings pivot on the perturbed zero diagonal (1e-9), so rho is ~4e8 and
  rows below such pivots hold entries up to ~2e9; two correct FP64
  factorizations that sum in different orders differ by O(u rho) (the BLAS-3
  restatement and the oracle's scalar loop differ by 1.4e-9 on this matrix);
  a wrong update or pivot step moves entries by O(1e-3 .. 1) relative.
"""
import numpy as np
import pytest

from nasoq_amd import lbl
from nasoq_amd import problems as P

pytestmark = pytest.mark.gpu


def test_kkt64_configuration(oracle):
    n, Ap, Ai, Ax, B = P.kkt_config(64, 16384, 8, seed=0, nrhs=4)
    print(f"case kkt64 n={n}", flush=True)
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = 1e-9
    s.cost_param = 16
    s.target_bwd_err = 1e-10
    s.max_refine_rounds = 20
    assert s.symbolic_analysis() == 1
    assert s.numerical_factorization() == 1
    assert s.inertia() == (64 ** 3, 16384, 0)
    _, d1, rows1, perm1 = s.get_factor()
    X, bwd = s.solve_only(B, return_bwd=True)
    for c in range(B.shape[1]):
        assert oracle.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c]) <= 1e-10
    assert s.numerical_factorization() == 1
    _, d2, rows2, perm2 = s.get_factor()
    assert np.array_equal(d1[n:] != 0, d2[n:] != 0)
    assert np.array_equal(perm1, perm2) and np.array_equal(rows1, rows2)
    assert s.inertia() == (64 ** 3, 16384, 0)
    assert np.abs(d1[:n] - d2[:n]).max() <= 1e-9 * np.abs(d1[:n]).max()
    s.close()


@pytest.mark.parametrize("tall", [0, 1])
def test_kkt32_factor_entries(oracle, tall, monkeypatch):
    monkeypatch.setenv("LBLG_GROUP_W", "256")
    monkeypatch.setenv("LBLG_WIDE_TALL", str(tall))
    n, Ap, Ai, Ax, B = P.kkt_config(32, 2048, 8, seed=0, nrhs=2)
    print(f"case kkt32 entries n={n} tall={tall}", flush=True)
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = 1e-9
    s.cost_param = 16
    assert s.symbolic_analysis() == 1
    sup, pi = np.array(s.sym_ints("super")), np.array(s.sym_longs("pi"))
    w, r = np.diff(sup), np.diff(pi)
    big = w > 256
    rows_per_group = 448 if tall else 192
    groups = np.ceil((r[big] - 64) / rows_per_group)
    assert groups.max() >= 10, groups  # cooperative panels with tens of workgroups
    assert s.numerical_factorization() == 1
    valL, dval, rows, permf = s.get_factor()
    args = (n, Ap, Ai, Ax, 1e-9, s.sym_ints("Perm"), s.sym_ints("super"), s.sym_longs("pi"), s.sym_ints("s"),
            s.sym_longs("p"))
    if oracle.cpu_lib() is not None:
        F = oracle.CpuFactor(*args, s.sym_ints("sParent"), 16)
        ref_valL, ref_d, ref_rows, ref_perm = F.valL, F.d_val, F.s_lab, F.perm
    else:
        F = oracle.Factor(*args)
        ref_valL, ref_d, ref_rows, ref_perm = F.valL, F.d_val, F.s_lab, F.perm
    assert np.array_equal(dval[n:] != 0, ref_d[n:] != 0)
    assert np.array_equal(permf, ref_perm)
    # row labels of the rows below each diagonal block (the diagonal block's
    # own rows are its columns, in order)
    below = np.ones(len(rows), dtype=bool)
    for t in range(len(w)):
        below[pi[t]:pi[t] + w[t]] = False
    assert np.array_equal(rows[below], ref_rows[below])
    u = 2.0 ** -53
    max_d, max_k = np.abs(ref_d).max(), np.abs(Ax).max()
    rho = max_d / max_k
    dd = np.abs(dval - ref_d).max()
    assert dd <= 8 * u * max_d, (dd, max_d)
    # every stored entry of L (unit diagonal blocks; upper parts of the
    # diagonal blocks are not part of L)
    p = s.sym_longs("p")
    worst = (0.0, -1)
    for t in range(len(w)):
        c0, wt, rt = sup[t], w[t], r[t]
        px = int(p[c0])
        a = valL[px:px + wt * rt].reshape(wt, rt).T
        b = ref_valL[px:px + wt * rt].reshape(wt, rt).T
        low = np.tril(np.ones((rt, wt), dtype=bool))
        rel = (np.abs(a - b) / np.maximum(1.0, np.abs(b)))[low]
        worst = max(worst, (float(rel.max()), t))
    print(f"growth rho {rho:.2e}; worst |dL| / max(1, |L|) {worst[0]:.2e} (supernode {worst[1]}), |dD| {dd:.2e}, "
          f"bar 8 u rho = {8 * u * rho:.2e}", flush=True)
    assert worst[0] <= 8 * u * rho, worst
    s.close()
