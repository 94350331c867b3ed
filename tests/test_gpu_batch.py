"""Batched KKT systems of uneven size factored and solved together in one
block-diagonal handle (one launch per level and kernel family for all
blocks), FP64 and FP32-factor, against the oracle's per-block backward
error and the solutions of the blocks solved one by one."""
import numpy as np
import pytest

from nasoq_amd import lbl
from nasoq_amd import problems as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_matches_one_by_one(oracle, prec):
    # sizes from a few hundred to a few thousand rows: LDS, panel and chain paths
    inst = P.batch_qp_kkts(12, seed=11, nmin=8, nmax=64, meq_max=40)
    print(f"case batch prec={prec} sizes={[x[0] for x in inst]}", flush=True)
    b = lbl.BatchSolverSettings([(n, Ap, Ai, Ax) for n, Ap, Ai, Ax, _ in inst])
    b.reg_diag = 1e-9
    b.precision = prec
    b.target_bwd_err = 1e-10  # refinement runs until the worst block meets it
    b.max_refine_rounds = 30
    assert b.symbolic_analysis() == 1
    assert b.numerical_factorization() == 1
    rhs = np.concatenate([x[4] for x in inst])
    X = b.solve_only(rhs)
    pos = neg = 0
    for k, (n, Ap, Ai, Ax, bk) in enumerate(inst):
        o = b.offsets[k]
        xk = X[o:o + n]
        assert oracle.backward_error(n, Ap, Ai, Ax, xk, bk) <= 1e-10, k
        s = lbl.SolverSettings(n, Ap, Ai, Ax)
        s.reg_diag = 1e-9
        s.target_bwd_err = 1e-12
        assert s.symbolic_analysis() == 1 and s.numerical_factorization() == 1
        x1 = s.solve_only(bk)
        assert np.linalg.norm(xk - x1) <= 1e-7 * max(1.0, np.linalg.norm(x1)), k
        if prec == 0:
            p, q, z = s.inertia()
            pos, neg = pos + p, neg + q
        s.close()
    if prec == 0:
        assert b.inertia() == (pos, neg, 0)
    e = b.backward_errors(rhs, X)
    assert e.max() <= 1e-10
    b.close()
