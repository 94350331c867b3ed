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


def _swept(b, rhs):
    """Solve with every launch timed; returns X and the launches of the
    per-system sweeps."""
    import ctypes
    lib = b.lib
    assert lib.lblg_set_kernel_timing(b.h, 1) == 0
    X = b.solve_only(rhs)
    tot, cnt = ctypes.c_double(), ctypes.c_int64()
    lib.lblg_kernel_time(b.h, b"solve_batch", ctypes.byref(tot), ctypes.byref(cnt))
    assert lib.lblg_set_kernel_timing(b.h, 0) == 0
    return X, cnt.value


@pytest.mark.parametrize("prec,nrhs", [(0, 1), (0, 3), (1, 1)])
def test_batch_sweeps_one_workgroup_per_system(oracle, prec, nrhs, monkeypatch):
    # one workgroup sweeps each system (its block in LDS, or in memory for the
    # 118 x 118 system, beyond the LDS at any nrhs) against the level sweeps of
    # the same batch: per-system backward errors and the refined solutions
    inst = P.batch_qp_kkts(7, seed=21, nmin=8, nmax=48, meq_max=30) + \
        P.batch_qp_kkts(1, seed=22, nmin=118, nmax=118, meq_max=20)
    sizes = [x[0] for x in inst]
    print(f"case batch sweeps prec={prec} nrhs={nrhs} sizes={sizes}", flush=True)
    rng = np.random.default_rng(5)
    rhs = np.concatenate([np.column_stack([x[4]] + [rng.standard_normal(x[0]) for _ in range(nrhs - 1)])
                          for x in inst])

    def run(level_sweeps):
        if level_sweeps:
            monkeypatch.setenv("LBLG_BATCH_SWEEPS", "0")
        else:
            monkeypatch.delenv("LBLG_BATCH_SWEEPS", raising=False)
        b = lbl.BatchSolverSettings([(n, Ap, Ai, Ax) for n, Ap, Ai, Ax, _ in inst])
        b.reg_diag = 1e-9
        b.precision = prec
        b.target_bwd_err = 1e-12 if prec == 0 else 1e-10
        b.max_refine_rounds = 30
        assert b.symbolic_analysis() == 1 and b.numerical_factorization() == 1
        X, launches = _swept(b, rhs)
        X = np.asarray(X).reshape(rhs.shape)
        e = b.backward_errors(rhs, X)
        offs = list(b.offsets)
        b.close()
        return X, launches, e, offs

    Xs, ls, es, offs = run(False)
    Xl, ll, el, _ = run(True)
    assert ls >= 1 and ll == 0, (ls, ll)
    assert es.max() <= 1e-10 and el.max() <= 1e-10
    for k, (n, Ap, Ai, Ax, _) in enumerate(inst):
        o = offs[k]
        for c in range(nrhs):
            assert oracle.backward_error(n, Ap, Ai, Ax, Xs[o:o + n, c], rhs[o:o + n, c]) <= 1e-10, (k, c)
            ref = Xl[o:o + n, c]
            assert np.linalg.norm(Xs[o:o + n, c] - ref) <= 1e-7 * max(1.0, np.linalg.norm(ref)), (k, c)
    print(f"per-system sweeps: {ls} launches; worst backward error {es.max():.2e} (level sweeps {el.max():.2e})",
          flush=True)
