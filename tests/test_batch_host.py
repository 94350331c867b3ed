"""Batched systems as one block-diagonal handle: the symbolic result of the
batch is the union of the blocks' own results, and the per-block backward
error follows the reference's definition (host-side, no GPU needed)."""
import numpy as np
import pytest

from nasoq_amd import lbl
from nasoq_amd import problems as P


def _blocks(count=6, seed=3):
    return P.batch_qp_kkts(count, seed=seed, nmin=6, nmax=24, meq_max=8)


def test_batch_symbolic_is_union_of_blocks():
    inst = _blocks()
    b = lbl.BatchSolverSettings([(n, Ap, Ai, Ax) for n, Ap, Ai, Ax, _ in inst])
    b.reg_diag = 1e-9
    assert b.symbolic_analysis() == 1
    assert b.n == sum(x[0] for x in inst)
    perm = b.sym_ints("Perm")
    fl_b, lnz_b = b.flops()
    fl, lnz, nsup = 0.0, 0.0, 0
    for k, (n, Ap, Ai, Ax, _) in enumerate(inst):
        s = lbl.SolverSettings(n, Ap, Ai, Ax)
        s.reg_diag = 1e-9
        assert s.symbolic_analysis() == 1
        f, z = s.flops()
        fl += f
        lnz += z
        nsup += len(s.sym_ints("super")) - 1
        o = b.offsets[k]
        # the block's ordering is its own nested dissection, shifted
        assert np.array_equal(np.sort(perm[o:o + n]), np.arange(o, o + n))
        assert np.array_equal(perm[o:o + n] - o, s.sym_ints("Perm"))
    assert fl_b == fl and lnz_b == lnz
    assert len(b.sym_ints("super")) - 1 == nsup


def test_batch_backward_errors_match_definition(oracle):
    inst = _blocks(4, seed=5)
    b = lbl.BatchSolverSettings([(n, Ap, Ai, Ax) for n, Ap, Ai, Ax, _ in inst])
    rng = np.random.default_rng(0)
    X = rng.standard_normal((b.n, 2))
    B = rng.standard_normal((b.n, 2))
    err = b.backward_errors(B, X)
    for k, (n, Ap, Ai, Ax, _) in enumerate(inst):
        o = b.offsets[k]
        for c in range(2):
            e = oracle.backward_error(n, Ap, Ai, Ax, X[o:o + n, c], B[o:o + n, c])
            assert abs(err[k, c] - e) <= 1e-14 * max(1.0, e)


def test_batch_rejects_invalid_blocks():
    n, Ap, Ai, Ax, _ = _blocks(1)[0]
    bad = Ai.copy()
    bad[Ap[1]] = 0  # diagonal not first in column 1
    with pytest.raises(RuntimeError):
        lbl.BatchSolverSettings([(n, Ap, Ai, Ax), (n, Ap, bad, Ax)])
