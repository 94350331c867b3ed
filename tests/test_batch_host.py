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


def test_batch_supernodes_stay_within_their_system():
    # the per-system sweeps (one workgroup per system) rely on each system's
    # supernodes forming one contiguous range, in system order, whose row
    # indices stay within the system's columns; their levels (leaves 0, a
    # parent one above its highest child) are then the system's own
    inst = _blocks(count=9, seed=7)
    b = lbl.BatchSolverSettings([(n, Ap, Ai, Ax) for n, Ap, Ai, Ax, _ in inst])
    b.reg_diag = 1e-9
    assert b.symbolic_analysis() == 1
    sup = np.array(b.sym_ints("super"))
    pi = np.array(b.sym_longs("pi"))
    rows = np.array(b.sym_ints("s"))
    par = np.array(b.sym_ints("sParent"))
    offs = np.array(b.offsets)  # first column of each system, then n
    sysof = np.searchsorted(offs, sup[:-1], side="right") - 1
    assert np.all(np.diff(sysof) >= 0) and sysof[0] == 0 and sysof[-1] == len(inst) - 1
    for s in range(len(sup) - 1):
        g = sysof[s]
        assert offs[g] <= sup[s] and sup[s + 1] <= offs[g + 1]
        r = rows[pi[s]:pi[s + 1]]
        assert r.min() >= offs[g] and r.max() < offs[g + 1]
        if par[s] >= 0:
            assert sysof[par[s]] == g and par[s] > s


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
