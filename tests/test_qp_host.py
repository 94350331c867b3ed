"""Host-side QP layer (no GPU): the SMP reader, the super-KKT of
build_super_matrix and the refinement settings each NASOQ variant derives."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from nasoq_amd import qp, smp

from qp_cases import GOLDEN, nasoq_main_problem, out_correct


@pytest.mark.parametrize("name", ["matt_qp1", "test05_0"])
def test_smp_reader_sizes(name):
    # the problem dimensions out_correct.csv records for each file
    row = out_correct()[("NASOQ-Fixed", name)]
    d = smp.load_smp(os.path.join(GOLDEN, name + ".yml"))
    H = d["H"]
    full = H + sp.tril(H, -1).T
    assert H.shape[0] == int(row["Hessian dim"])
    assert full.nnz == int(row["Hessian NNZ"]) or H.nnz == int(row["Hessian NNZ"])
    assert d["A"].shape[0] == int(row["# of Eq Constraints"]) and d["A"].nnz == int(row["Eq Constraint NNZ"])
    assert d["B"].shape[0] == int(row["# of Ineq Const"]) and d["B"].nnz == int(row["Ineq Constraint NNZ"])


def test_super_kkt_structure():
    # SM = lower part of [H A' B'; A 0 0; B 0 0] with the inequality values held
    # at zero until activated, one explicit diagonal slot per appended row
    rng = np.random.default_rng(3)
    n, me, mi = 12, 3, 7
    R = sp.random(n, n, density=0.3, random_state=rng)
    H = sp.csc_matrix(R @ R.T + sp.identity(n))
    A = sp.random(me, n, density=0.4, random_state=rng, format="csc")
    B = sp.random(mi, n, density=0.4, random_state=rng, format="csc")
    m = qp.Nasoq(H, rng.standard_normal(n), A, rng.standard_normal(me), B, rng.standard_normal(mi))
    m.variant = qp.PREDET
    m.symbolic_QP()
    N, p, i, x = m.super_kkt()
    S = sp.csc_matrix((x, i, p), shape=(N, N)).toarray()
    want = np.zeros((N, N))
    want[:n, :n] = np.tril(H.toarray())
    want[n:n + me, :n] = A.toarray()
    assert np.array_equal(S, want)
    pattern = sp.csc_matrix((np.ones_like(x), i, p), shape=(N, N)).toarray() != 0
    want_pat = np.zeros((N, N), bool)
    want_pat[:n, :n] = np.tril(H.toarray()) != 0
    want_pat[n:n + me, :n] = A.toarray() != 0
    want_pat[n + me:, :n] = B.toarray() != 0
    want_pat[np.arange(N), np.arange(N)] = True
    assert np.array_equal(pattern, want_pat)
    # rows sorted, diagonal first in every column (the engine's CSC contract)
    for j in range(N):
        assert i[p[j]] == j and np.all(np.diff(i[p[j]:p[j + 1]]) > 0)
    m.close()


@pytest.mark.parametrize("name", ["matt_qp1", "test05_0"])
@pytest.mark.parametrize("tool,variant,max_iter", [("NASOQ-Fixed", qp.FIXED, 0), ("NASOQ", qp.PREDET, 2)])
def test_variant_refinement_settings(name, tool, variant, max_iter):
    # determine_iterations / PREDET: the GMRES budget, tolerance and
    # perturbation out_correct.csv records for each run
    row = out_correct()[(tool, name)]
    d = smp.load_smp(os.path.join(GOLDEN, name + ".yml"))
    m = qp.Nasoq(d["H"], d["q"], d["A"] if d["A"].shape[0] else None, d["a"], d["B"], d["b"])
    m.eps_abs = float(row["eps_abs"])
    m.variant = variant
    m.max_iter = max_iter
    m.symbolic_QP()
    assert m.outer_iter_ref == int(row["Outer GMRES Iter"]) and m.inner_iter_ref == int(row["Inner GMRES Iter"])
    assert m.stop_tol == float(row["GMRES Tol"]) and m.diag_perturb == float(row["Diagonal Pert"])
    m.close()


def test_invalid_problem_rejected():
    import ctypes
    from nasoq_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    q = np.array([1.0, 1.0])
    Hx = np.array([2.0, 1.0, 2.0])
    # an upper-triangle entry (row 0 of column 1) is rejected
    Hp, Hi = np.array([0, 1, 3], np.int32), np.array([0, 0, 1], np.int32)
    rc = lib.lblg_qp_create(ctypes.byref(h), 2, _lib.iptr(Hp), _lib.iptr(Hi), _lib.dptr(Hx), _lib.dptr(q), 0, None,
                            None, None, None, 0, None, None, None, None)
    assert rc == -1
    # unsorted rows are rejected
    Hp, Hi = np.array([0, 2, 3], np.int32), np.array([1, 0, 1], np.int32)
    rc = lib.lblg_qp_create(ctypes.byref(h), 2, _lib.iptr(Hp), _lib.iptr(Hi), _lib.dptr(Hx), _lib.dptr(q), 0, None,
                            None, None, None, 0, None, None, None, None)
    assert rc == -1
