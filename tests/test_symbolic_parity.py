"""Host symbolic analysis vs the reference's own compiled symbolic code.

oracle/_ref/libnasoq_ref.so is built from the reference's sources
(etreeC, postOrderC, rowcolcounts, super_node_detection, computeCostperBlock,
getCoarseLevelSet_6) with a driver that follows symbolic_analysis_lin_solve
(src/symbolic/symbolic_phase.cpp:167-890) for a given permutation.  Every
output must agree bit for bit.
"""
import ctypes

import numpy as np
import pytest

from nasoq_amd import lbl
from nasoq_amd import problems as P

SCALARS = ["n", "nsuper", "n_level", "n_par", "max_sup_wid", "max_col", "xsize", "ssize"]


def run_ref(O, ref, n, Ap, Ai, perm, extra, cost):
    nr = np.array([4, 16, 48], np.int32)
    zr = np.array([0.8, 0.1, 0.05])
    ex = None if extra is None else np.ascontiguousarray(extra, np.int32)
    R = ref.ref_symbolic(n, O._ip(Ap), O._ip(Ai), O._ip(perm), O._ip(ex), cost, -3, 50, O._ip(nr), O._dp(zr))
    out = {}
    for k, name in enumerate(SCALARS):
        out[name] = ref.ref_symbolic_scalar(R, k)
    for k, name in enumerate(lbl.SYM_INTS):
        ln = ref.ref_symbolic_ints(R, k, None)
        a = np.zeros(max(ln, 1), np.int32)
        ref.ref_symbolic_ints(R, k, O._ip(a))
        out[name] = a[:ln]
    for k, name in enumerate(lbl.SYM_LONGS):
        ln = ref.ref_symbolic_longs(R, k, None)
        a = np.zeros(ln, np.int64)
        ref.ref_symbolic_longs(R, k, O._lp(a))
        out[name] = a
    ref.ref_symbolic_free(R)
    return out


def compare(O, ref, n, Ap, Ai, Ax, perm, extra=None, cost=4):
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.cost_param = cost
    assert s.symbolic_analysis(perm=perm, extra_cols=extra) == 1
    r = run_ref(O, ref, n, Ap, Ai, perm, extra, cost)
    for name in SCALARS:
        assert s.sym_scalar(name) == r[name], name
    for name in lbl.SYM_INTS:
        assert np.array_equal(s.sym_ints(name), r[name]), name
    for name in lbl.SYM_LONGS:
        assert np.array_equal(s.sym_longs(name), r[name]), name
    return s


@pytest.mark.parametrize("seed", range(24))
def test_random_kkt_random_perm(oracle, ref, seed):
    rng = np.random.default_rng(100 + seed)
    nh = int(rng.integers(5, 260))
    m = int(rng.integers(0, nh // 2 + 1))
    n, Ap, Ai, Ax, _ = P.kkt_small(nh, m, seed=seed, density=float(rng.uniform(0.005, 0.1)))
    perm = rng.permutation(n).astype(np.int32)
    extra = (rng.random(n) < 0.15).astype(np.int32) if seed % 3 == 0 else None
    compare(oracle, ref, n, Ap, Ai, Ax, perm, extra, int(rng.integers(1, 24)))


@pytest.mark.parametrize("g,cost", [(8, 2), (12, 8), (16, 16), (20, 3)])
def test_grid3d_nd(oracle, ref, g, cost):
    n, Ap, Ai, Ax = P.lower_csc(P.laplacian_3d(g, 1e-3))
    perm = lbl.nd_order(n, Ap, Ai)
    compare(oracle, ref, n, Ap, Ai, Ax, perm, None, cost)


@pytest.mark.parametrize("r,c,cost", [(40, 40, 8), (64, 33, 5), (100, 100, 16)])
def test_grid2d_nd_and_random(oracle, ref, r, c, cost):
    n, Ap, Ai, Ax = P.lower_csc(P.laplacian_2d(r, c, 1e-2))
    compare(oracle, ref, n, Ap, Ai, Ax, lbl.nd_order(n, Ap, Ai), None, cost)
    perm = np.random.default_rng(r * c).permutation(n).astype(np.int32)
    compare(oracle, ref, n, Ap, Ai, Ax, perm, None, cost)


def test_kkt_config_shape_nd(oracle, ref):
    # the plain-LBL KKT construction at a reduced grid (16^3, 512 constraint rows)
    n, Ap, Ai, Ax, _ = P.kkt_config(16, 512, 8, seed=3)
    compare(oracle, ref, n, Ap, Ai, Ax, lbl.nd_order(n, Ap, Ai), None, 8)


def test_kkt_configuration_64(oracle, ref):
    # the full plain-LBL configuration (64^3 Laplacian + 16,384 constraint
    # rows, N = 278,528) with the bench's level-schedule parameter
    n, Ap, Ai, Ax, _ = P.kkt_config(64, 16384, 8, seed=0)
    compare(oracle, ref, n, Ap, Ai, Ax, lbl.nd_order(n, Ap, Ai), None, 16)


def test_natural_order_tridiagonal_closed_form(oracle, ref):
    # tridiagonal: etree parent(i) = i+1, column counts 2 (last 1), one supernode
    n = 50
    import scipy.sparse as sp
    T = sp.diags([-np.ones(n - 1), 4 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    n, Ap, Ai, Ax = P.lower_csc(T)
    s = compare(oracle, ref, n, Ap, Ai, Ax, np.arange(n, dtype=np.int32), None, 1)
    assert np.array_equal(s.sym_ints("Parent")[:-1], np.arange(1, n))
    cc = s.sym_ints("ColCount")
    assert np.all(cc[:-1] == 2) and cc[-1] == 1
    # fundamental supernodes are single columns here; relaxation merges them
    assert 1 <= s.sym_scalar("nsuper") < n


def test_colcounts_match_dense_symbolic_elimination():
    # brute force: pattern of the Cholesky factor by dense symbolic elimination
    rng = np.random.default_rng(7)
    for t in range(6):
        n, Ap, Ai, Ax, _ = P.kkt_small(40, 10, seed=t, density=0.08)
        perm = rng.permutation(n).astype(np.int32)
        s = lbl.SolverSettings(n, Ap, Ai, Ax)
        s.symbolic_analysis(perm=perm)
        Pm = s.sym_ints("Perm")
        K = P.lower_to_full(n, Ap, Ai, Ax).toarray() != 0
        K = K[np.ix_(Pm, Pm)]
        F = K.copy()
        for k in range(n):
            nz = np.where(F[k + 1:, k])[0] + k + 1
            F[np.ix_(nz, nz)] = True
        cc = np.array([1 + np.count_nonzero(F[j + 1:, j]) for j in range(n)])
        assert np.array_equal(cc, s.sym_ints("ColCount"))
        par = s.sym_ints("Parent")
        for j in range(n):
            below = np.where(F[j + 1:, j])[0]
            assert par[j] == (below[0] + j + 1 if len(below) else -1)
