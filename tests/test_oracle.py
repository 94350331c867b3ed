"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the restatement (a wrong
pivot rule, a dropped 2x2 sub-diagonal, a transposed panel, a wrong norm)
fails it:

  * oracle_sytf2 vs LAPACK dsytrf through scipy (same pivot sequence, same
    factors) and the reference's own known-answer test (tests/test_clapacke.cpp:
    4x4 reversed Hilbert, compared through the inverse);
  * oracle_factor: P L D L^T P^T = K~ to rounding, inertia vs eigvalsh
    (Sylvester), a dense single-supernode factor equal to dsytrf's;
  * solves vs numpy.linalg.solve, GMRES with an exact preconditioner;
  * spmv / backward_error vs dense arithmetic and vs the reference's own
    spmv_CSC.cpp / Norm.cpp compiled in oracle/_ref (skipped where absent);
  * planted mistakes in the oracle's outputs are caught by the same checks.
"""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from nasoq_amd import lbl
from nasoq_amd import problems as P


def sym_indef(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    return a + a.T


def unpack_sytf2(a, ipiv):
    """Dense (U, D) with A = U D U^T, U = P1 L1 P2 L2 ..., from LAPACK 'L' dsytrf storage."""
    n = a.shape[0]
    D = np.zeros((n, n))
    k = 0
    # LAPACK applies interchange k <-> ipiv(k)-1 before eliminating column(s) k
    Lfac = []
    while k < n:
        if ipiv[k] > 0:
            kp = ipiv[k] - 1
            Lk = np.eye(n)
            Lk[k + 1:, k] = a[k + 1:, k]
            D[k, k] = a[k, k]
            Lfac.append((k, kp, Lk))
            k += 1
        else:
            kp = -ipiv[k] - 1
            Lk = np.eye(n)
            Lk[k + 2:, k] = a[k + 2:, k]
            Lk[k + 2:, k + 1] = a[k + 2:, k + 1]
            D[k, k] = a[k, k]
            D[k + 1, k] = D[k, k + 1] = a[k + 1, k]
            D[k + 1, k + 1] = a[k + 1, k + 1]
            Lfac.append((k + 1, kp, Lk))
            k += 2
    # A = (P1 L1 P2 L2 ...) D (...)^T
    U = np.eye(n)
    for piv_row, kp, Lk in Lfac:
        Pk = np.eye(n)
        Pk[[piv_row, kp]] = Pk[[kp, piv_row]]
        U = U @ Pk @ Lk
    return U, D


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (5, 2), (17, 3), (40, 4), (63, 5)])
def test_sytf2_matches_lapack_dsytrf(oracle, n, seed):
    A = sym_indef(n, seed)
    a, ipiv, info = oracle.sytf2(A)
    ldu, ipiv_l, info_l = lapack.dsytrf(A, lower=1)
    assert info == info_l == 0
    assert np.array_equal(ipiv, ipiv_l)
    low = np.tril_indices(n)
    assert np.allclose(a[low], ldu[low], rtol=1e-10, atol=1e-12)
    U, D = unpack_sytf2(a, ipiv)
    assert np.max(np.abs(U @ D @ U.T - A)) <= 1e-12 * np.abs(A).max() * n


def test_sytf2_reversed_hilbert_inverse(oracle):
    # tests/test_clapacke.cpp:24-34,55-70 (col-major, 'L'): the factors are
    # compared through the inverse they give (there with dsytri; here the
    # factors solve A X = I with LAPACK dsytrs)
    n = 4
    A = np.array([[1.0 / (2 * n - i - j - 1) for j in range(n)] for i in range(n)])
    a, ipiv, info = oracle.sytf2(A)
    assert info == 0
    inv, info_i = lapack.dsytrs(a, ipiv, np.eye(n), lower=1)
    assert info_i == 0
    assert np.allclose(inv, np.linalg.inv(A), rtol=1e-8, atol=1e-8)
    ldu, ipiv_l, _ = lapack.dsytrf(A, lower=1)
    inv_l, _ = lapack.dsytrs(ldu, ipiv_l, np.eye(n), lower=1)
    assert np.allclose(inv, inv_l, rtol=1e-10, atol=1e-10)


def test_sytf2_two_by_two_pivot_and_zero_column(oracle):
    # zero diagonal forces a 2x2 pivot (BK rule absakk < alpha*colmax)
    A = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 3.0]])
    a, ipiv, info = oracle.sytf2(A)
    assert list(ipiv[:2]) == [-2, -2] and ipiv[2] == 3 and info == 0
    # an all-zero column is a zero pivot: info = its 1-based index
    Z = np.diag([1.0, 0.0, 2.0])
    _, ipiv_z, info_z = oracle.sytf2(Z)
    _, ipiv_lz, info_lz = lapack.dsytrf(Z, lower=1)
    assert info_z == info_lz == 2 and np.array_equal(ipiv_z, ipiv_lz)


def analyse(n, Ap, Ai, Ax, perm=None, **kw):
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    for k, v in kw.items():
        setattr(s, k, v)
    assert s.symbolic_analysis(perm=perm) == 1
    return s


def oracle_factor(O, s, n, Ap, Ai, Ax, reg):
    return O.Factor(n, Ap, Ai, Ax, reg, s.sym_ints("Perm"), s.sym_ints("super"), s.sym_longs("pi"),
                    s.sym_ints("s"), s.sym_longs("p"))


def perturbed_dense(n, Ap, Ai, Ax, reg):
    K = P.lower_to_full(n, Ap, Ai, Ax).toarray()
    d = np.diag(K).copy()
    K[np.arange(n), np.arange(n)] = np.where(np.abs(d) < reg, reg, d)
    return K


def recon_error(F, K):
    Pm = np.eye(F.n)[:, F.perm]
    L, D = F.dense_L(), F.dense_D()
    R = Pm @ L @ D @ L.T @ Pm.T
    growth = (np.abs(L) @ np.abs(D) @ np.abs(L).T).max()
    return np.max(np.abs(R - K)) / max(1.0, growth)


CASES = [("kkt", s) for s in range(4)] + [("blk", (8, 10, None)), ("blk", (20, 12, "nat")),
                                          ("blk", (6, 40, "nat")), ("blk", (6, 40, None))]


def make_case(kind, arg):
    if kind == "kkt":
        n, Ap, Ai, Ax, B = P.kkt_small(40 + 25 * arg, 10 + 5 * arg, seed=arg, nrhs=1)
        return n, Ap, Ai, Ax, B, None
    nb, bs, order = arg
    n, Ap, Ai, Ax, B = P.block_indefinite(nb, bs, seed=1)
    return n, Ap, Ai, Ax, B, (np.arange(n, dtype=np.int32) if order == "nat" else None)


@pytest.mark.parametrize("kind,arg", CASES)
def test_factor_reconstructs_and_inertia(oracle, kind, arg):
    n, Ap, Ai, Ax, B, perm = make_case(kind, arg)
    reg = 1e-9
    s = analyse(n, Ap, Ai, Ax, perm=perm)
    F = oracle_factor(oracle, s, n, Ap, Ai, Ax, reg)
    K = perturbed_dense(n, Ap, Ai, Ax, reg)
    assert recon_error(F, K) <= 1e-12
    ev = np.linalg.eigvalsh(K)
    tol = 1e-12 * np.abs(ev).max()
    assert F.inertia() == (int((ev > tol).sum()), int((ev < -tol).sum()), int((np.abs(ev) <= tol).sum()))
    if kind == "blk":
        # zero-diagonal dense blocks: the 2x2 pivot path (D sub-diagonal) is exercised
        assert np.count_nonzero(F.d_val[n:]) > 0
    # unrefined solve: small backward error w.r.t. K~ (pivots limited to
    # supernodes admit growth); refined on the unperturbed K: the dense solution
    b = B[:, 0]
    x = F.solve(b)
    assert np.linalg.norm(K @ x - b) <= 1e-8 * (np.linalg.norm(K) * np.linalg.norm(x) + np.linalg.norm(b))
    K0 = P.lower_to_full(n, Ap, Ai, Ax).toarray()
    xr, _ = F.refine(b, x, itr_max=4, mr=4)
    xd = np.linalg.solve(K0, b)
    assert np.linalg.norm(xr - xd) <= 1e-9 * np.linalg.cond(K0) * np.linalg.norm(xd)
    s.close()


def test_factor_planted_mistakes_are_caught(oracle):
    n, Ap, Ai, Ax, _, perm = make_case("blk", (6, 40, "nat"))
    reg = 1e-9
    s = analyse(n, Ap, Ai, Ax, perm=perm)
    K = perturbed_dense(n, Ap, Ai, Ax, reg)
    F = oracle_factor(oracle, s, n, Ap, Ai, Ax, reg)
    assert recon_error(F, K) <= 1e-12
    # dropped 2x2 sub-diagonal term
    i = int(np.flatnonzero(F.d_val[n:])[0])
    sub = F.d_val[n + i]
    F.d_val[n + i] = 0.0
    assert recon_error(F, K) > 1e-6
    F.d_val[n + i] = sub
    # wrong sign on one pivot
    j = int(np.argmax(np.abs(F.d_val[:n])))
    F.d_val[j] = -F.d_val[j]
    assert recon_error(F, K) > 1e-6
    F.d_val[j] = -F.d_val[j]
    # a transposed panel (rows and columns of one wide supernode swapped)
    sup = F.sup
    t = int(np.argmax(np.diff(sup)))
    c0, w = sup[t], sup[t + 1] - sup[t]
    r = int(F.pi[t + 1] - F.pi[t])
    panel = F.valL[F.p[c0]:F.p[c0] + r * w].reshape(w, r).copy()
    blk = panel[:, :w].copy()
    F.valL[F.p[c0]:F.p[c0] + r * w] = np.concatenate([blk.T, panel[:, w:]], axis=1).ravel()
    assert w > 1 and recon_error(F, K) > 1e-6
    s.close()


def test_dense_single_supernode_equals_dsytrf(oracle):
    # a fully dense symmetric matrix in natural order is one supernode; the
    # supernodal factor's D blocks equal dsytrf's on the same matrix
    n = 24
    A = sym_indef(n, 7)
    rows, cols = np.tril_indices(n)
    order = np.lexsort((rows, cols))
    Ai = rows[order].astype(np.int32)
    Ap = np.concatenate([[0], np.cumsum(np.bincount(cols, minlength=n))]).astype(np.int32)
    Ax = A[rows[order], cols[order]]
    s = analyse(n, Ap, Ai, Ax, perm=np.arange(n, dtype=np.int32))
    assert len(s.sym_ints("super")) == 2
    F = oracle_factor(oracle, s, n, Ap, Ai, Ax, 0.0)
    ldu, ipiv, _ = lapack.dsytrf(A, lower=1)
    # same 1x1/2x2 structure, same D values
    d_l = np.diag(ldu).copy()
    sub_l = np.zeros(n)
    k = 0
    while k < n:
        if ipiv[k] < 0:
            sub_l[k] = ldu[k + 1, k]
            k += 2
        else:
            k += 1
    assert np.array_equal(F.d_val[n:] != 0, sub_l != 0)
    assert np.allclose(F.d_val[:n], d_l, rtol=1e-10, atol=1e-12)
    assert np.allclose(F.d_val[n:], sub_l, rtol=1e-10, atol=1e-12)
    assert recon_error(F, A) <= 1e-12
    s.close()


def test_lbl_main_example(oracle):
    # examples/lbl_main.cpp:26-57: H = 2I; q = (-4,-4) -> (-2,-2); q = (8,8) -> (4,4)
    Ap = np.array([0, 1, 2], np.int32)
    Ai = np.array([0, 1], np.int32)
    Ax = np.array([2.0, 2.0])
    s = analyse(2, Ap, Ai, Ax)
    F = oracle_factor(oracle, s, 2, Ap, Ai, Ax, 1e-9)
    assert np.array_equal(F.solve(np.array([-4.0, -4.0])), [-2.0, -2.0])
    assert np.array_equal(F.solve(np.array([8.0, 8.0])), [4.0, 4.0])
    s.close()


def test_gmres_refinement(oracle):
    n, Ap, Ai, Ax, B = P.kkt_config(6, 40, 6, seed=3, nrhs=2)
    reg = 1e-6  # a visible perturbation: the unrefined solve is off, refinement recovers
    s = analyse(n, Ap, Ai, Ax)
    F = oracle_factor(oracle, s, n, Ap, Ai, Ax, reg)
    K = P.lower_to_full(n, Ap, Ai, Ax).toarray()
    for c in range(2):
        b = B[:, c]
        x0 = F.solve(b)
        e0 = oracle.backward_error(n, Ap, Ai, Ax, x0, b)
        x1, _ = F.refine(b, x0, itr_max=4, mr=4)
        e1 = oracle.backward_error(n, Ap, Ai, Ax, x1, b)
        assert e1 < e0 and e1 <= 1e-14
        assert np.allclose(x1, np.linalg.solve(K, b), rtol=1e-8, atol=1e-9)
    # exact preconditioner (no perturbation): converged after one inner step
    F2 = oracle_factor(oracle, s, n, Ap, Ai, Ax, 0.0)
    x2, _ = F2.refine(B[:, 0], F2.solve(B[:, 0]), itr_max=1, mr=1)
    assert oracle.backward_error(n, Ap, Ai, Ax, x2, B[:, 0]) <= 1e-15
    s.close()


def test_spmv_and_backward_error_dense(oracle):
    n, Ap, Ai, Ax, B = P.kkt_small(70, 20, seed=4, nrhs=1)
    K = P.lower_to_full(n, Ap, Ai, Ax).toarray()
    x = np.random.default_rng(2).uniform(-1, 1, n)
    assert np.allclose(oracle.spmv(n, Ap, Ai, Ax, x), K @ x, rtol=1e-14, atol=1e-13)
    b = B[:, 0]
    # ||b - Kx||_1 / (||K||_inf ||x||_1 + ||b||_1): norm_dense(1, n, v, 0) is the sum
    # of magnitudes (Norm.cpp:15-54), the matrix norm the largest row sum (Norm.cpp:115-185)
    want = np.abs(b - K @ x).sum() / (np.abs(K).sum(axis=1).max() * np.abs(x).sum() + np.abs(b).sum())
    assert np.isclose(oracle.backward_error(n, Ap, Ai, Ax, x, b), want, rtol=1e-12)
    # a plausible mistake (infinity norms instead) gives a different value
    wrong = np.abs(b - K @ x).max() / (np.abs(K).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())
    assert not np.isclose(want, wrong, rtol=1e-3)


def test_spmv_and_backward_error_vs_reference_build(oracle, ref):
    n, Ap, Ai, Ax, B = P.kkt_small(90, 25, seed=8, nrhs=1)
    x = np.random.default_rng(5).uniform(-1, 1, n)
    yo = oracle.spmv(n, Ap, Ai, Ax, x)
    yr = np.zeros(n)
    ref.ref_spmv_sym(n, oracle._ip(Ap), oracle._ip(Ai), oracle._dp(Ax), oracle._dp(x), oracle._dp(yr))
    assert np.allclose(yo, yr, rtol=1e-14, atol=1e-14)
    eo = oracle.backward_error(n, Ap, Ai, Ax, x, B[:, 0])
    er = ref.ref_backward_error(n, oracle._ip(Ap), oracle._ip(Ai), oracle._dp(Ax), oracle._dp(x),
                                oracle._dp(np.ascontiguousarray(B[:, 0])))
    assert np.isclose(eo, er, rtol=1e-13)


def test_blas3_cpu_restatement_matches_oracle(oracle):
    # the bench's CPU baseline (cpu_lbl.c: DGEMM / DSYTRF / DTRSM, level-parallel)
    # factors the same matrix as oracle_factor: same D block structure and
    # inertia, P L D L^T P^T = K~, solutions equal to the oracle's
    from nasoq_amd import lbl
    from nasoq_amd import problems as P
    if oracle.cpu_lib() is None:
        pytest.skip("no BLAS library for the CPU restatement")
    for n_, Ap, Ai, Ax, B, perm in [P.kkt_config(8, 60, 6, seed=2, nrhs=3) + (None,),
                                    P.block_indefinite(4, 60, seed=3, nrhs=3) + ("nat",)]:
        s = lbl.SolverSettings(n_, Ap, Ai, Ax)
        s.reg_diag = 1e-9
        s.symbolic_analysis(perm=np.arange(n_, dtype=np.int32) if perm else None)
        args = (n_, Ap, Ai, Ax, 1e-9, s.sym_ints("Perm"), s.sym_ints("super"), s.sym_longs("pi"), s.sym_ints("s"),
                s.sym_longs("p"))
        F = oracle.Factor(*args)
        from conftest import dense_from_factor
        K = P.lower_to_full(n_, Ap, Ai, Ax).toarray()
        d = np.diag(K).copy()
        K[np.arange(n_), np.arange(n_)] = np.where(np.abs(d) < 1e-9, 1e-9, d)
        sup, pi = s.sym_ints("super"), s.sym_longs("pi")
        px = s.sym_longs("p")[sup[:-1]]
        for threads in (1, 4):
            C = oracle.CpuFactor(*args, s.sym_ints("sParent"), threads)
            assert np.array_equal(C.d_val[n_:] != 0, F.d_val[n_:] != 0)
            assert C.inertia() == F.inertia()
            L, D = dense_from_factor(n_, sup, pi, px, C.valL, C.s_lab, C.d_val)
            Pm = np.eye(n_)[:, C.perm]
            growth = (np.abs(L) @ np.abs(D) @ np.abs(L).T).max()
            assert np.abs(Pm @ L @ D @ L.T @ Pm.T - K).max() <= 1e-12 * max(1.0, growth)
            X = C.solve(B)
            for c in range(B.shape[1]):  # unrefined solves: rounding amplified by the 1e-9 pivots
                xo = F.solve(B[:, c])
                assert np.linalg.norm(X[:, c] - xo) <= 1e-6 * max(1.0, np.linalg.norm(xo))
        s.close()
