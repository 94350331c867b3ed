"""FP32 factor with FP64 refinement (precision = LBLG_F32) vs the CPU oracle.

The factor is of S K S, S = diag(s) with s_i = 1 / sqrt(max_j |K_ij|) (one
max-norm equilibration step); the factor, D and the L / D / L^T sweeps run
in FP32 (FP32 MFMA for the dense updates); the residual b - K x, GMRES and
the solution update stay FP64.  Diagonals of S K S below 2^-12 are raised
to it (static pivoting at sqrt(eps_f32)); the refinement works on the
reference's unperturbed, unscaled K0.  Bar: the same inertia as the FP64
oracle's factor of the same scaled, perturbed matrix; L D L^T reproduces
P S K~ S P^T to an FP32 rounding bound scaled by the growth; target-driven
refinement reaches a backward error <= 1e-10 on K (recomputed by the
oracle) and the solution agrees with the oracle's refined FP64 solution.
"""
import numpy as np
import pytest

from nasoq_amd import lbl
from nasoq_amd import problems as P

from conftest import dense_from_factor

pytestmark = pytest.mark.gpu

F32 = 1
EPS32 = 2.0 ** -24


def engine32(n, Ap, Ai, Ax, reg=1e-9, perm=None):
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = reg
    s.precision = F32
    assert s.symbolic_analysis(perm=perm) == 1
    assert s.numerical_factorization() == 1
    return s


def oracle_factor(O, s, n, Ap, Ai, Ax, reg):
    return O.Factor(n, Ap, Ai, Ax, reg, s.sym_ints("Perm"), s.sym_ints("super"), s.sym_longs("pi"),
                    s.sym_ints("s"), s.sym_longs("p"))


REG_HAT = 2.0 ** -12  # the static-pivot threshold of the scaled matrix


def check_f32(O, s, n, Ap, Ai, Ax, B, reg0):
    # the oracle factors the same scaled, perturbed matrix S K~ S (s_i =
    # 1 / sqrt(max_j |K_ij|), recomputed here from K)
    sc = s.scaling()
    Kf = P.lower_to_full(n, Ap, Ai, Ax).tocsr()
    rmax = np.asarray(abs(Kf).max(axis=1).todense()).ravel()
    assert np.allclose(sc, np.where(rmax > 0, 1.0 / np.sqrt(np.where(rmax > 0, rmax, 1.0)), 1.0), rtol=1e-14)
    Axs = sc[Ai] * Ax * sc[np.repeat(np.arange(n), np.diff(Ap))]
    reg = REG_HAT
    F = oracle_factor(O, s, n, Ap, Ai, Axs, reg)
    assert s.inertia() == F.inertia()
    valL, dval, rows, permf = s.get_factor()
    sup, pi = s.sym_ints("super"), s.sym_longs("pi")
    px = s.sym_longs("p")[sup[:-1]]
    L, D = dense_from_factor(n, sup, pi, px, valL, rows, dval)
    K = P.lower_to_full(n, Ap, Ai, Ax).toarray()
    Kt = P.lower_to_full(n, Ap, Ai, Axs).toarray()
    d = np.diag(Kt).copy()
    Kt[np.arange(n), np.arange(n)] = np.where(np.abs(d) < reg, reg, d)
    Pm = np.eye(n)[:, permf]
    R = Pm @ L @ D @ L.T @ Pm.T
    growth = (np.abs(L) @ np.abs(D) @ np.abs(L).T).max()
    # FP32 rounding of every stored entry and of each update, over at most n terms
    assert np.max(np.abs(R - Kt)) <= 4 * n * EPS32 * max(1.0, growth)
    # refinement in FP64 to the 1e-10 backward error of the reference's definition
    s.target_bwd_err = 1e-10
    s.max_refine_rounds = 30
    X, bwd = s.solve_only(B, return_bwd=True)
    Fo = oracle_factor(O, s, n, Ap, Ai, Ax, reg0)  # the FP64 factor of K~ itself
    cond_est = np.linalg.cond(K)
    for c in range(B.shape[1]):
        e = O.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c])
        assert e <= 1e-10, (c, e)
        assert abs(e - bwd[c]) <= max(1e-3 * e, 1e-16)
        xr, _ = Fo.refine(B[:, c], Fo.solve(B[:, c]))
        assert np.linalg.norm(X[:, c] - xr) <= max(1e-8, 1e-9 * cond_est) * max(1.0, np.linalg.norm(xr))
    return F


def test_small_kkt_f32(oracle):
    for seed in range(3):
        n, Ap, Ai, Ax, B = P.kkt_small(60 + 40 * seed, 15 + 5 * seed, seed=seed, nrhs=2)
        print(f"case f32 small_kkt seed={seed} n={n}", flush=True)
        s = engine32(n, Ap, Ai, Ax)
        check_f32(oracle, s, n, Ap, Ai, Ax, B, 1e-9)


@pytest.mark.parametrize("group_w,grid,m", [(2048, 9, 60), (96, 12, 300)])
def test_wide_supernodes_f32(oracle, group_w, grid, m, monkeypatch):
    # panel path (w <= group_w) and the cooperative wide panel (w > group_w)
    monkeypatch.setenv("LBLG_GROUP_W", str(group_w))
    n, Ap, Ai, Ax, B = P.kkt_config(grid, m, 6, seed=5, nrhs=3)
    print(f"case f32 wide n={n} group_w={group_w} grid={grid}", flush=True)
    s = engine32(n, Ap, Ai, Ax)
    check_f32(oracle, s, n, Ap, Ai, Ax, B, 1e-9)


@pytest.mark.parametrize("nb,bs,group_w", [(8, 10, 2048), (4, 200, 2048), (4, 200, 96)])
def test_two_by_two_pivots_f32(oracle, nb, bs, group_w, monkeypatch):
    # zero-diagonal blocks: 2x2 pivots on the LDS, panel and cooperative paths
    monkeypatch.setenv("LBLG_GROUP_W", str(group_w))
    n, Ap, Ai, Ax, B = P.block_indefinite(nb, bs, seed=1, nrhs=2)
    print(f"case f32 2x2 n={n} nb={nb} bs={bs} group_w={group_w}", flush=True)
    s = engine32(n, Ap, Ai, Ax, perm=np.arange(n, dtype=np.int32))
    _, dval, _, _ = s.get_factor()
    assert np.count_nonzero(dval[n:]) > 0  # 2x2 pivots were taken
    check_f32(oracle, s, n, Ap, Ai, Ax, B, 1e-9)


def test_precision_switch_on_one_handle(oracle):
    # FP64 -> FP32 -> FP64 on the same handle: the value buffers follow the setting
    n, Ap, Ai, Ax, B = P.kkt_small(150, 40, seed=7, nrhs=1)
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = 1e-9
    s.target_bwd_err = 1e-10
    s.max_refine_rounds = 30
    assert s.symbolic_analysis() == 1
    out = []
    for prec in (0, 1, 0):
        s.precision = prec
        assert s.numerical_factorization() == 1
        x, bwd = s.solve_only(B[:, 0], return_bwd=True)
        assert oracle.backward_error(n, Ap, Ai, Ax, x, B[:, 0]) <= 1e-10
        out.append(x)
    assert np.linalg.norm(out[0] - out[2]) <= 1e-12 * max(1.0, np.linalg.norm(out[0]))
