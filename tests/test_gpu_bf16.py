"""BF16 factor with FP64 refinement (precision = LBLG_BF16) vs the CPU oracle.

The factor is of S K~ S, S = diag(s) with s_i = 1 / sqrt(max_j |K_ij|) (one
max-norm equilibration step of every supernode's rows and columns); values,
panels and sweeps are FP32, the 128-row-tile Schur updates and the wide
panels' trailing updates take BF16 MFMA operands with FP32 accumulation.
Tiny diagonals of S K S below 2^-12 are raised to it (static pivoting).
The residual, GMRES and the solution update stay FP64 on the reference's
unperturbed, unscaled K0.

Bar: L D L^T reproduces P S K~ S P^T within a BF16 bound (2^-6 of the
growth |L||D||L^T|); its inertia equals that of the oracle's FP64 factor of
the same scaled, perturbed matrix except for eigenvalues within the
reconstruction error's 2-norm of zero (Weyl), and exactly when there are
none; on the cases where the BF16 updates run the reconstruction is
visibly worse than the FP32 factor of the same matrix (the BF16 rounding
did happen); refinement reaches a backward error <= 1e-10 on K
(recomputed by the oracle) and the solution agrees with the oracle's refined
FP64 solution.  The ill-conditioned variant (H's diagonal scaled by
10^U[-3,3]) reaches 1e-10 at FP64, FP32 and BF16.
"""
import numpy as np
import pytest

from nasoq_amd import lbl
from nasoq_amd import problems as P

from conftest import dense_from_factor

pytestmark = pytest.mark.gpu

F64, F32, BF16 = 0, 1, 2
REG_HAT = 2.0 ** -12  # the static-pivot threshold of the scaled matrix


def factored(n, Ap, Ai, Ax, prec, perm=None, reg=1e-9):
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = reg
    s.precision = prec
    assert s.symbolic_analysis(perm=perm) == 1
    assert s.numerical_factorization() == 1
    return s


def scaled_values(n, Ap, Ai, Ax, sc):
    col = np.repeat(np.arange(n), np.diff(Ap))
    return sc[Ai] * Ax * sc[col]


def reconstruction(s, n, Ap, Ai, Axs, reg):
    """E = P L D L^T P^T - Ks~ of the handle's factor of Axs (dense), Ks~ and the growth max(|L||D||L^T|)."""
    valL, dval, rows, permf = s.get_factor()
    sup, pi = s.sym_ints("super"), s.sym_longs("pi")
    px = s.sym_longs("p")[sup[:-1]]
    L, D = dense_from_factor(n, sup, pi, px, valL, rows, dval)
    K = P.lower_to_full(n, Ap, Ai, Axs).toarray()
    d = np.diag(K).copy()
    K[np.arange(n), np.arange(n)] = np.where(np.abs(d) < reg, reg, d)
    Pm = np.eye(n)[:, permf]
    R = Pm @ L @ D @ L.T @ Pm.T
    return R - K, K, (np.abs(L) @ np.abs(D) @ np.abs(L).T).max()


def reconstruction_error(s, n, Ap, Ai, Axs, reg):
    E, _, growth = reconstruction(s, n, Ap, Ai, Axs, reg)
    return np.abs(E).max(), growth


def check_bf16(O, n, Ap, Ai, Ax, B, perm=None, bf16_ran=True):
    s = factored(n, Ap, Ai, Ax, BF16, perm=perm)
    sc = s.scaling()
    # s_i = 1 / sqrt(max_j |K_ij|): recomputed here from K
    K = P.lower_to_full(n, Ap, Ai, Ax).tocsr()
    rmax = np.asarray(abs(K).max(axis=1).todense()).ravel()
    assert np.allclose(sc, np.where(rmax > 0, 1.0 / np.sqrt(np.where(rmax > 0, rmax, 1.0)), 1.0), rtol=1e-14)
    Axs = scaled_values(n, Ap, Ai, Ax, sc)
    F = O.Factor(n, Ap, Ai, Axs, REG_HAT, s.sym_ints("Perm"), s.sym_ints("super"), s.sym_longs("pi"),
                 s.sym_ints("s"), s.sym_longs("p"))
    E, Kt, growth = reconstruction(s, n, Ap, Ai, Axs, REG_HAT)
    err16 = np.abs(E).max()
    assert err16 <= 2.0 ** -6 * max(1.0, growth), (err16, growth)
    # inertia: the factor's is that of R = Ks~ + E (Sylvester); by Weyl's
    # theorem it can differ from Ks~'s (the oracle's) only by eigenvalues of
    # Ks~ within ||E||_2 of zero
    lam = np.linalg.eigvalsh(Kt)
    e2 = np.linalg.norm(E, 2)
    near = int(np.sum(np.abs(lam) <= e2))
    pos, neg, zero = s.inertia()
    fpos, fneg, fzero = F.inertia()
    assert (fpos, fneg) == (int(np.sum(lam > 0)), int(np.sum(lam < 0)))
    print(f"bf16 inertia {(pos, neg, zero)} oracle {(fpos, fneg, fzero)} ||E||_2 {e2:.2e} "
          f"eigenvalues within it {near} (min |lambda| {np.abs(lam).min():.2e})", flush=True)
    assert abs(pos - fpos) <= near and abs(neg - fneg) <= near and zero == 0
    if near == 0:
        assert (pos, neg, zero) == (fpos, fneg, fzero)
    if bf16_ran:
        # the FP32 factor of K is the factor of the same S K~ S (same scaling,
        # same threshold 2^-12): its error is FP32 rounding, the BF16
        # factor's is that of the BF16 terms
        s32 = factored(n, Ap, Ai, Ax, F32, perm=perm)
        assert np.array_equal(s32.scaling(), sc)
        err32, _ = reconstruction_error(s32, n, Ap, Ai, Axs, REG_HAT)
        s32.close()
        print(f"reconstruction bf16 {err16:.2e} f32 {err32:.2e} growth {growth:.2e}", flush=True)
        assert err16 > 8 * err32, (err16, err32)
    s.target_bwd_err = 1e-10
    s.max_refine_rounds = 40
    X, bwd = s.solve_only(B, return_bwd=True)
    Fo = O.Factor(n, Ap, Ai, Ax, 1e-9, s.sym_ints("Perm"), s.sym_ints("super"), s.sym_longs("pi"),
                  s.sym_ints("s"), s.sym_longs("p"))
    cond_est = np.linalg.cond(K.toarray())
    for c in range(B.shape[1]):
        e = O.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c])
        assert e <= 1e-10, (c, e)
        assert abs(e - bwd[c]) <= max(1e-3 * e, 1e-16)
        xr, _ = Fo.refine(B[:, c], Fo.solve(B[:, c]))
        assert np.linalg.norm(X[:, c] - xr) <= max(1e-8, 1e-9 * cond_est) * max(1.0, np.linalg.norm(xr))
    rounds = s.num_ref_iter
    s.close()
    return rounds


def _launches(s, name):
    import ctypes
    tot, cnt = ctypes.c_double(), ctypes.c_int64()
    s.lib.lblg_kernel_time(s.h, name.encode(), ctypes.byref(tot), ctypes.byref(cnt))
    return cnt.value


def test_small_kkt_bf16(oracle):
    # no 128-tile updates at these sizes: the scaled FP32 path with its refinement
    for seed in range(3):
        n, Ap, Ai, Ax, B = P.kkt_small(60 + 40 * seed, 15 + 5 * seed, seed=seed, nrhs=2)
        print(f"case bf16 small_kkt seed={seed} n={n}", flush=True)
        check_bf16(oracle, n, Ap, Ai, Ax, B, bf16_ran=False)


@pytest.mark.parametrize("group_w,grid,m", [(96, 12, 300), (256, 14, 500)])
def test_bf16_updates(oracle, group_w, grid, m, monkeypatch):
    # cooperative wide panels (trailing updates) and 128-row-tile Schur
    # updates (from 128 off-diagonal rows) run with BF16 operands
    monkeypatch.setenv("LBLG_GROUP_W", str(group_w))
    monkeypatch.setenv("LBLG_SCHUR_BIG_M", "128")
    n, Ap, Ai, Ax, B = P.kkt_config(grid, m, 6, seed=5, nrhs=3)
    print(f"case bf16 updates n={n} group_w={group_w}", flush=True)
    s = factored(n, Ap, Ai, Ax, BF16)
    assert s.lib.lblg_set_kernel_timing(s.h, 1) == 0
    s.numerical_factorization()
    assert _launches(s, "schur_big") > 0 and _launches(s, "trailing") > 0
    s.close()
    rounds = check_bf16(oracle, n, Ap, Ai, Ax, B)
    print(f"bf16 refinement rounds {rounds}", flush=True)


@pytest.mark.parametrize("nb,bs,group_w", [(4, 200, 96), (3, 600, 256)])
def test_two_by_two_pivots_bf16(oracle, nb, bs, group_w, monkeypatch):
    # zero-diagonal blocks force 2x2 pivots in the cooperative panels
    monkeypatch.setenv("LBLG_GROUP_W", str(group_w))
    monkeypatch.setenv("LBLG_SCHUR_BIG_M", "128")
    n, Ap, Ai, Ax, B = P.block_indefinite(nb, bs, seed=1, nrhs=2)
    print(f"case bf16 2x2 n={n} nb={nb} bs={bs}", flush=True)
    check_bf16(oracle, n, Ap, Ai, Ax, B, perm=np.arange(n, dtype=np.int32))


@pytest.mark.parametrize("prec", [F64, F32, BF16])
def test_ill_conditioned_variant(oracle, prec, monkeypatch):
    # H's diagonal scaled by 10^U[-3,3]: refinement on K reaches 1e-10 at
    # every precision (the BF16 factor with its diagonal scaling)
    monkeypatch.setenv("LBLG_GROUP_W", "96")
    monkeypatch.setenv("LBLG_SCHUR_BIG_M", "128")
    n, Ap, Ai, Ax, B = P.kkt_config(12, 300, 6, seed=7, nrhs=2, ill=True)
    print(f"case ill n={n} prec={prec}", flush=True)
    s = factored(n, Ap, Ai, Ax, prec)
    s.target_bwd_err = 1e-10
    s.max_refine_rounds = 60
    X, bwd = s.solve_only(B, return_bwd=True)
    print(f"ill prec={prec} rounds={s.num_ref_iter} bwd={bwd.max():.2e}", flush=True)
    for c in range(B.shape[1]):
        assert oracle.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c]) <= 1e-10
    Fo = oracle.Factor(n, Ap, Ai, Ax, 1e-9, s.sym_ints("Perm"), s.sym_ints("super"), s.sym_longs("pi"),
                       s.sym_ints("s"), s.sym_longs("p"))
    assert s.inertia() == Fo.inertia()
    s.close()
