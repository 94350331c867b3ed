"""The SoMod update on the device (lblg_update_values + lblg_factor; need a
gfx950 GPU).

After entries of K change, the next factorization refactors only the
supernodes on the elimination-tree paths from the changed entries to the
root (the reference's update_factorization / update_ldl_left_sn_02_v2 with
the modified supernodes add_del_matrix_qp collects,
src/QP/linear_solver_wrapper.cpp:1058-1119).  Checked here:

- the number of refactored supernodes equals the size of that path set,
  recomputed from the symbolic tree (sParent, Perm, col2sup);
- the updated factor equals a full factorization of the same values on a
  fresh handle: the same pivots, 2x2 blocks, permutation and row labels,
  values to the larger of 64 u rho max(1, |entry|) and 16 times the spread
  between two fresh full factorizations (the scatter-adds are atomic, so
  sums round differently from run to run; rho = max|D| / max|K| is the
  growth);
- the refined solve reaches 1e-10 (backward error recomputed by the oracle);
- chains of add / drop edits (constraint rows switched on and off, as the
  active-set loop does), with cooperative panels and 128-row Schur tiles on
  the paths;
- the FP32 factor updates the same way and keeps the diagonal scaling of
  its last full factorization; refinement still reaches 1e-10;
- the cases that must refactor everything do (a changed reg_diag, values
  replaced with lblg_set_values, a factorization with nothing changed).
"""
import numpy as np
import pytest

from nasoq_amd import lbl
from nasoq_amd import problems as P

pytestmark = pytest.mark.gpu

U = 2.0 ** -53


def analysed(n, Ap, Ai, Ax, prec=0, reg=1e-9):
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = reg
    s.precision = prec
    assert s.symbolic_analysis() == 1
    return s


def path_set(s, n, Ap, Ai, idx):
    """Supernodes on the root paths of the changed entries, from the symbolic tree."""
    sup = np.array(s.sym_ints("super"))
    par = np.array(s.sym_ints("sParent"))
    iperm = np.argsort(np.array(s.sym_ints("Perm")))
    col2sup = np.repeat(np.arange(len(sup) - 1), np.diff(sup))
    col = np.repeat(np.arange(n), np.diff(Ap))
    out = set()
    for q in idx:
        t = col2sup[min(iperm[Ai[q]], iperm[col[q]])]
        while t >= 0 and t not in out:
            out.add(int(t))
            t = par[t]
    return out


def same_factor(s, n, Ap, Ai, Ax):
    """The factor of s against a full factorization of Ax on a new handle.

    The Schur updates are scatter-added with atomics, so two full
    factorizations of the same values differ by rounding in an order that
    changes from run to run; with pivots of reg_diag = 1e-9 beside O(1)
    couplings (growth rho ~ 1e8 here) unit-size entries are sums of terms
    near rho, and that spread reaches 8 u rho.  The bar is therefore the
    larger of 64 u rho and 16 times the spread measured between two fresh
    full factorizations (same pivots), relative to max(1, |entry|): two
    full factorizations can happen to sum in one order (spread 0) while the
    update groups its sums differently.  An error of the update itself (a
    missed sender, a stale panel) is many orders larger."""
    f = analysed(n, Ap, Ai, Ax, reg=s.reg_diag)
    assert f.numerical_factorization() == 1
    assert f.refactored_supernodes() == len(f.sym_ints("super")) - 1
    f2 = analysed(n, Ap, Ai, Ax, reg=s.reg_diag)
    assert f2.numerical_factorization() == 1
    valL, dval, rows, permf = s.get_factor()
    rvalL, rd, rrows, rperm = f.get_factor()
    v2, d2, rows2, perm2 = f2.get_factor()
    f.close()
    f2.close()
    assert np.array_equal(dval[n:] != 0, rd[n:] != 0)
    assert np.array_equal(permf, rperm)
    assert np.array_equal(rows, rrows)
    rho = np.abs(rd).max() / np.abs(Ax).max()
    sd, sl = np.maximum(1.0, np.abs(rd)), np.maximum(1.0, np.abs(rvalL))
    spread = 0.0
    if np.array_equal(d2[n:] != 0, rd[n:] != 0) and np.array_equal(perm2, rperm):
        spread = max((np.abs(d2 - rd) / sd).max(), (np.abs(v2 - rvalL) / sl).max())
    bar = max(64 * U * max(rho, 1.0), 16 * spread)
    dd = (np.abs(dval - rd) / sd).max()
    dl = (np.abs(valL - rvalL) / sl).max()
    assert dd <= bar and dl <= bar, (dd, dl, bar, spread, U * rho)


def coupling_positions(n, Ap, Ai, row):
    """Positions in the lower CSC of the entries of `row` left of the diagonal."""
    col = np.repeat(np.arange(n), np.diff(Ap))
    return np.nonzero((Ai == row) & (col < row))[0]


def refined_ok(oracle, s, n, Ap, Ai, Ax, B, dropped=()):
    # a dropped constraint row leaves K singular (zero couplings, zero
    # diagonal); its right-hand side is zero, as downdate_somod sets it
    # (linear_solver_wrapper.cpp:1008-1018)
    B = np.array(B, dtype=np.float64).reshape(n, -1)
    B[list(dropped)] = 0.0
    s.target_bwd_err = 1e-10
    s.max_refine_rounds = 20
    X = s.solve_only(B)
    X = X.reshape(n, -1)
    for c in range(X.shape[1]):
        assert oracle.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c]) <= 1e-10


@pytest.mark.parametrize("grid,m,env", [(8, 120, {}), (12, 300, {"LBLG_GROUP_W": "96", "LBLG_SCHUR_BIG_M": "128"})])
def test_path_update_matches_full_factor(oracle, grid, m, env, monkeypatch):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    n, Ap, Ai, Ax, B = P.kkt_config(grid, m, 6, seed=3, nrhs=1)
    nh = grid ** 3
    print(f"case somod n={n} env={env}", flush=True)
    s = analysed(n, Ap, Ai, Ax)
    ns = len(s.sym_ints("super")) - 1
    assert s.numerical_factorization() == 1
    assert s.refactored_supernodes() == ns
    rng = np.random.default_rng(0)
    Axc = Ax.copy()
    active = np.ones(m, dtype=bool)
    counts = []
    for step in range(6):
        # drop or re-add one or two constraint rows (their couplings to H)
        rows = rng.choice(m, size=1 + step % 2, replace=False)
        idx = np.concatenate([coupling_positions(n, Ap, Ai, nh + r) for r in rows])
        vals = np.where(np.repeat(active[rows], [len(coupling_positions(n, Ap, Ai, nh + r)) for r in rows]),
                        0.0, Ax[idx])
        active[rows] = ~active[rows]
        Axc[idx] = vals
        s.update_values(idx, vals)
        assert s.numerical_factorization() == 1
        expect = path_set(s, n, Ap, Ai, idx)
        assert s.refactored_supernodes() == len(expect)
        counts.append(len(expect))
        same_factor(s, n, Ap, Ai, Axc)
        refined_ok(oracle, s, n, Ap, Ai, Axc, B, dropped=nh + np.nonzero(~active)[0])
    print(f"refactored supernodes per update {counts} of {ns}", flush=True)
    assert max(counts) < ns
    s.close()


def test_full_refactor_cases(oracle):
    n, Ap, Ai, Ax, B = P.kkt_config(8, 120, 6, seed=4, nrhs=1)
    nh = 8 ** 3
    idx = coupling_positions(n, Ap, Ai, nh + 5)
    # FP32 factor: restricted too, with the scaling of the full factorization
    s = analysed(n, Ap, Ai, Ax, prec=1)
    ns = len(s.sym_ints("super")) - 1
    assert s.numerical_factorization() == 1
    sc = s.scaling()
    Axc = Ax.copy()
    Axc[idx] = 0.0
    s.update_values(idx, Axc[idx])
    assert s.numerical_factorization() == 1
    assert s.refactored_supernodes() == len(path_set(s, n, Ap, Ai, idx)) < ns
    assert np.array_equal(s.scaling(), sc)
    refined_ok(oracle, s, n, Ap, Ai, Axc, B, dropped=[nh + 5])
    s.close()
    # FP64: a changed reg_diag refactors everything, then updates are restricted again
    s = analysed(n, Ap, Ai, Ax)
    assert s.numerical_factorization() == 1
    Axc = Ax.copy()
    Axc[idx] = 0.0
    s.update_values(idx, Axc[idx])
    s.reg_diag = 1e-8
    assert s.numerical_factorization() == 1
    assert s.refactored_supernodes() == ns
    same_factor(s, n, Ap, Ai, Axc)
    Axc[idx] = Ax[idx]
    s.update_values(idx, Axc[idx])
    assert s.numerical_factorization() == 1
    assert 0 < s.refactored_supernodes() < ns
    same_factor(s, n, Ap, Ai, Axc)
    # lblg_set_values: everything again
    s.set_values(Axc)
    assert s.numerical_factorization() == 1
    assert s.refactored_supernodes() == ns
    # no change since the last factorization: everything again, as the
    # reference's numerical_factorization does
    assert s.numerical_factorization() == 1
    assert s.refactored_supernodes() == ns
    refined_ok(oracle, s, n, Ap, Ai, Axc, B)
    # out-of-range positions are rejected and change nothing
    with pytest.raises(RuntimeError):
        s.update_values(np.array([len(Ax)]), np.array([1.0]))
    s.close()
