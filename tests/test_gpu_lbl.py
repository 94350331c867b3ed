"""Device LBL engine vs the CPU oracle (parity tests; need a gfx950 GPU).

Bar: the 1x1/2x2 block structure of D and the inertia are identical to the
oracle's; L D L^T reproduces P K~ P^T to within a growth-scaled rounding
bound; solutions agree with the oracle's and the refined backward error
(SolverSettings::backward_error) reaches 1e-10 or better.
"""
import os

import numpy as np
import scipy.sparse as sp
import pytest

from nasoq_amd import lbl
from nasoq_amd import problems as P

from conftest import dense_from_factor

pytestmark = pytest.mark.gpu


def engine(n, Ap, Ai, Ax, reg=1e-9, perm=None, extra_cols=None, **kw):
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = reg
    for k, v in kw.items():
        setattr(s, k, v)
    assert s.symbolic_analysis(perm=perm, extra_cols=extra_cols) == 1
    assert s.numerical_factorization() == 1
    return s


QP_ROWS = 20 * 20 + 20  # variables and equality rows of the small QP below


def qp_super_kkt(nrhs=1):
    """The super-KKT pattern [H A' C'; A . .; C . .] of a QP of the plain
    NASOQ shape (20 x 20 Laplacian, 20 equality rows, x >= -0.5), made
    quasi-definite (-1e-2 on the 420 constraint diagonals; more constraints
    than variables would leave the zero block singular)."""
    H, q, A, b, C, d = P.qp_config(grid=20, meq=20)
    nc = A.shape[0] + C.shape[0]
    K = sp.bmat([[H, A.T, C.T], [A, None, None], [C, None, None]], format="csc")
    K = (K - 1e-2 * sp.diags(np.r_[np.zeros(H.shape[0]), np.ones(nc)])).tocsc()
    n, Ap, Ai, Ax = P.lower_csc(K)
    B = np.random.default_rng(11).uniform(-1, 1, (n, nrhs))
    return n, Ap, Ai, Ax, B


def oracle_factor(O, s, n, Ap, Ai, Ax, reg):
    return O.Factor(n, Ap, Ai, Ax, reg, s.sym_ints("Perm"), s.sym_ints("super"), s.sym_longs("pi"),
                    s.sym_ints("s"), s.sym_longs("p"))


def check_factor(O, s, n, Ap, Ai, Ax, reg):
    F = oracle_factor(O, s, n, Ap, Ai, Ax, reg)
    valL, dval, rows, permf = s.get_factor()
    # identical pivot block structure and inertia
    assert np.array_equal(dval[n:] != 0, F.d_val[n:] != 0)
    assert s.inertia() == F.inertia()
    sup, pi = s.sym_ints("super"), s.sym_longs("pi")
    px = s.sym_longs("p")[sup[:-1]]
    L, D = dense_from_factor(n, sup, pi, px, valL, rows, dval)
    K = P.lower_to_full(n, Ap, Ai, Ax).toarray()
    Kt = K.copy()
    d = np.diag(K)
    Kt[np.arange(n), np.arange(n)] = np.where(np.abs(d) < reg, reg, d)
    Pm = np.eye(n)[:, permf]
    R = Pm @ L @ D @ L.T @ Pm.T
    growth = np.abs(L) @ np.abs(D) @ np.abs(L).T
    assert np.max(np.abs(R - Kt)) <= 1e-12 * max(1.0, growth.max())
    # the oracle's factor obeys the same bound
    Lo, Do = F.dense_L(), F.dense_D()
    Po = np.eye(n)[:, F.perm]
    Ro = Po @ Lo @ Do @ Lo.T @ Po.T
    assert np.max(np.abs(Ro - Kt)) <= 1e-12 * max(1.0, (np.abs(Lo) @ np.abs(Do) @ np.abs(Lo).T).max())
    return F


@pytest.mark.parametrize("precision", [0, 1])
def test_mfma_operand_layout(precision):
    # the lane layout the tile engines assume, for the FP64 and the FP32 16x16x4 MFMA
    from nasoq_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 4))
    B = rng.standard_normal((4, 16))
    C = np.zeros((16, 16))
    assert lib.lblg_selftest_mfma(_lib.dptr(np.ascontiguousarray(A)), _lib.dptr(np.ascontiguousarray(B)),
                                  _lib.dptr(C), precision) == 0
    if precision == 0:
        assert np.allclose(C, A @ B, rtol=1e-14, atol=1e-14)
    else:
        ref = A.astype(np.float32).astype(np.float64) @ B.astype(np.float32).astype(np.float64)
        assert np.allclose(C, ref, rtol=1e-5, atol=1e-5)


def test_reference_example_lbl_main():
    # examples/lbl_main.cpp:26-57: H = 2I, q = (-4,-4) -> x = (-2,-2); q = (8,8) -> x = (4,4)
    Ap = np.array([0, 1, 2], np.int32)
    Ai = np.array([0, 1], np.int32)
    Ax = np.array([2.0, 2.0])
    s = engine(2, Ap, Ai, Ax, reg=1e-9)
    assert np.allclose(s.solve_only(np.array([-4.0, -4.0])), [-2, -2], atol=1e-14)
    assert np.allclose(s.solve_only(np.array([8.0, 8.0])), [4, 4], atol=1e-14)


# sweeps: "small" = one workgroup with the block in LDS (the default for these
# sizes), "global" = one workgroup with the block in memory (the form for
# blocks beyond the LDS), "levels" = the level-scheduled launches
SWEEPS = {"small": {"LBLG_SMALL_SOLVE_NNZ": str(1 << 30), "LBLG_SMALL_SOLVE_GLOBAL": "0"},
          "global": {"LBLG_SMALL_SOLVE_NNZ": str(1 << 30), "LBLG_SMALL_SOLVE_GLOBAL": "1"},
          "levels": {"LBLG_SMALL_SOLVE_NNZ": "0", "LBLG_SMALL_SOLVE_GLOBAL": "0"}}


def set_sweep(monkeypatch, sweep):
    for k, v in SWEEPS[sweep].items():
        monkeypatch.setenv(k, v)


@pytest.mark.parametrize("sweep", ["small", "global", "levels"])
@pytest.mark.parametrize("seed", range(6))
def test_small_kkt_parity(oracle, seed, sweep, monkeypatch):
    set_sweep(monkeypatch, sweep)
    n, Ap, Ai, Ax, B = P.kkt_small(50 + 30 * seed, 12 + 6 * seed, seed=seed, nrhs=2)
    print(f"case small_kkt seed={seed} n={n} sweep={sweep}", flush=True)
    reg = 1e-9
    s = engine(n, Ap, Ai, Ax, reg=reg)
    F = check_factor(oracle, s, n, Ap, Ai, Ax, reg)
    # unrefined solve agrees with the oracle's, refined solve reaches 1e-10
    s.req_ref_iter = 0
    x0 = s.solve_only(B[:, 0])
    xo = F.solve(B[:, 0])
    assert np.linalg.norm(x0 - xo) <= 1e-6 * max(1.0, np.linalg.norm(xo))
    s.req_ref_iter = 2
    X, bwd = s.solve_only(B, return_bwd=True)
    for c in range(B.shape[1]):
        e = oracle.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c])
        assert e <= 1e-10, e
        # device-reported backward error agrees (both are rounding-level near 1e-16)
        assert abs(e - bwd[c]) <= max(1e-3 * e, 1e-16)
        xr, _ = F.refine(B[:, c], F.solve(B[:, c]))
        assert np.linalg.norm(X[:, c] - xr) <= 1e-8 * max(1.0, np.linalg.norm(xr))


@pytest.mark.parametrize("group_w,grid,m,tall", [(2048, 9, 60, 0), (96, 9, 60, 0), (96, 12, 300, 0),
                                                  (256, 14, 500, 0), (96, 12, 300, 1), (256, 14, 500, 1)])
def test_wide_supernodes(oracle, group_w, grid, m, tall, monkeypatch):
    # 3-D Laplacian KKTs whose separators give supernodes wider than 64
    # columns (panel path); small group_w sends them to the cooperative
    # multi-workgroup LDS panel kernel.  tall=1 forces its tall launch shape
    # (512-thread workgroups with 448 far rows each and 32-column panels),
    # which supernodes beyond 256 x 192 far rows take
    monkeypatch.setenv("LBLG_GROUP_W", str(group_w))
    monkeypatch.setenv("LBLG_WIDE_TALL", str(tall))
    n, Ap, Ai, Ax, B = P.kkt_config(grid, m, 6, seed=5, nrhs=3)
    print(f"case wide n={n} group_w={group_w} grid={grid}", flush=True)
    reg = 1e-9
    s = engine(n, Ap, Ai, Ax, reg=reg)
    assert s.sym_scalar("max_sup_wid") > 64
    check_factor(oracle, s, n, Ap, Ai, Ax, reg)
    X, bwd = s.solve_only(B, return_bwd=True)
    for c in range(3):
        assert oracle.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c]) <= 1e-10


@pytest.mark.parametrize("nb,bs,order,group_w,tall", [(8, 10, None, 2048, 0), (20, 12, "nat", 2048, 0),
                                                       (6, 40, "nat", 2048, 0), (4, 200, "nat", 2048, 0),
                                                       (4, 200, "nat", 96, 0), (3, 600, "nat", 256, 0),
                                                       (4, 200, "nat", 96, 1), (3, 600, "nat", 256, 1)])
@pytest.mark.parametrize("sweep", ["small", "levels"])
def test_two_by_two_pivots(oracle, nb, bs, order, group_w, tall, sweep, monkeypatch):
    # dense zero-diagonal blocks force Bunch-Kaufman 2x2 pivots in every panel
    # path: w <= 64 (LDS kernel), 64 < w <= group_w (single-workgroup panel),
    # w > group_w (cooperative multi-workgroup panel, in both launch shapes)
    monkeypatch.setenv("LBLG_GROUP_W", str(group_w))
    monkeypatch.setenv("LBLG_WIDE_TALL", str(tall))
    set_sweep(monkeypatch, sweep)
    n, Ap, Ai, Ax, B = P.block_indefinite(nb, bs, seed=1, nrhs=2)
    perm = np.arange(n, dtype=np.int32) if order == "nat" else None
    print(f"case 2x2 n={n} nb={nb} bs={bs} order={order} group_w={group_w}", flush=True)
    reg = 1e-9
    s = engine(n, Ap, Ai, Ax, reg=reg, perm=perm)
    F = check_factor(oracle, s, n, Ap, Ai, Ax, reg)
    _, dval, _, _ = s.get_factor()
    assert np.count_nonzero(dval[n:]) > 0 and np.count_nonzero(F.d_val[n:]) > 0
    X, bwd = s.solve_only(B, return_bwd=True)
    for c in range(2):
        assert oracle.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c]) <= 1e-10


def test_refine_to_target(oracle):
    n, Ap, Ai, Ax, B = P.kkt_config(10, 150, 8, seed=9, nrhs=4)
    s = engine(n, Ap, Ai, Ax, reg=1e-9, target_bwd_err=1e-13)
    X, bwd = s.solve_only(B, return_bwd=True)
    assert np.all(bwd <= 1e-13)
    for c in range(4):
        assert oracle.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c]) <= 1e-12


def test_spd_matches_dense_solve():
    n, Ap, Ai, Ax = P.lower_csc(P.laplacian_2d(23, 17, 1e-2))
    s = engine(n, Ap, Ai, Ax, reg=1e-9)
    assert s.inertia() == (n, 0, 0)
    b = np.random.default_rng(1).uniform(-1, 1, n)
    x = s.solve_only(b)
    K = P.lower_to_full(n, Ap, Ai, Ax).toarray()
    assert np.allclose(x, np.linalg.solve(K, b), rtol=1e-12, atol=1e-12)


def test_cpp_dropin_lbl_main(tmp_path):
    # the reference's examples/lbl_main.cpp flow, compiled against the C++
    # drop-in nasoq::SolverSettings header and run as its own process
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "lbl_main")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "cpp", "lbl_main.cpp"), "-L", os.path.join(root, "nasoq_amd"),
                    "-llblg", "-Wl,-rpath," + os.path.join(root, "nasoq_amd"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "Solution: 4,4" in out.stdout


def _launches(s, name):
    import ctypes
    tot, cnt = ctypes.c_double(), ctypes.c_int64()
    s.lib.lblg_kernel_time(s.h, name.encode(), ctypes.byref(tot), ctypes.byref(cnt))
    return cnt.value


def test_kernel_timing_filter():
    # lblg_set_kernel_timing_only brackets only the named families with events;
    # lblg_set_kernel_timing(h, 0) ends it and the next full timing starts afresh
    n, Ap, Ai, Ax, B = P.kkt_config(12, 300, 6, seed=5, nrhs=2)
    print(f"case timing filter n={n}", flush=True)
    s = engine(n, Ap, Ai, Ax)
    assert s.lib.lblg_set_kernel_timing_only(s.h, b"spmv,fwd_tri") == 0
    s.target_bwd_err = 1e-10
    s.max_refine_rounds = 10
    s.numerical_factorization()
    s.solve_only(B)
    assert _launches(s, "spmv") >= 1  # the residual of every backward error
    for name in ("bk_small", "bk_panel", "bk_panel_group", "schur", "schur_big", "bwd_tri", "bwd_gemv"):
        assert _launches(s, name) == 0, name
    assert s.lib.lblg_set_kernel_timing_only(s.h, b"") != 0
    assert s.lib.lblg_set_kernel_timing_only(s.h, b",") != 0
    assert s.lib.lblg_set_kernel_timing(s.h, 0) == 0
    assert s.lib.lblg_set_kernel_timing(s.h, 1) == 0
    s.numerical_factorization()
    assert _launches(s, "spmv") == 0
    assert _launches(s, "bk_small") + _launches(s, "bk_panel") + _launches(s, "bk_panel_group") > 0
    s.close()


@pytest.mark.parametrize("chain_w,group_w,grid,m,prec", [(129, 2048, 12, 300, 0), (129, 96, 14, 500, 0),
                                                         (129, 2048, 14, 500, 1), (100000, 2048, 14, 500, 0)])
def test_chain_sweeps(oracle, chain_w, group_w, grid, m, prec, monkeypatch):
    # supernodes wider than chain_w are swept by one cooperative launch each
    # (units of 128 rows chained by per-piece flags), the others piece by piece
    monkeypatch.setenv("LBLG_CHAIN_W", str(chain_w))
    monkeypatch.setenv("LBLG_GROUP_W", str(group_w))
    set_sweep(monkeypatch, "levels")
    n, Ap, Ai, Ax, B = P.kkt_config(grid, m, 6, seed=3, nrhs=3)
    print(f"case chain n={n} chain_w={chain_w} group_w={group_w} prec={prec}", flush=True)
    reg = 1e-9
    s = engine(n, Ap, Ai, Ax, reg=reg, precision=prec)
    assert s.lib.lblg_set_kernel_timing(s.h, 1) == 0
    if prec == 0:
        F = oracle_factor(oracle, s, n, Ap, Ai, Ax, reg)
        s.req_ref_iter = 0
        X0 = s.solve_only(B)
        for c in range(B.shape[1]):
            xo = F.solve(B[:, c])
            assert np.linalg.norm(X0[:, c] - xo) <= 1e-6 * max(1.0, np.linalg.norm(xo))
    s.target_bwd_err = 1e-10
    s.max_refine_rounds = 30
    X, bwd = s.solve_only(B, return_bwd=True)
    for c in range(B.shape[1]):
        assert oracle.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c]) <= 1e-10
    chains = _launches(s, "fwd_chain")
    assert (chains > 0) == (chain_w < 1000), chains
    assert _launches(s, "bwd_chain") == chains


@pytest.mark.parametrize("ss_mode", ["auto", "0", "1"])
@pytest.mark.parametrize("prec", [0, 1])
def test_small_sweeps_match_level_sweeps(oracle, prec, ss_mode, monkeypatch):
    # the one-workgroup sweeps against the level-scheduled launches on the
    # same factor, for 1..8 right-hand sides: supernodes wider than 64
    # columns (chunked triangles), and 2x2 pivots in natural order.  The
    # unrefined solutions' backward errors for the factored matrix K~ (K with
    # the perturbed diagonal; tiny pivots make them growth-sized) must be no
    # worse than the level sweeps', and refinement must reach 1e-10 on K.
    # ss_mode: levels swept per wave (0), by the whole workgroup (1), or as
    # the cost model picks (auto).
    if ss_mode != "auto":
        monkeypatch.setenv("LBLG_SMALL_SOLVE_MODE", ss_mode)
    cases = [(P.kkt_config(10, 150, 6, seed=2, nrhs=8), None),
             (P.block_indefinite(4, 150, seed=3, nrhs=8), "nat"),
             (qp_super_kkt(nrhs=8), "extra")]
    for (n, Ap, Ai, Ax, B), order in cases:
        perm = np.arange(n, dtype=np.int32) if order == "nat" else None
        extra = None
        if order == "extra":  # the QP's inequality rows as singleton supernodes ("thin" sweeps)
            extra = np.zeros(n, dtype=np.int32)
            extra[QP_ROWS:] = 1
        reg = 1e-9 if prec == 0 else max(1e-9, np.ldexp(np.abs(Ax).max(), -12))
        K = P.lower_to_full(n, Ap, Ai, Ax).toarray()
        d = np.diag(K).copy()
        K[np.arange(n), np.arange(n)] = np.where(np.abs(d) < reg, reg, d)
        knorm = np.abs(K).sum(1).max()
        eta = {}
        for sweep in ("small", "global", "levels"):
            set_sweep(monkeypatch, sweep)
            print(f"case sweeps n={n} prec={prec} sweep={sweep}", flush=True)
            s = engine(n, Ap, Ai, Ax, reg=1e-9, perm=perm, precision=prec, extra_cols=extra)
            assert s.lib.lblg_set_kernel_timing(s.h, 1) == 0
            s.req_ref_iter = 0
            e = []
            for k in range(1, 9):
                X = s.solve_only(B[:, :k].copy())
                for c in range(k):
                    x, b = X[:, c], B[:, c]
                    e.append(np.abs(b - K @ x).max() / (knorm * np.abs(x).max() + np.abs(b).max()))
            eta[sweep] = np.array(e)
            assert (_launches(s, "solve_small") > 0) == (sweep != "levels")
            s.target_bwd_err = 1e-10
            s.max_refine_rounds = 30
            X = s.solve_only(B)
            for c in range(B.shape[1]):
                assert oracle.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c]) <= 1e-10
        print(f"eta small {eta['small'].max():.2e} global {eta['global'].max():.2e} "
              f"levels {eta['levels'].max():.2e}", flush=True)
        assert eta["small"].max() <= 4 * eta["levels"].max() + 1e-15
        assert eta["global"].max() <= 4 * eta["levels"].max() + 1e-15


def test_small_sweeps_block_beyond_lds(oracle, monkeypatch):
    # a factor whose ordered block (n doubles) does not fit the 104 KB of LDS,
    # swept by one workgroup with the block in memory (LBLG_SMALL_SOLVE_GLOBAL):
    # refinement reaches 1e-10 and the unrefined solution matches the level
    # sweeps' to the same backward error
    n, Ap, Ai, Ax, B = P.kkt_config(24, 800, 6, seed=4, nrhs=1)
    assert 8 * n > 104 * 1024
    print(f"case sweeps beyond lds n={n}", flush=True)
    eta = {}
    for sweep in ("global", "levels"):
        set_sweep(monkeypatch, sweep)
        s = engine(n, Ap, Ai, Ax)
        assert s.lib.lblg_set_kernel_timing(s.h, 1) == 0
        s.req_ref_iter = 0
        x = s.solve_only(B[:, 0])
        eta[sweep] = oracle.backward_error(n, Ap, Ai, Ax, x, B[:, 0])
        assert (_launches(s, "solve_small") > 0) == (sweep == "global")
        s.target_bwd_err = 1e-10
        s.max_refine_rounds = 30
        x = s.solve_only(B[:, 0])
        assert oracle.backward_error(n, Ap, Ai, Ax, x, B[:, 0]) <= 1e-10
        s.close()
    print(f"unrefined backward error global {eta['global']:.2e} levels {eta['levels']:.2e}", flush=True)
    assert eta["global"] <= 4 * eta["levels"] + 1e-15


def test_residual_norms_recomputed(oracle):
    # SolverSettings::compute_norms / backward_error (linear_solver_wrapper.cpp:
    # 1462-1498) of caller-given x, recomputed on the device: equal to the
    # oracle's backward error for the refined solution and for an edited one,
    # and the device-pointer entry point agrees bit for bit with the host one
    import ctypes
    import torch
    n, Ap, Ai, Ax, B = P.kkt_config(8, 40, 6, seed=5, nrhs=3)
    print(f"case residual_norms n={n}", flush=True)
    s = engine(n, Ap, Ai, Ax)
    X = s.solve_only(B)
    Xe = X.copy()
    Xe[::7, :] += 1e-3  # an edited solution (every column)
    for XX in (X, Xe):
        nm = s.residual_norms(B, XX)
        for c in range(B.shape[1]):
            want = oracle.backward_error(n, Ap, Ai, Ax, XX[:, c], B[:, c])
            if XX is X:  # rounding-level residuals: only their size is fixed
                assert nm[c, 3] <= 1e-14 and want <= 1e-14, (c, nm[c, 3], want)
            else:  # cancellation leaves the summation order visible at ~1e-12
                assert abs(nm[c, 3] - want) <= 1e-10 * want, (c, nm[c, 3], want)
            assert abs(nm[c, 1] - np.abs(XX[:, c]).sum()) <= 1e-13 * nm[c, 1]
            assert abs(nm[c, 2] - np.abs(B[:, c]).sum()) <= 1e-13 * nm[c, 2]
        dB = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()  # column-major n x nrhs
        dX = torch.from_numpy(np.ascontiguousarray(XX.T)).cuda()
        dE = torch.zeros(B.shape[1], dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        rc = s.lib.lblg_backward_error_device(s.h, B.shape[1], ctypes.c_void_p(dB.data_ptr()), n,
                                              ctypes.c_void_p(dX.data_ptr()), n, ctypes.c_void_p(dE.data_ptr()))
        assert rc == 0
        assert s.lib.lblg_synchronize(s.h) == 0
        assert np.array_equal(dE.cpu().numpy(), nm[:, 3])
    s.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_supernodes_beyond_256_workgroups_of_192_rows(oracle, prec, monkeypatch):
    # a 100-column supernode with 50,100 rows and the 50,100-column clique
    # above it: both have more far rows than 256 workgroups of 192 (the
    # 256-thread launch shape), so the cooperative panel takes its tall shape
    # with 50000 / 448 -> 112 workgroups.  The narrow supernode's panel
    # depends on its own columns only: L11 D1 L11' = T1 (dense SPD, no
    # pivoting) and L21 = B1 L11^-T D1^-1, computed here in numpy; the whole
    # factor gives the closed-form inertia (n1, m2, 0) and the refined solve
    # reaches 1e-10 (backward error recomputed by the oracle)
    import ctypes
    monkeypatch.setenv("LBLG_GROUP_W", "64")
    n1, m2 = 200, 50000
    n, Ap, Ai, Ax, B = P.tall_border(n1, m2, chains=2, seed=1, nrhs=2)
    print(f"case tall border n={n} prec={prec}", flush=True)
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.n_relax, s.z_relax = [0, 0, 0], [0.0, 0.0, 0.0]
    s.reg_diag = 1e-9
    s.precision = prec
    assert s.symbolic_analysis(perm=np.arange(n, dtype=np.int32)) == 1
    sup, pi = np.array(s.sym_ints("super")), np.array(s.sym_longs("pi"))
    w, r = np.diff(sup), np.diff(pi)
    assert list(w) == [100, 50100] and list(r) == [50100, 50100]
    assert np.all(r - 64 > 256 * 192)
    assert s.numerical_factorization() == 1
    panels, piv = ctypes.c_int64(), ctypes.c_int64()
    assert s.lib.lblg_panel_stats(s.h, ctypes.byref(panels), ctypes.byref(piv)) == 0
    assert panels.value > 50100 // (32 if prec == 0 else 64)  # tall-shape panel widths: 32 FP64, 64 FP32
    assert s.inertia() == (n1, m2, 0)
    # the narrow supernode's panel against dense arithmetic (the FP32 factor
    # is that of S K S, S the max-norm diagonal scaling the handle reports)
    c = 100
    T1 = np.full((c, c), 0.5) + np.eye(c) * (c / 2.0 + 0.5)
    K = P.lower_to_full(n, Ap, Ai, Ax)
    B1 = K[n1:, :c].toarray()
    if prec != 0:
        sc = s.scaling()
        T1 = sc[:c, None] * T1 * sc[None, :c]
        B1 = sc[n1:, None] * B1 * sc[None, :c]
    Lc = np.linalg.cholesky(T1)
    d1 = np.diag(Lc) ** 2
    L11 = Lc / np.diag(Lc)
    L21 = np.linalg.solve(L11, B1.T).T / d1
    valL, dval, rows, permf = s.get_factor()
    panel = valL[:c * (c + m2)].reshape(c, c + m2).T
    tol = 1e-12 if prec == 0 else 2e-5
    assert np.array_equal(permf[:c], np.arange(c))
    assert np.abs(dval[:c] - d1).max() <= tol * d1.max()
    assert np.abs(np.tril(panel[:c], -1) - np.tril(L11, -1)).max() <= tol
    assert np.abs(panel[c:] - L21).max() <= tol * max(1.0, np.abs(L21).max())
    s.target_bwd_err = 1e-10
    s.max_refine_rounds = 20
    X, bwd = s.solve_only(B, return_bwd=True)
    for k in range(B.shape[1]):
        assert oracle.backward_error(n, Ap, Ai, Ax, X[:, k], B[:, k]) <= 1e-10
    s.close()
