"""NASOQ QP solver on the device engine vs the reference's own expected
results: examples/nasoq_main.cpp, tests/test_1.cpp (quadprog_1.json) and
data/out_correct.csv (matt_qp1 / test05_0), plus KKT optimality conditions."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from nasoq_amd import problems as P
from nasoq_amd import qp, smp

from qp_cases import GOLDEN, nasoq_main_problem, out_correct, quadprog_1

pytestmark = pytest.mark.gpu


def test_nasoq_main_example():
    # examples/nasoq_main.cpp:47-68: x = {0.4, 1.2}, z = {1.6, 0, 0, 0}
    H, q, C, d = nasoq_main_problem()
    m = qp.Nasoq(H, q, B=C, b=d)
    m.diag_perturb = 1e-9
    m.eps_abs = 1e-3
    m.max_iter = 0
    m.variant = qp.PREDET
    assert m.solve() == qp.OPTIMAL
    assert np.allclose(m.primal_vars, [0.4, 1.2], atol=1e-12)
    assert np.allclose(m.dual_vars, [1.6, 0, 0, 0], atol=1e-12)
    m.close()


def test_quadprog_1():
    # tests/test_1.cpp:82-97: x, y, z of the fixture (Eigen isApprox)
    t = quadprog_1()
    st, x, y, z = qp.quadprog(t["H"], t["q"], t["A"], t["b"], t["C"], t["d"])
    assert st > 0
    for got, want in ((x, t["x"]), (z, t["z"])):
        assert np.linalg.norm(got - want) <= 1e-9 * max(np.linalg.norm(want), 1e-300)
    assert len(y) == len(t["y"]) == 0


def _smp_solver(name, variant, max_iter, eps):
    d = smp.load_smp(os.path.join(GOLDEN, name + ".yml"))
    m = qp.Nasoq(d["H"], d["q"], d["A"] if d["A"].shape[0] else None, d["a"], d["B"], d["b"])
    m.eps_abs = eps
    m.variant = variant
    m.max_iter = max_iter
    return m, d


@pytest.mark.parametrize("separate_rows", [1, 0])
@pytest.mark.parametrize("name", ["matt_qp1", "test05_0"])
@pytest.mark.parametrize("tool,variant,max_iter", [("NASOQ-Fixed", qp.FIXED, 0), ("NASOQ", qp.PREDET, 2)])
def test_smp_out_correct(name, tool, variant, max_iter, separate_rows):
    # separate_rows = 1: every inequality row its own supernode, as in the
    # reference; 0: rows join supernodes (same matrix, other blocking)
    row = out_correct()[(tool, name)]
    print(f"case smp {name} {tool} separate_rows={separate_rows}", flush=True)
    m, d = _smp_solver(name, variant, max_iter, float(row["eps_abs"]))
    m.separate_rows = separate_rows
    st = m.solve()
    assert st == int(row["Status"])
    assert m.num_iter == int(row["# of Iterations"])
    assert len(m.active_set) == int(row["Active-set Size"])
    # values printed with 6 significant digits in the csv
    assert np.isclose(m.objective, float(row["Obj Value"]), rtol=1e-5)
    assert np.isclose(m.primal_obj, float(row["Primal Obj"]), rtol=1e-5)
    cs = float(row["Constraint Satisfaction Inf"])
    assert m.cons_sat_norm <= max(2 * cs, 1e-12) and (cs < 1e-12 or np.isclose(m.cons_sat_norm, cs, rtol=1e-4))
    assert m.lag_res <= 1e-10 and m.non_negativity_infn == float(row["Non-negativity Inf"])
    m.close()


def _kkt_check(m, H, q, A, a, B, b, tol):
    x, y, z = m.primal_vars, m.dual_vars_eq, m.dual_vars
    Hf = (H + sp.tril(H, -1).T).toarray()
    r = Hf @ x + q + (B.T @ z if B is not None else 0) + (A.T @ y if A is not None else 0)
    assert np.abs(r).max() <= tol                                    # stationarity
    if B is not None:
        s = B @ x - b
        assert s.max() <= tol and z.min() >= -tol                     # primal / dual feasibility
        assert np.abs(z * s).max() <= tol                             # complementarity
    if A is not None:
        assert np.abs(A @ x - a).max() <= tol


@pytest.mark.parametrize("seed", range(3))
def test_random_qp_kkt_and_warm_start(seed):
    rng = np.random.default_rng(seed)
    n, me, mi = 60, 5, 80
    R = sp.random(n, n, density=0.08, random_state=rng)
    H = sp.tril(sp.csc_matrix(R @ R.T + sp.identity(n)), format="csc")
    q = rng.standard_normal(n)
    A = sp.random(me, n, density=0.2, random_state=rng, format="csc")
    a = A @ rng.uniform(-0.5, 0.5, n)
    B = sp.random(mi, n, density=0.1, random_state=rng, format="csc")
    b = B @ rng.uniform(-0.5, 0.5, n) + rng.uniform(0.0, 0.2, mi)
    m = qp.Nasoq(H, q, A, a, B, b)
    m.eps_abs = 1e-8
    m.variant = qp.FIXED
    assert m.solve() == qp.OPTIMAL
    _kkt_check(m, H, q, A, a, B, b, 1e-7)
    # the solution solves the KKT system of its active set (dense check)
    act = m.active_set
    Ba = B.toarray()[act]
    Hf = (H + sp.tril(H, -1).T).toarray()
    K = np.block([[Hf, A.toarray().T, Ba.T], [A.toarray(), np.zeros((me, me)), np.zeros((me, len(act)))],
                  [Ba, np.zeros((len(act), me)), np.zeros((len(act), len(act)))]])
    sol = np.linalg.solve(K, np.concatenate([-q, a, b[act]]))
    assert np.allclose(m.primal_vars, sol[:n], atol=1e-8)
    x, z, y = m.primal_vars.copy(), m.dual_vars.copy(), m.dual_vars_eq.copy()
    # warm start from the solution: converged without further iterations
    w = qp.Nasoq(H, q, A, a, B, b)
    w.eps_abs = 1e-8
    assert w.solve(x, z, act, y) == 1
    assert w.num_iter == 0 and np.allclose(w.primal_vars, x, atol=1e-12)
    m.close()
    w.close()


def _build_cpp(name, tmp_path):
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "cpp", name + ".cpp"), "-L", os.path.join(root, "nasoq_amd"),
                    "-llblg", "-Wl,-rpath," + os.path.join(root, "nasoq_amd"), "-o", exe], check=True)
    return exe


def test_cpp_dropin_nasoq_main(tmp_path):
    import subprocess
    out = subprocess.run([_build_cpp("nasoq_main", tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr


def test_cpp_dropin_readable_state(tmp_path):
    # the reference's readable SolverSettings state (L, valL, d_val, A_ord,
    # AT_ord, perm_piv, level sets) is consistent: L D L' rebuilt from it equals
    # the perturbed A_ord, and backward_error() recomputes the residual of an
    # edited x (linear_solver_wrapper.cpp:1432-1498)
    import subprocess
    out = subprocess.run([_build_cpp("lbl_state_main", tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), (out.returncode, out.stdout + out.stderr)


@pytest.mark.parametrize("zero", [True, False])
def test_cpp_dropin_somod(tmp_path, zero):
    # examples/lbl_somod_main.cpp: with zero right-hand sides for the added
    # rows the solutions are the example's listed vectors {-2,-2}, {0,-2,-4},
    # {0,0,-4,-4}, {-2,0,0,-4}; with d = (2, 1) they solve the same bordered
    # systems with those right-hand sides
    import subprocess
    exe = _build_cpp("lbl_somod_main", tmp_path)
    out = subprocess.run([exe] + (["zero"] if zero else []), capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    got = [np.array([float(v) for v in ln.split()]) for ln in out.stdout.strip().splitlines()]
    if zero:
        want = [[-2, -2], [0, -2, -4, 0], [0, 0, -4, -4], [-2, 0, 0, -4]]
    else:
        want = [[-2, -2], [2, -2, -8, 0], [2, 1, -8, -6], [-2, 1, 0, -6]]
    assert len(got) == 4
    for g, w in zip(got, want):
        assert np.allclose(g, w, atol=1e-12), (g, w)


def test_plain_nasoq_configuration():
    # the reference's CPU-runnable configuration (40 x 40 Laplacian + 1e-2 I,
    # 100 equality rows of 5 N(0,1), x >= -0.5), eps 1e-8: NASOQ on the GPU
    # must return a point that satisfies the KKT conditions, and that point
    # must be the solution of the KKT system of its own active set
    H, q, A, b, C, d = P.qp_config()
    n, me = H.shape[0], A.shape[0]
    print(f"case qp_config n={n} me={me} mi={C.shape[0]}", flush=True)
    m = qp.Nasoq(sp.tril(H, format="csc"), q, A, b, C, d)
    m.eps_abs = 1e-8
    m.variant = qp.FIXED
    assert m.solve() == qp.OPTIMAL
    _kkt_check(m, sp.tril(H, format="csc"), q, A, b, C, d, 1e-6)
    act = np.array(m.active_set, dtype=int)
    assert len(act) > 0
    # active bounds x_i = -0.5: eliminate them and solve the reduced equality-constrained KKT densely
    free = np.setdiff1d(np.arange(n), act)
    Hd = H.toarray()
    Ad = A.toarray()
    xa = np.full(len(act), -0.5)
    Kr = np.block([[Hd[np.ix_(free, free)], Ad[:, free].T], [Ad[:, free], np.zeros((me, me))]])
    rhs = np.concatenate([-q[free] - Hd[np.ix_(free, act)] @ xa, b - Ad[:, act] @ xa])
    sol = np.linalg.solve(Kr, rhs)
    x = np.zeros(n)
    x[free] = sol[:len(free)]
    x[act] = xa
    assert np.linalg.norm(m.primal_vars - x) <= 1e-7 * max(1.0, np.linalg.norm(x))
    m.close()


def test_qp_autotuned_schedule_keeps_the_solution():
    # a KKT device schedule picked by the autotuner gives the same optimum
    H, q, A, b, C, d = P.qp_config(grid=24, meq=30, seed=3)
    Hl = sp.tril(H, format="csc")
    sols = []
    for tune in (False, True):
        m = qp.Nasoq(Hl, q, A, b, C, d)
        m.eps_abs = 1e-8
        m.variant = qp.FIXED
        if tune:
            t = m.autotune()
            print(f"case qp autotune {t}", flush=True)
            assert t["schedules_tried"] >= 2
        assert m.solve() == qp.OPTIMAL
        _kkt_check(m, Hl, q, A, b, C, d, 1e-6)
        sols.append(m.primal_vars.copy())
        m.close()
    assert np.linalg.norm(sols[0] - sols[1]) <= 1e-7 * max(1.0, np.linalg.norm(sols[0]))
