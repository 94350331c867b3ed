"""The device-schedule autotuner (lblg_autotune): it times candidate
schedules on the handle's matrix, keeps the fastest and leaves the handle
factored with it; the result must solve as every schedule does (parity with
the oracle), and an explicit schedule set by lblg_set_tuning must too."""
import numpy as np
import pytest

from nasoq_amd import lbl
from nasoq_amd import problems as P

pytestmark = pytest.mark.gpu


def test_autotune_keeps_a_correct_schedule(oracle):
    n, Ap, Ai, Ax, B = P.kkt_config(12, 300, 6, seed=8, nrhs=2)
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = 1e-9
    s.target_bwd_err = 1e-10
    s.max_refine_rounds = 30
    assert s.symbolic_analysis() == 1
    print(f"case autotune n={n} maxw={s.sym_scalar('max_sup_wid')}", flush=True)
    res = s.autotune(nrhs=2, reps=1)
    print(res, flush=True)
    assert res["schedules_tried"] >= 3 and res["factor_ms"] > 0 and res["solve_ms"] > 0
    assert s.tuning() == {k: res[k] for k in ("group_w", "chain_w", "small_solve_nnz", "schur_big_m", "wide_sub")}
    X = s.solve_only(B)  # the handle is factored with the chosen schedule
    for c in range(2):
        assert oracle.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c]) <= 1e-10
    s.close()


@pytest.mark.parametrize("knobs", [dict(group_w=96, schur_big_m=128, chain_w=129, small_solve_nnz=0, wide_sub=2),
                                   dict(group_w=96, schur_big_m=256, chain_w=512, small_solve_nnz=0, wide_sub=8),
                                   dict(group_w=4096, schur_big_m=512, chain_w=1 << 30, small_solve_nnz=1 << 22,
                                        wide_sub=4)])
def test_explicit_schedules(oracle, knobs):
    n, Ap, Ai, Ax, B = P.kkt_config(12, 300, 6, seed=9, nrhs=3)
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = 1e-9
    s.target_bwd_err = 1e-10
    s.max_refine_rounds = 30
    assert s.symbolic_analysis() == 1
    s.set_tuning(**knobs)
    assert s.tuning() == knobs
    print(f"case schedule {knobs}", flush=True)
    assert s.numerical_factorization() == 1
    X = s.solve_only(B)
    for c in range(3):
        assert oracle.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c]) <= 1e-10
    s.close()
