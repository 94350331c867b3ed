"""Captured launch sequences (HIP graphs) against the eager launches.

The factor, the solve with fixed GMRES cycles and the cycles of
target-driven refinement are captured on their second call and replayed.
Each case factors and solves a sequence of matrices with the same pattern
(changing values, as the active-set loop does), with eager host and device
work between the replays, once with capture on and once with LBLG_GRAPH=0,
and checks every result against the oracle's backward error and against the
eager run."""
import numpy as np
import pytest
import torch

from nasoq_amd import lbl
from nasoq_amd import problems as P

pytestmark = pytest.mark.gpu


def run_sequence(n, Ap, Ai, Ax, B, target, sweep_nnz, monkeypatch, graph):
    monkeypatch.setenv("LBLG_GRAPH", "1" if graph else "0")
    monkeypatch.setenv("LBLG_SMALL_SOLVE_NNZ", sweep_nnz)
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = 1e-9
    if target:
        s.target_bwd_err = 1e-12
        s.max_refine_rounds = 20
    else:
        s.req_ref_iter = 2
        s.max_inner_iter = 3
    assert s.symbolic_analysis() == 1
    diag = np.asarray(Ap[:-1])
    out = []
    for k in range(6):
        Axk = Ax.copy()
        Axk[diag] *= 1.0 + 0.1 * k  # same pattern, new values
        s.set_values(Axk)
        assert s.numerical_factorization() == 1
        X = s.solve_only(B)
        # eager work between replays: a torch kernel and a synchronisation
        t = torch.ones(1000, device="cuda").sum().item()
        assert t == 1000.0
        out.append((Axk, X))
    stats = s.graph_stats()
    s.close()
    return out, stats


@pytest.mark.parametrize("target", [False, True])
@pytest.mark.parametrize("sweep", ["small", "levels"])
def test_graph_replays_match_eager(oracle, target, sweep, monkeypatch):
    n, Ap, Ai, Ax, B = P.kkt_config(9, 60, 6, seed=4, nrhs=2)
    nnz = str(1 << 30) if sweep == "small" else "0"
    print(f"case graph n={n} target={target} sweep={sweep}", flush=True)
    eg, st_e = run_sequence(n, Ap, Ai, Ax, B, target, nnz, monkeypatch, graph=False)
    gr, st_g = run_sequence(n, Ap, Ai, Ax, B, target, nnz, monkeypatch, graph=True)
    assert st_e == (0, 0)
    caps, launches = st_g
    assert caps >= 2 and launches >= 8, st_g  # factor and solve captured, replays from the third call
    for (Axk, Xe), (_, Xg) in zip(eg, gr):
        for c in range(B.shape[1]):
            e = oracle.backward_error(n, Ap, Ai, Axk, Xg[:, c], B[:, c])
            assert e <= 1e-10, e
            assert np.linalg.norm(Xg[:, c] - Xe[:, c]) <= 1e-8 * max(1.0, np.linalg.norm(Xe[:, c]))
