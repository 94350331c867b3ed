"""Batched box-constrained QPs (need a gfx950 GPU).

The batched QP configuration: independent QPs min 1/2 x'Hx + q'x with
Ax = b and |x| <= 1 (H a 5-point Laplacian + 1e-2 I of an r x c mesh),
every one solved by NASOQ's dual active-set loop (src/nasoq.cpp:478-610)
with the KKT systems factored on the GPU and updated along the tree paths
of the edited constraint rows.  qp.solve_many runs them on several host
threads, each problem with its own engine handle and HIP stream.

Checked: every problem is OPTIMAL by NASOQ's own KKT test
(check_solve_status, nasoq.cpp:712-728); the concurrent run gives the same
iteration counts, active sets and objectives as one problem at a time;
the updates refactor fewer supernodes than full factorizations would; the
FP32 KKT factor (FP64 refinement) reaches the same optima.
"""
import numpy as np
import pytest

from nasoq_amd import problems as P
from nasoq_amd import qp

pytestmark = pytest.mark.gpu

SETTINGS = {"eps_abs": 1e-6, "variant": qp.FIXED}


def test_concurrent_matches_sequential():
    probs = P.box_qp_batch(6, seed=11, lo=16, hi=28, meq_max=12)
    print("case box qps", [p[0].shape[0] for p in probs], flush=True)
    seq = qp.solve_many(probs, SETTINGS, threads=1, keep=("primal_vars", "active_set"))
    par = qp.solve_many(probs, SETTINGS, threads=3, keep=("primal_vars", "active_set"))
    for a, b in zip(seq, par):
        assert a["status"] == qp.OPTIMAL and b["status"] == qp.OPTIMAL
        assert a["num_iter"] == b["num_iter"]
        assert sorted(a["active_set"].tolist()) == sorted(b["active_set"].tolist())
        assert abs(a["objective"] - b["objective"]) <= 1e-10 * max(1.0, abs(a["objective"]))
        assert np.abs(a["primal_vars"] - b["primal_vars"]).max() <= 1e-8
        # the box: |x| <= 1 within eps
        assert np.abs(a["primal_vars"]).max() <= 1 + 1e-6
        # the SoMod updates refactor part of the tree per iteration
        assert a["refactored_supernodes"] < a["factorizations"] * a["supernodes"]
    print("iterations", [r["num_iter"] for r in seq], "refactored / full",
          [round(r["refactored_supernodes"] / (r["factorizations"] * r["supernodes"]), 3) for r in seq], flush=True)


def test_fp32_factor_same_optimum():
    # small QPs (the variant refines three GMRES(3) cycles) and larger ones
    # (one GMRES(1) cycle for an FP64 factor: the FP32 factor's solves refine
    # to a backward error target instead)
    probs = P.box_qp_batch(3, seed=12, lo=16, hi=24, meq_max=8) + P.box_qp_batch(2, seed=13, lo=60, hi=72, meq_max=8)
    print("case box qps fp32", [p[0].shape[0] for p in probs], flush=True)
    f64 = qp.solve_many(probs, SETTINGS, threads=2)
    f32 = qp.solve_many(probs, dict(SETTINGS, precision=1), threads=2)
    for a, b in zip(f64, f32):
        assert a["status"] == qp.OPTIMAL and b["status"] == qp.OPTIMAL
        assert abs(a["objective"] - b["objective"]) <= 1e-7 * max(1.0, abs(a["objective"]))
        assert np.abs(a["primal_vars"] - b["primal_vars"]).max() <= 1e-5
    print("iterations f64", [r["num_iter"] for r in f64], "f32", [r["num_iter"] for r in f32], flush=True)
