"""The C-ABI library loads and exports every entry point include/lblg.h declares."""
import os
import re
import subprocess

from nasoq_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "lblg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lblg_[a-z_0-9]+)\s*\(", txt)))


def test_library_loads_and_exports_all_declared():
    lib = _lib.load()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (lblg_\w+)", out))
    decl = declared_symbols()
    assert len(decl) >= 20
    missing = [s for s in decl if s not in exported]
    assert not missing, missing
    for s in decl:
        getattr(lib, s)


def test_python_binding_covers_declarations():
    names = {n for n, _, _ in _lib.SIGNATURES}
    assert set(declared_symbols()) <= names


def test_settings_struct_matches_defaults():
    import ctypes
    lib = _lib.load()
    st = _lib.Settings()
    lib.lblg_default_settings(ctypes.byref(st))
    # SolverSettings::default_setting (linear_solver_wrapper.cpp:205-244)
    assert st.req_ref_iter == 2 and st.max_inner_iter == 2
    assert st.tol_abs == 1e-15 and st.tol_rel == 1e-15 and st.reg_diag == 1e-8
    assert list(st.n_relax) == [4, 16, 48] and list(st.z_relax) == [0.8, 0.1, 0.05]
    assert st.level_param == -3 and st.final_seq_node == 50 and st.regularization == 1


def test_invalid_inputs_are_rejected():
    import ctypes
    import numpy as np
    lib = _lib.load()
    h = ctypes.c_void_p()
    Ap = np.array([0, 1, 2], np.int32)
    Ai = np.array([1, 1], np.int32)  # diagonal missing in column 0
    Ax = np.ones(2)
    assert lib.lblg_create(ctypes.byref(h), 2, _lib.iptr(Ap), _lib.iptr(Ai), _lib.dptr(Ax)) == -1
    Ai = np.array([0, 1], np.int32)
    assert lib.lblg_create(ctypes.byref(h), 2, _lib.iptr(Ap), _lib.iptr(Ai), _lib.dptr(Ax)) == 0
    # numeric calls before analysis are state errors, never crashes
    assert lib.lblg_factor(h) == -2
    B = np.ones(2)
    X = np.zeros(2)
    assert lib.lblg_solve(h, 1, _lib.dptr(B), 2, _lib.dptr(X), 2, None) in (-1, -2)
    lib.lblg_destroy(h)


def test_tuning_knobs_on_the_host():
    # the schedule knobs resolve and validate without a device
    from nasoq_amd import lbl
    from nasoq_amd import problems as P
    n, Ap, Ai, Ax, _ = P.kkt_small(40, 10, seed=1)
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    t = s.tuning()
    assert t["group_w"] >= 64 and t["chain_w"] >= 1 and t["small_solve_nnz"] >= 0 and t["schur_big_m"] >= 1
    s.set_tuning(group_w=1024, schur_big_m=128)
    assert s.tuning()["group_w"] == 1024 and s.tuning()["schur_big_m"] == 128
    import pytest as _pt
    with _pt.raises(RuntimeError):
        s.set_tuning(group_w=8)
    s.close()
