"""Python mirror of the reference's LBL operator interface, nasoq::SolverSettings
(include/nasoq/QP/linear_solver_wrapper.h:53-306), over the engine's C-ABI.

    s = SolverSettings(n, Ap, Ai, Ax)      # SolverSettings(CSC* A, double* rhs)
    s.reg_diag = 1e-9                      # public fields as in the reference
    s.symbolic_analysis()                  # -> 1 on success
    s.numerical_factorization()            # -> 1 on success (device factor)
    x = s.solve_only(rhs)                  # solve + GMRES refinement
    s.backward_error(x, rhs)               # reference residual formula

Every numeric call runs on the GPU through nasoq_amd/liblblg.so; there is no
host fallback.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr, lptr

SYM_SCALARS = ["n", "nsuper", "n_level", "n_par", "max_sup_wid", "max_col", "xsize", "ssize"]
SYM_INTS = ["Perm", "IPerm", "Parent", "ColCount", "super", "col2Sup", "sParent", "s",
            "level_ptr", "par_ptr", "par_set", "extra_cols"]
SYM_LONGS = ["pi", "p", "i_ptr"]

ORDER_ND, ORDER_GIVEN, ORDER_NATURAL = 0, 1, 2
F64, F32, BF16 = 0, 1, 2


class SolverSettings:
    def __init__(self, n, Ap, Ai, Ax):
        self.lib = _lib.load()
        self.n = int(n)
        self.Ap = np.ascontiguousarray(Ap, dtype=np.int32)
        self.Ai = np.ascontiguousarray(Ai, dtype=np.int32)
        self.Ax = np.ascontiguousarray(Ax, dtype=np.float64)
        self.h = ctypes.c_void_p()
        check(self.lib.lblg_create(ctypes.byref(self.h), self.n, iptr(self.Ap), iptr(self.Ai),
                                   dptr(self.Ax)), "lblg_create")
        st = _lib.Settings()
        self.lib.lblg_default_settings(ctypes.byref(st))
        # public fields of SolverSettings (default_setting values)
        self.ldl_variant = st.ldl_variant
        self.req_ref_iter = st.req_ref_iter
        self.max_inner_iter = st.max_inner_iter
        self.tol_abs = st.tol_abs
        self.tol_rel = st.tol_rel
        self.reg_diag = st.reg_diag
        self.regularization = st.regularization
        self.cost_param = st.cost_param
        self.level_param = st.level_param
        self.final_seq_node = st.final_seq_node
        self.n_relax = list(st.n_relax)
        self.z_relax = list(st.z_relax)
        self.ordering = st.ordering
        self.precision = st.precision
        self.target_bwd_err = st.target_bwd_err
        self.max_refine_rounds = st.max_refine_rounds
        self.perm_in = None
        self.extra_cols = None
        self.num_ref_iter = 0

    def _settings(self):
        st = _lib.Settings()
        self.lib.lblg_default_settings(ctypes.byref(st))
        for f, _ in _lib.Settings._fields_:
            v = getattr(self, f)
            if f in ("n_relax", "z_relax"):
                arr = getattr(st, f)
                for k in range(3):
                    arr[k] = v[k]
            else:
                setattr(st, f, v)
        return st

    def close(self):
        if self.h:
            self.lib.lblg_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- symbolic ---------------------------------------------------------
    def symbolic_analysis(self, perm=None, extra_cols=None):
        st = self._settings()
        if perm is not None:
            self.ordering = ORDER_GIVEN
            st.ordering = ORDER_GIVEN
            self.perm_in = np.ascontiguousarray(perm, dtype=np.int32)
        ex = None if extra_cols is None else np.ascontiguousarray(extra_cols, dtype=np.int32)
        rc = self.lib.lblg_analyze(self.h, ctypes.byref(st), iptr(self.perm_in), iptr(ex))
        return 1 if rc == 0 else 0

    def sym_scalar(self, name):
        return int(self.lib.lblg_symbolic_scalar(self.h, SYM_SCALARS.index(name)))

    def sym_ints(self, name):
        w = SYM_INTS.index(name)
        ln = self.lib.lblg_symbolic_ints(self.h, w, None)
        if ln < 0:
            raise RuntimeError("symbolic analysis not available")
        out = np.zeros(max(ln, 1), dtype=np.int32)
        self.lib.lblg_symbolic_ints(self.h, w, iptr(out))
        return out[:ln]

    def sym_longs(self, name):
        w = SYM_LONGS.index(name)
        ln = self.lib.lblg_symbolic_longs(self.h, w, None)
        out = np.zeros(max(ln, 1), dtype=np.int64)
        self.lib.lblg_symbolic_longs(self.h, w, lptr(out))
        return out[:ln]

    def flops(self):
        fl, lnz = ctypes.c_double(), ctypes.c_double()
        check(self.lib.lblg_symbolic_flops(self.h, ctypes.byref(fl), ctypes.byref(lnz)), "flops")
        return fl.value, lnz.value


class BatchSolverSettings(SolverSettings):
    """Many independent symmetric systems (lower CSC, diagonal first) in one
    handle: a block-diagonal matrix whose blocks are ordered by their own
    nested dissection.  Factor and solve as SolverSettings; right-hand sides
    and solutions are the blocks' vectors concatenated in order."""

    def __init__(self, blocks):
        self.lib = _lib.load()
        self.blocks = [(int(n), np.ascontiguousarray(Ap, dtype=np.int32), np.ascontiguousarray(Ai, dtype=np.int32),
                        np.ascontiguousarray(Ax, dtype=np.float64)) for n, Ap, Ai, Ax in blocks]
        cnt = len(self.blocks)
        ns = np.array([b[0] for b in self.blocks], dtype=np.int32)
        P = (ctypes.c_void_p * cnt)(*[b[1].ctypes.data for b in self.blocks])
        I = (ctypes.c_void_p * cnt)(*[b[2].ctypes.data for b in self.blocks])
        X = (ctypes.c_void_p * cnt)(*[b[3].ctypes.data for b in self.blocks])
        self.h = ctypes.c_void_p()
        check(self.lib.lblg_batch_create(ctypes.byref(self.h), cnt, iptr(ns), ctypes.cast(P, ctypes.c_void_p),
                                         ctypes.cast(I, ctypes.c_void_p), ctypes.cast(X, ctypes.c_void_p)),
              "lblg_batch_create")
        c = ctypes.c_int()
        self.offsets = np.zeros(cnt + 1, dtype=np.int64)
        check(self.lib.lblg_batch_info(self.h, ctypes.byref(c), lptr(self.offsets)), "lblg_batch_info")
        self.n = int(self.offsets[-1])
        st = _lib.Settings()
        self.lib.lblg_default_settings(ctypes.byref(st))
        for f, _ in _lib.Settings._fields_:
            setattr(self, f, list(getattr(st, f)) if f in ("n_relax", "z_relax") else getattr(st, f))
        self.perm_in = None
        self.extra_cols = None
        self.num_ref_iter = 0

    def symbolic_analysis(self, perm=None, extra_cols=None):
        assert perm is None and extra_cols is None, "batches are ordered block by block"
        st = self._settings()
        ok = self.lib.lblg_batch_analyze(self.h, ctypes.byref(st)) == 0
        if ok:
            self.ordering = ORDER_GIVEN  # the handle records the composed block ordering
        return 1 if ok else 0

    def backward_errors(self, B, X):
        """(count, nrhs) backward errors of every block and column."""
        B = np.asfortranarray(np.asarray(B, dtype=np.float64).reshape(self.n, -1))
        X = np.asfortranarray(np.asarray(X, dtype=np.float64).reshape(self.n, -1))
        cnt = len(self.blocks)
        err = np.zeros(cnt * B.shape[1])
        check(self.lib.lblg_batch_backward_errors(self.h, B.shape[1], dptr(B), self.n, dptr(X), self.n, dptr(err)),
              "lblg_batch_backward_errors")
        return err.reshape(B.shape[1], cnt).T


def nd_order(n, Ap, Ai):
    """The engine's nested-dissection ordering (perm[new] = old)."""
    lib = _lib.load()
    perm = np.zeros(n, dtype=np.int32)
    check(lib.lblg_order_nd(n, iptr(np.ascontiguousarray(Ap, dtype=np.int32)),
                            iptr(np.ascontiguousarray(Ai, dtype=np.int32)), iptr(perm)), "order_nd")
    return perm


def _numeric_methods():
    def numerical_factorization(self):
        """SolverSettings::numerical_factorization -> 1 on success (raises on device errors)."""
        # settings may have changed since analysis (reg_diag, refinement): push them
        check(self._push_settings(self._settings()), "lblg_set_settings")
        check(self.lib.lblg_factor(self.h), "lblg_factor")
        return 1

    def _push_settings(self, st):
        return self.lib.lblg_set_settings(self.h, ctypes.byref(st))

    def refactor_resident(self):
        check(self.lib.lblg_refactor_resident(self.h), "lblg_refactor_resident")
        return 1

    def solve_only(self, rhs, return_bwd=False):
        """Solve K x = rhs (rhs: n or n x nrhs) with the configured refinement."""
        st = self._settings()
        check(self._push_settings(st), "lblg_set_settings")
        b = np.asarray(rhs, dtype=np.float64)
        one = b.ndim == 1
        B = np.asfortranarray(b.reshape(self.n, -1))
        nrhs = B.shape[1]
        X = np.zeros((self.n, nrhs), dtype=np.float64, order="F")
        bwd = np.zeros(nrhs, dtype=np.float64)
        check(self.lib.lblg_solve(self.h, nrhs, dptr(B), self.n, dptr(X), self.n, dptr(bwd)), "lblg_solve")
        self.num_ref_iter = int(self.lib.lblg_num_ref_iter(self.h))
        x = X[:, 0].copy() if one else X
        return (x, bwd) if return_bwd else x

    def set_values(self, Ax):
        """New values for the same pattern (the next numerical_factorization uses them)."""
        self.Ax = np.ascontiguousarray(Ax, dtype=np.float64).copy()
        check(self.lib.lblg_set_values(self.h, dptr(self.Ax)), "lblg_set_values")

    def update_values(self, idx, vals):
        """Change entries Ax[idx] = vals in place; the next numerical_factorization
        refactors only the supernodes on their tree paths to the root (the
        SoMod update) when the handle holds an FP64 factor of the old values."""
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        assert idx.shape == vals.shape
        check(self.lib.lblg_update_values(self.h, len(idx), idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                          dptr(vals)), "lblg_update_values")
        self.Ax = np.array(self.Ax, dtype=np.float64)
        self.Ax[idx] = vals

    def refactored_supernodes(self):
        """Supernodes the last numerical_factorization refactored."""
        return int(self.lib.lblg_refactored_supernodes(self.h))

    def tuning(self):
        """The handle's device schedule knobs (dict)."""
        t = _lib.Tuning()
        check(self.lib.lblg_get_tuning(self.h, ctypes.byref(t)), "lblg_get_tuning")
        return {f: getattr(t, f) for f in ("group_w", "chain_w", "small_solve_nnz", "schur_big_m", "wide_sub")}

    def set_tuning(self, **knobs):
        """Change schedule knobs; the next numerical_factorization rebuilds the device schedule."""
        t = _lib.Tuning()
        check(self.lib.lblg_get_tuning(self.h, ctypes.byref(t)), "lblg_get_tuning")
        for k, v in knobs.items():
            setattr(t, k, int(v))
        check(self.lib.lblg_set_tuning(self.h, ctypes.byref(t)), "lblg_set_tuning")

    def autotune(self, nrhs=1, reps=2):
        """Time candidate device schedules on the current values and keep the
        fastest (the handle is left factored with it); returns the knobs and
        the chosen schedule's factor / solve times."""
        check(self._push_settings(self._settings()), "lblg_set_settings")
        t = _lib.Tuning()
        check(self.lib.lblg_autotune(self.h, nrhs, reps, ctypes.byref(t)), "lblg_autotune")
        return {f: getattr(t, f) for f, _ in _lib.Tuning._fields_}

    def graph_stats(self):
        """(captures, graph launches) of the handle's captured launch sequences."""
        c, l = ctypes.c_int64(), ctypes.c_int64()
        check(self.lib.lblg_graph_stats(self.h, ctypes.byref(c), ctypes.byref(l)), "lblg_graph_stats")
        return c.value, l.value

    def inertia(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(self.lib.lblg_inertia(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "inertia")
        return a.value, b.value, c.value

    def get_factor(self):
        """(valL, d_val, rows_pivoted, perm_final) copied from the device."""
        xs = self.sym_scalar("xsize")
        ss = int(self.sym_longs("pi")[-1])
        valL = np.zeros(int(self.sym_longs("p")[-1]), dtype=np.float64)
        d_val = np.zeros(2 * self.n, dtype=np.float64)
        rows = np.zeros(max(ss, 1), dtype=np.int32)
        perm = np.zeros(self.n, dtype=np.int32)
        check(self.lib.lblg_get_factor(self.h, dptr(valL), dptr(d_val), iptr(rows), iptr(perm)), "get_factor")
        del xs
        return valL, d_val, rows[:ss], perm

    def residual_norms(self, rhs, x):
        """(res_l1, x_l1, rhs_l1, bwd_err) per column of the caller's x for rhs,
        recomputed on the device with the unperturbed K0 of the last
        factorization (SolverSettings::compute_norms / backward_error,
        linear_solver_wrapper.cpp:1462-1498); an (nrhs, 4) array."""
        B = np.asfortranarray(np.asarray(rhs, dtype=np.float64).reshape(self.n, -1))
        X = np.asfortranarray(np.asarray(x, dtype=np.float64).reshape(self.n, -1))
        nrhs = B.shape[1]
        out = np.zeros((nrhs, 4), dtype=np.float64)
        check(self.lib.lblg_residual_norms(self.h, nrhs, dptr(B), self.n, dptr(X), self.n, dptr(out)),
              "lblg_residual_norms")
        return out

    def backward_error(self, rhs, x):
        """Backward error of x for rhs (one column: a float; several: an array)."""
        e = self.residual_norms(rhs, x)[:, 3]
        return float(e[0]) if np.asarray(rhs).ndim == 1 else e

    def scaling(self):
        """The diagonal scaling s of the factor (S K~ S is factored; ones for the FP64 factor)."""
        out = np.zeros(self.n, dtype=np.float64)
        check(self.lib.lblg_get_scaling(self.h, dptr(out)), "lblg_get_scaling")
        return out

    def factor_info(self):
        z, nrm, sec = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        check(self.lib.lblg_factor_info(self.h, ctypes.byref(z), ctypes.byref(nrm), ctypes.byref(sec)), "info")
        return {"zero_pivot": z.value, "k0_norm_inf": nrm.value, "factor_seconds": sec.value}

    for f in (numerical_factorization, _push_settings, refactor_resident, solve_only, set_values, update_values,
              refactored_supernodes, graph_stats, tuning, set_tuning, autotune, inertia, get_factor, residual_norms,
              backward_error, scaling, factor_info):
        setattr(SolverSettings, f.__name__, f)


_numeric_methods()


def device_count():
    lib = _lib.load()
    c = ctypes.c_int()
    lib.lblg_device_count(ctypes.byref(c))
    return c.value
