"""Python mirror of the reference's QP interface, nasoq::Nasoq
(include/nasoq/nasoq.h:61-493), over the engine's C-ABI (lblg_qp_*).

    qm = Nasoq(H, q, A=None, a=None, B=B, b=b)   # H lower CSC (scipy), A/B: m x n
    qm.eps_abs = 1e-3; qm.variant = FIXED          # public settings as in the reference
    status = qm.solve()                            # OPTIMAL == 1
    qm.primal_vars, qm.dual_vars, qm.dual_vars_eq, qm.num_iter, qm.objective, ...

Every KKT factorization and solve runs on the GPU (nasoq_amd/liblblg.so).
"""
import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import check, dptr, iptr

FIXED, AUTO, TUNED, PREDET = 0, 1, 2, 3
INFEASIBLE, OPTIMAL, INACCURATE, NOT_CONVERGED = 0, 1, 2, 3
ORDER_ND, ORDER_NATURAL = 0, 2


def _csc(M, nrow, ncol):
    M = sp.csc_matrix(M, shape=(nrow, ncol))
    M.sort_indices()
    return (np.ascontiguousarray(M.indptr, dtype=np.int32), np.ascontiguousarray(M.indices, dtype=np.int32),
            np.ascontiguousarray(M.data, dtype=np.float64))


class Nasoq:
    def __init__(self, H, q, A=None, a=None, B=None, b=None):
        self.lib = _lib.load()
        H = sp.tril(sp.csc_matrix(H), format="csc")
        n = H.shape[0]
        self.n = n
        self.me = 0 if A is None else A.shape[0]
        self.mi = 0 if B is None else B.shape[0]
        self._H = _csc(H, n, n)
        self._q = np.ascontiguousarray(q, dtype=np.float64)
        self._A = _csc(A, self.me, n) if self.me else (None, None, None)
        self._a = np.ascontiguousarray(a, dtype=np.float64) if self.me else None
        self._B = _csc(B, self.mi, n) if self.mi else (None, None, None)
        self._b = np.ascontiguousarray(b, dtype=np.float64) if self.mi else None
        self.h = ctypes.c_void_p()
        check(self.lib.lblg_qp_create(ctypes.byref(self.h), n, iptr(self._H[0]), iptr(self._H[1]), dptr(self._H[2]),
                                      dptr(self._q), self.me, iptr(self._A[0]), iptr(self._A[1]), dptr(self._A[2]),
                                      dptr(self._a), self.mi, iptr(self._B[0]), iptr(self._B[1]), dptr(self._B[2]),
                                      dptr(self._b)), "lblg_qp_create")
        st = _lib.QPSettings()
        self.lib.lblg_qp_default_settings(ctypes.byref(st))
        for f, _ in _lib.QPSettings._fields_:
            setattr(self, f, getattr(st, f))
        self.primal_vars = np.zeros(n)
        self.dual_vars = np.zeros(self.mi)
        self.dual_vars_eq = np.zeros(self.me)
        self.active_set = []

    def close(self):
        if self.h:
            self.lib.lblg_qp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _push(self):
        st = _lib.QPSettings()
        for f, _ in _lib.QPSettings._fields_:
            setattr(st, f, getattr(self, f))
        check(self.lib.lblg_qp_set_settings(self.h, ctypes.byref(st)), "lblg_qp_set_settings")

    def _pull(self):
        info = _lib.QPInfo()
        check(self.lib.lblg_qp_get_info(self.h, ctypes.byref(info)), "lblg_qp_get_info")
        for f, _ in _lib.QPInfo._fields_:
            setattr(self, f, getattr(info, f))
        act = np.zeros(max(info.n_active, 1), dtype=np.int32)
        check(self.lib.lblg_qp_get(self.h, dptr(self.primal_vars), dptr(self.dual_vars), dptr(self.dual_vars_eq),
                                   iptr(act)), "lblg_qp_get")
        self.active_set = [int(c) for c in act[:info.n_active]]
        self.ret_val = info.status

    def symbolic_QP(self):
        self._push()
        check(self.lib.lblg_qp_analyze(self.h), "lblg_qp_analyze")
        self._pull()
        return 0

    def solve(self, primal_in=None, dual_in=None, active_set=None, dual_eq_in=None):
        """Nasoq::solve(); with primal_in/dual_in/active_set the warm-started
        solve(primal, dual, active_set), which returns 1 when converged."""
        self._push()
        if primal_in is None:
            st = ctypes.c_int()
            check(self.lib.lblg_qp_solve(self.h, ctypes.byref(st)), "lblg_qp_solve")
            self._pull()
            return st.value
        x0 = np.ascontiguousarray(primal_in, dtype=np.float64)
        z0 = np.ascontiguousarray(dual_in if dual_in is not None else np.zeros(self.mi), dtype=np.float64)
        act = np.ascontiguousarray(active_set if active_set is not None else [], dtype=np.int32)
        y0 = None if dual_eq_in is None else np.ascontiguousarray(dual_eq_in, dtype=np.float64)
        conv = ctypes.c_int()
        check(self.lib.lblg_qp_solve_warm(self.h, dptr(x0), dptr(z0), iptr(act) if len(act) else None, len(act),
                                          dptr(y0), ctypes.byref(conv)), "lblg_qp_solve_warm")
        self._pull()
        return conv.value

    def solve_predicted(self, active_set):
        """Warm start from a predicted active set (kept where its multipliers
        are non-negative); returns (status, KKT solves spent on the prediction)."""
        self._push()
        act = np.ascontiguousarray(active_set, dtype=np.int32)
        st, pr = ctypes.c_int(), ctypes.c_int()
        check(self.lib.lblg_qp_solve_predicted(self.h, iptr(act) if len(act) else None, len(act), ctypes.byref(st),
                                               ctypes.byref(pr)), "lblg_qp_solve_predicted")
        self._pull()
        return st.value, pr.value

    def autotune(self, reps=1):
        """Pick the device schedule of the QP's KKT handle by timing candidates
        (lblg_qp_autotune); kept for the solves that follow."""
        self._push()
        t = _lib.Tuning()
        check(self.lib.lblg_qp_autotune(self.h, reps, ctypes.byref(t)), "lblg_qp_autotune")
        return {f: getattr(t, f) for f, _ in _lib.Tuning._fields_}

    def set_linear(self, q=None, a=None, b=None):
        """New linear terms for the next solve (None keeps a vector); the
        super-KKT pattern, its analysis and the device buffers are kept."""
        if q is not None:
            self._q = np.ascontiguousarray(q, dtype=np.float64).copy()
            assert self._q.shape == (self.n,)
        if a is not None and self.me:
            self._a = np.ascontiguousarray(a, dtype=np.float64).copy()
            assert self._a.shape == (self.me,)
        if b is not None and self.mi:
            self._b = np.ascontiguousarray(b, dtype=np.float64).copy()
            assert self._b.shape == (self.mi,)
        check(self.lib.lblg_qp_set_linear(self.h, dptr(self._q) if q is not None else None,
                                          dptr(self._a) if a is not None and self.me else None,
                                          dptr(self._b) if b is not None and self.mi else None),
              "lblg_qp_set_linear")

    def super_kkt(self):
        """(N, p, i, x) of the current super-KKT (lower CSC)."""
        nnz = self.lib.lblg_qp_super_kkt(self.h, None, None, None)
        if nnz < 0:
            raise RuntimeError("super-KKT not built (call symbolic_QP or solve first)")
        N = self.n + self.me + self.mi
        p = np.zeros(N + 1, np.int32)
        i = np.zeros(max(nnz, 1), np.int32)
        x = np.zeros(max(nnz, 1), np.float64)
        self.lib.lblg_qp_super_kkt(self.h, iptr(p), iptr(i), dptr(x))
        return N, p, i[:nnz], x[:nnz]


def solve_many(problems, settings=None, threads=8, keep=("primal_vars",), log=None):
    """Solve independent QPs (tuples (H, q, A, b, C, d), A / C may be None)
    concurrently: `threads` host threads each run NASOQ's active-set loop on
    one problem at a time, every problem with its own engine handle and HIP
    stream, so the KKT factorizations and solves of different problems
    overlap on the GPU (ctypes releases the interpreter lock during the
    library calls).  The threads take the problems largest first (by
    variables plus constraints), so that a large problem does not start
    last and run alone at the end.  settings: dict of public Nasoq fields;
    log(message) is called after every problem.  Returns one dict per
    problem, in the order given: status, num_iter, objective, the KKT
    checks, the factorization counters and the fields named in `keep`."""
    from concurrent.futures import ThreadPoolExecutor
    done = [0]

    def one(prob):
        H, q, A, b, C, d = prob
        m = Nasoq(sp.tril(H, format="csc"), q, A, b, C, d)
        try:
            for k, v in (settings or {}).items():
                setattr(m, k, v)
            st = m.solve()
            out = {"status": st, "num_iter": m.num_iter, "objective": m.objective, "lag_res": m.lag_res,
                   "cons_sat_norm": m.cons_sat_norm, "n_active": len(m.active_set),
                   "factorizations": int(m.factorizations), "kkt_solves": int(m.kkt_solves),
                   "refactored_supernodes": int(m.refactored_supernodes),
                   "supernodes": int(m.supernodes), "factor_seconds": m.factor_seconds,
                   "solve_seconds": m.solve_seconds}
            for k in keep:
                out[k] = np.array(getattr(m, k), copy=True)
            done[0] += 1
            if log:
                log(f"solved {done[0]}/{len(problems)}: n={m.n} status {st} {m.num_iter} iterations")
            return out
        finally:
            m.close()

    if threads <= 1:
        return [one(p) for p in problems]

    def size(prob):
        H, q, A, b, C, d = prob
        return H.shape[0] + sum(M.shape[0] for M in (A, C) if M is not None)

    order = sorted(range(len(problems)), key=lambda k: -size(problems[k]))
    out = [None] * len(problems)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        for k, r in zip(order, ex.map(one, [problems[k] for k in order])):
            out[k] = r
    return out


def quadprog(H, q, A, b, C, d, settings=None):
    """The reference's Eigen quadprog (eigen_interface/include/nasoq/nasoq_eigen.h:13-73):
    returns (status, x, y, z).  settings: dict of QPSettings fields (eps,
    diag_perturb, max_iter, stop_tol, max_iter_nas, nasoq_variant); as in the
    reference, a variant name other than "fixed"/"tuned"/"auto" selects PREDET."""
    s = {"eps": 1e-3, "diag_perturb": 1e-9, "max_iter": 0, "stop_tol": 1e-15, "max_iter_nas": 4000,
         "nasoq_variant": "Fixed"}
    s.update(settings or {})
    qm = Nasoq(H, q, A if A is not None and A.shape[0] else None, b, C if C is not None and C.shape[0] else None, d)
    qm.diag_perturb = s["diag_perturb"]
    qm.eps_abs = s["eps"]
    qm.max_iter = s["max_iter"]
    qm.stop_tol = s["stop_tol"]
    qm.max_iter_nas = s["max_iter_nas"]
    qm.variant = {"fixed": FIXED, "tuned": TUNED, "auto": AUTO}.get(s["nasoq_variant"], PREDET)
    st = qm.solve()
    out = (st, qm.primal_vars.copy(), qm.dual_vars_eq.copy(), qm.dual_vars.copy())
    qm.close()
    return out
