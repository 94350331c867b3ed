"""TEST INFRASTRUCTURE ONLY: ctypes access to the CPU oracle.

  liboracle.so          -- the FP64 CPU restatement of the numeric LBL path
                           (lbl_oracle.c; reference file:line citations there)
  _ref/libnasoq_ref.so  -- the reference's own symbolic, SpMV and norm sources,
                           compiled in place (Makefile); present only where the
                           reference tree exists.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, always as the checker, never as the thing measured or shipped.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libnasoq_ref.so")

_i32p = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)


def _ip(a):
    return a.ctypes.data_as(_i32p) if a is not None else None


def _lp(a):
    return a.ctypes.data_as(_i64p) if a is not None else None


def _dp(a):
    return a.ctypes.data_as(_f64p) if a is not None else None


_o = None
_r = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            build()
        o = ctypes.CDLL(ORACLE_SO)
        o.oracle_sytf2.restype = ctypes.c_int
        o.oracle_sytf2.argtypes = [ctypes.c_int, _f64p, ctypes.c_int, _i32p]
        o.oracle_factor.restype = ctypes.c_int
        o.oracle_factor.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, ctypes.c_double, _i32p,
                                    ctypes.c_int, _i32p, _i64p, _i32p, _i64p, _f64p, _f64p, _i32p,
                                    _i32p, _i32p]
        o.oracle_solve_ordered.restype = None
        o.oracle_solve_ordered.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i64p, _i32p, _i64p,
                                           _f64p, _f64p, _f64p]
        o.oracle_spmv.restype = None
        o.oracle_spmv.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        o.oracle_backward_error.restype = ctypes.c_double
        o.oracle_backward_error.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        o.oracle_gmres.restype = ctypes.c_int
        o.oracle_gmres.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, ctypes.c_int, _i32p, _i64p,
                                   _i32p, _i64p, _f64p, _f64p, _f64p, _f64p, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_double]
        o.oracle_permute_lower.restype = None
        o.oracle_permute_lower.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _i32p, _i32p, _i32p,
                                           _f64p]
        o.oracle_inertia.restype = None
        o.oracle_inertia.argtypes = [ctypes.c_int, _f64p, _i64p, _i64p, _i64p]
        _o = o
    return _o


def ref_lib():
    """The reference's own compiled symbolic/SpMV/norm code, or None if absent."""
    global _r
    if _r is None:
        if not os.path.exists(REF_SO):
            return None
        r = ctypes.CDLL(REF_SO)
        r.ref_symbolic.restype = ctypes.c_void_p
        r.ref_symbolic.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _i32p, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, _i32p, _f64p]
        r.ref_symbolic_scalar.restype = ctypes.c_longlong
        r.ref_symbolic_scalar.argtypes = [ctypes.c_void_p, ctypes.c_int]
        r.ref_symbolic_ints.restype = ctypes.c_longlong
        r.ref_symbolic_ints.argtypes = [ctypes.c_void_p, ctypes.c_int, _i32p]
        r.ref_symbolic_longs.restype = ctypes.c_longlong
        r.ref_symbolic_longs.argtypes = [ctypes.c_void_p, ctypes.c_int, _i64p]
        r.ref_symbolic_cost.restype = ctypes.c_longlong
        r.ref_symbolic_cost.argtypes = [ctypes.c_void_p, _f64p]
        r.ref_symbolic_free.restype = None
        r.ref_symbolic_free.argtypes = [ctypes.c_void_p]
        r.ref_spmv_sym.restype = None
        r.ref_spmv_sym.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        r.ref_backward_error.restype = ctypes.c_double
        r.ref_backward_error.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        _r = r
    return _r


def sytf2(a):
    """Bunch-Kaufman of dense symmetric a (lower used). Returns (factored, ipiv, info)."""
    a = np.array(a, dtype=np.float64, order="F")
    n = a.shape[0]
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().oracle_sytf2(n, _dp(a), n, _ip(ipiv))
    return a, ipiv, info


class Factor:
    """Oracle factorization of lower-CSC K for a given symbolic structure."""

    def __init__(self, n, Ap, Ai, Ax, reg_diag, perm, sup, pi, s, p):
        self.n = n
        self.sup, self.pi, self.p = sup, pi, p
        self.nsuper = len(sup) - 1
        self.valL = np.zeros(int(p[-1]), dtype=np.float64)
        self.d_val = np.zeros(2 * n, dtype=np.float64)
        self.s_lab = np.zeros(len(s), dtype=np.int32)
        self.perm = np.zeros(n, dtype=np.int32)
        self.perturbed = np.zeros(n, dtype=np.int32)
        self.npiv = lib().oracle_factor(n, _ip(Ap), _ip(Ai), _dp(Ax), reg_diag, _ip(perm),
                                        self.nsuper, _ip(sup), _lp(pi), _ip(s), _lp(p),
                                        _dp(self.valL), _dp(self.d_val), _ip(self.s_lab),
                                        _ip(self.perm), _ip(self.perturbed))
        self.Ap, self.Ai, self.Ax = Ap, Ai, Ax

    def solve(self, b):
        """x = K~^{-1} b (one LDL^T solve, no refinement), original ordering."""
        xo = np.ascontiguousarray(b[self.perm], dtype=np.float64)
        lib().oracle_solve_ordered(self.n, self.nsuper, _ip(self.sup), _lp(self.pi), _ip(self.s_lab),
                                   _lp(self.p), _dp(self.valL), _dp(self.d_val), _dp(xo))
        x = np.empty_like(xo)
        x[self.perm] = xo
        return x

    def refine(self, b, x0, itr_max=2, mr=2, tol_abs=1e-15, tol_rel=1e-15):
        """GMRES refinement on the unperturbed K (perturbed diagonals zeroed)."""
        n = self.n
        Kx = self.Ax.copy()
        Kx[self.Ap[:-1][self.perturbed.astype(bool)]] = 0.0
        Bp = np.zeros(n + 1, np.int32)
        Bi = np.zeros(self.Ap[-1], np.int32)
        Bx = np.zeros(self.Ap[-1], np.float64)
        lib().oracle_permute_lower(n, _ip(self.Ap), _ip(self.Ai), _dp(Kx), _ip(self.perm), _ip(Bp),
                                   _ip(Bi), _dp(Bx))
        xo = np.ascontiguousarray(x0[self.perm], dtype=np.float64)
        bo = np.ascontiguousarray(b[self.perm], dtype=np.float64)
        it = lib().oracle_gmres(n, _ip(Bp), _ip(Bi), _dp(Bx), self.nsuper, _ip(self.sup),
                                _lp(self.pi), _ip(self.s_lab), _lp(self.p), _dp(self.valL),
                                _dp(self.d_val), _dp(xo), _dp(bo), itr_max, mr, tol_abs, tol_rel)
        x = np.empty_like(xo)
        x[self.perm] = xo
        return x, it

    def inertia(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        lib().oracle_inertia(self.n, _dp(self.d_val), ctypes.byref(a), ctypes.byref(b),
                             ctypes.byref(c))
        return a.value, b.value, c.value

    def dense_L(self):
        """Unit lower L in the final (pivoted) ordering as a dense array (small n only)."""
        n = self.n
        L = np.zeros((n, n))
        for t in range(self.nsuper):
            c0, c1 = self.sup[t], self.sup[t + 1]
            w = c1 - c0
            r = int(self.pi[t + 1] - self.pi[t])
            panel = self.valL[self.p[c0]:self.p[c0] + r * w].reshape(w, r).T
            rows = self.s_lab[self.pi[t]:self.pi[t + 1]]
            for j in range(w):
                L[rows, c0 + j] = panel[:, j]
        return np.tril(L)

    def dense_D(self):
        n = self.n
        D = np.diag(self.d_val[:n].copy())
        for i in range(n - 1):
            if self.d_val[n + i] != 0.0:
                D[i + 1, i] = D[i, i + 1] = self.d_val[n + i]
        return D


def spmv(n, Ap, Ai, Ax, x):
    y = np.zeros(n)
    lib().oracle_spmv(n, _ip(Ap), _ip(Ai), _dp(Ax), _dp(np.ascontiguousarray(x, dtype=np.float64)),
                      _dp(y))
    return y


def backward_error(n, Ap, Ai, Ax, x, b):
    return lib().oracle_backward_error(n, _ip(Ap), _ip(Ai), _dp(Ax),
                                       _dp(np.ascontiguousarray(x, dtype=np.float64)),
                                       _dp(np.ascontiguousarray(b, dtype=np.float64)))


# --------------------------------------------------------------------------
# BLAS-3 CPU restatement (cpu_lbl.c): the same factorization on DGEMM /
# DSYTRF / DTRSM of a BLAS library loaded at run time, level-parallel with
# OpenMP.  The bench's CPU baseline ("CPU restatement, not the reference
# binary").
# --------------------------------------------------------------------------
CPU_SO = os.path.join(HERE, "liboracle_cpu.so")
BLAS_CANDIDATES = ["/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1"]
_c = None


def cpu_lib():
    """The BLAS-3 restatement with its BLAS loaded, or None when no BLAS library is found."""
    global _c
    if _c is None:
        if not os.path.exists(CPU_SO):
            build()
        c = ctypes.CDLL(CPU_SO)
        c.cpu_lbl_init.restype = ctypes.c_int
        c.cpu_lbl_init.argtypes = [ctypes.c_char_p]
        c.cpu_lbl_blas_name.restype = ctypes.c_char_p
        c.cpu_lbl_factor.restype = ctypes.c_int
        c.cpu_lbl_factor.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, ctypes.c_double, _i32p, ctypes.c_int, _i32p,
                                     _i64p, _i32p, _i64p, _i32p, _f64p, _f64p, _i32p, _i32p, ctypes.c_int]
        c.cpu_lbl_solve.restype = None
        c.cpu_lbl_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i64p, _i32p, _i64p, _f64p,
                                    _f64p, _f64p, ctypes.c_int]
        ok = False
        for path in BLAS_CANDIDATES:
            if os.path.exists(path) and c.cpu_lbl_init(path.encode()) == 0:
                ok = True
                break
        if not ok:
            return None
        _c = c
    return _c


class CpuFactor:
    """FP64 LDL^T of K~ by the BLAS-3 restatement on `threads` threads, with
    the same symbolic structure as Factor; solve() works on several
    right-hand sides at once in the original ordering."""

    def __init__(self, n, Ap, Ai, Ax, reg_diag, perm, sup, pi, s, p, sparent, threads):
        c = cpu_lib()
        self.n, self.threads = n, threads
        self.sup = np.ascontiguousarray(sup, dtype=np.int32)
        self.pi = np.ascontiguousarray(pi, dtype=np.int64)
        self.p = np.ascontiguousarray(p, dtype=np.int64)
        self.nsuper = len(self.sup) - 1
        self.valL = np.empty(int(self.p[-1]), dtype=np.float64)
        self.d_val = np.empty(2 * n, dtype=np.float64)
        self.s_lab = np.empty(int(self.pi[-1]), dtype=np.int32)
        self.perm = np.empty(n, dtype=np.int32)
        args = [np.ascontiguousarray(x, dtype=t) for x, t in
                ((Ap, np.int32), (Ai, np.int32), (Ax, np.float64), (perm, np.int32), (s, np.int32),
                 (sparent, np.int32))]
        self.pivoted = c.cpu_lbl_factor(n, _ip(args[0]), _ip(args[1]), _dp(args[2]), reg_diag, _ip(args[3]),
                                        self.nsuper, _ip(self.sup), _lp(self.pi), _ip(args[4]), _lp(self.p),
                                        _ip(args[5]), _dp(self.valL), _dp(self.d_val), _ip(self.s_lab),
                                        _ip(self.perm), threads)

    def solve(self, B):
        B = np.asarray(B, dtype=np.float64)
        one = B.ndim == 1
        B2 = B.reshape(self.n, -1)
        X = np.asfortranarray(B2[self.perm, :])
        cpu_lib().cpu_lbl_solve(self.n, X.shape[1], self.nsuper, _ip(self.sup), _lp(self.pi), _ip(self.s_lab),
                                _lp(self.p), _dp(self.valL), _dp(self.d_val), _dp(X), self.threads)
        out = np.empty_like(B2)
        out[self.perm, :] = X
        return out[:, 0] if one else out

    def inertia(self):
        a, b, z = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        lib().oracle_inertia(self.n, _dp(self.d_val), ctypes.byref(a), ctypes.byref(b), ctypes.byref(z))
        return a.value, b.value, z.value
