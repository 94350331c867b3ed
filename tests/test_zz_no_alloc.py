"""lblg_backward_error_device_noalloc against lblg_backward_error_device.

The no-allocation entry point shares the residual kernels with the
allocating one and keeps its partial sums in the library's fixed device
array instead of a per-handle buffer.  Two calls in a row, the second on a
smaller system with fewer right-hand sides (the array then holds the first
call's partials beyond what the second reads), must each equal the
allocating entry point's result bit for bit, and the oracle's backward error
to rounding.
"""
import ctypes

import numpy as np
import pytest

from nasoq_amd import lbl
from nasoq_amd import problems as P

pytestmark = pytest.mark.gpu


def _factored(grid, m, seed, nrhs):
    n, Ap, Ai, Ax, B = P.kkt_config(grid, m, 6, seed=seed, nrhs=nrhs)
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = 1e-9
    assert s.symbolic_analysis() == 1
    assert s.numerical_factorization() == 1
    X = s.solve_only(B)
    X[::5, :] += 1e-6 * (1.0 + np.arange(nrhs))  # not the refined solutions: visible residuals in every column
    return s, n, Ap, Ai, Ax, B, X


def _errors(s, entry, B, X):
    import torch
    n, nrhs = B.shape
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()  # column-major n x nrhs
    dX = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
    dE = torch.full((nrhs,), -1.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = getattr(s.lib, entry)(s.h, nrhs, ctypes.c_void_p(dB.data_ptr()), n, ctypes.c_void_p(dX.data_ptr()), n,
                               ctypes.c_void_p(dE.data_ptr()))
    assert rc == 0, (entry, rc)
    assert s.lib.lblg_synchronize(s.h) == 0
    return dE.cpu().numpy()


def test_noalloc_backward_error_matches_allocating_entry(oracle):
    cases = [(12, 300, 21, 4), (6, 30, 22, 2)]  # the second call: a smaller system, fewer columns
    for grid, m, seed, nrhs in cases:
        s, n, Ap, Ai, Ax, B, X = _factored(grid, m, seed, nrhs)
        print(f"case noalloc n={n} nrhs={nrhs}", flush=True)
        got = _errors(s, "lblg_backward_error_device_noalloc", B, X)
        want = _errors(s, "lblg_backward_error_device", B, X)
        assert np.array_equal(got, want), (got, want)
        for c in range(nrhs):
            ref = oracle.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c])
            # the residual cancels to ~1e-7 of its terms: summation order shows at ~1e-12
            assert abs(got[c] - ref) <= 1e-10 * ref, (c, got[c], ref)
        # beyond the limit: a clear error, nothing launched
        assert s.lib.lblg_backward_error_device_noalloc(s.h, 9, None, n, None, n, None) == -1
        s.close()
