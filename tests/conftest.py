import os
import sys

import numpy as np
import pytest

# torch's HIP runtime must be loaded before the engine library's when both
# are used in one process (tests hand torch device tensors to the C-ABI):
# with the engine's runtime initialised first, torch finds no HIP device
import torch  # noqa: F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run with -m gpu)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ref(oracle):
    r = oracle.ref_lib()
    if r is None:
        pytest.skip("reference build oracle/_ref not available (reference tree absent)")
    return r


def dense_from_factor(n, sup, pi, px, valL, rows, dval):
    """Dense unit-lower L (pivoted ordering) and block-diagonal D."""
    L = np.zeros((n, n))
    for t in range(len(sup) - 1):
        c0, c1 = sup[t], sup[t + 1]
        w = c1 - c0
        r = int(pi[t + 1] - pi[t])
        panel = valL[px[t]:px[t] + r * w].reshape(w, r).T
        lab = rows[pi[t]:pi[t + 1]].copy()
        lab[:w] = np.arange(c0, c1)
        for j in range(w):
            L[lab, c0 + j] = panel[:, j]
    L = np.tril(L)
    D = np.diag(dval[:n].copy())
    for i in range(n - 1):
        if dval[n + i] != 0.0:
            D[i + 1, i] = D[i, i + 1] = dval[n + i]
    return L, D
