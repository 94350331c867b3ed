"""The CPU oracle of the active-set predictor (oracle/mlp_oracle.py) pinned
against what fixes it independently of its own formulas: torch's BF16
rounding, torch autograd and central finite differences for the gradient of
the mean BCE, and the closed forms of Adam's first step and of Adam under a
constant gradient."""
import numpy as np
import pytest
import torch

import mlp_oracle as M

DIMS = (5, 7, 6, 9)


def rand_params(dims, seed):
    rng = np.random.default_rng(seed)
    n = M.layout(*dims)["n"]
    return rng.standard_normal(n) * 0.5


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 20000),
                        np.float32([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0, -0.0, 65504.0])])
    ours = M.to_bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
    # ties go to the even mantissa
    assert M.to_bf16(np.float32([1 + 2 ** -8]))[0] == 1.0
    assert M.to_bf16(np.float32([1 + 3 * 2 ** -8]))[0] == np.float32(1 + 2 ** -6)


def test_layout_covers_parameters_once():
    o = M.layout(*DIMS)
    n_in, h1, h2, out = DIMS
    assert o["n"] == h1 * n_in + h1 + h2 * h1 + h2 + out * h2 + out
    p = np.arange(o["n"], dtype=np.float64)
    parts = M.unpack(p, *DIMS)
    assert np.array_equal(np.sort(np.concatenate([a.ravel() for a in parts])), p)


def test_loss_and_gradient_match_torch_autograd():
    rng = np.random.default_rng(1)
    p = rand_params(DIMS, 2)
    X = rng.standard_normal((11, DIMS[0]))
    Y = (rng.uniform(size=(11, DIMS[3])) < 0.4).astype(np.float64)
    l, g = M.loss_grad(p, X, Y, DIMS, bf16=False)
    pt = torch.tensor(p, requires_grad=True)
    W1, b1, W2, b2, W3, b3 = M.unpack(pt, *DIMS)
    Xt = torch.tensor(X)
    Z = torch.relu(torch.relu(Xt @ W1.T + b1) @ W2.T + b2) @ W3.T + b3
    lt = torch.nn.functional.binary_cross_entropy_with_logits(Z, torch.tensor(Y), reduction="mean")
    lt.backward()
    assert abs(l - lt.item()) <= 1e-12 * max(1.0, abs(l))
    assert np.allclose(g, pt.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rand_params(DIMS, 4)
    X = rng.standard_normal((6, DIMS[0]))
    Y = (rng.uniform(size=(6, DIMS[3])) < 0.5).astype(np.float64)
    _, g = M.loss_grad(p, X, Y, DIMS, bf16=False)
    h = 1e-6
    fd = np.zeros_like(p)
    for k in range(len(p)):
        e = np.zeros_like(p)
        e[k] = h
        fd[k] = (M.loss(p + e, X, Y, DIMS) - M.loss(p - e, X, Y, DIMS)) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.abs(g).max())


def test_bf16_gradient_close_to_exact():
    # BF16 operands (8 significant bits) with FP32 accumulation: relative error ~1e-2
    dims = (32, 64, 48, 40)
    rng = np.random.default_rng(5)
    p = rand_params(dims, 6) * 0.3
    X = rng.standard_normal((37, dims[0]))
    Y = (rng.uniform(size=(37, dims[3])) < 0.3).astype(np.float64)
    l64, g64 = M.loss_grad(p, X, Y, dims, bf16=False)
    l16, g16 = M.loss_grad(p, X, Y, dims, bf16=True)
    assert g16.dtype == np.float32
    assert abs(l16 - l64) <= 1e-2 * abs(l64)
    assert np.linalg.norm(g16 - g64) <= 3e-2 * np.linalg.norm(g64)


@pytest.mark.parametrize("lr", [1e-3, 0.1])
def test_adam_closed_forms(lr):
    rng = np.random.default_rng(7)
    p0 = rng.standard_normal(50)
    g = rng.standard_normal(50)
    m = v = np.zeros(50)
    # first step: bias corrections make the move -lr g / (|g| + eps) exactly
    p1, m, v = M.adam_step(p0, g, m, v, 1, lr, eps=0.0)
    assert np.allclose(p1 - p0, -lr * np.sign(g), rtol=1e-12, atol=0)
    # a constant gradient keeps m_hat = g and v_hat = g^2 at every step
    p = p1
    for t in range(2, 8):
        pn, m, v = M.adam_step(p, g, m, v, t, lr, eps=1e-8)
        assert np.allclose(pn - p, -lr * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)
        p = pn
    # a gradient flipping sign at step 2: m_hat = (0.09 g - 0.1 g) / 0.19, v_hat = g^2
    p1, m, v = M.adam_step(p0, g, np.zeros(50), np.zeros(50), 1, lr, eps=0.0)
    p2, _, _ = M.adam_step(p1, -g, m, v, 2, lr, eps=0.0)
    assert np.allclose(p2 - p1, -lr * (-0.01 / 0.19) * np.sign(g), rtol=1e-10, atol=0)
