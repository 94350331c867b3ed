"""TEST INFRASTRUCTURE ONLY: CPU restatement of the active-set predictor of
the learned warm start (the MLP behind lblg_mlp_*), in numpy.

No reference counterpart (NASOQ has no learned component); the functions
restate the textbook definitions they are pinned against in
tests/test_mlp_oracle.py:

  forward     in -> h1 -> h2 -> out, ReLU hidden layers, logits out;
              parameters one vector [W1 (h1 x in, row-major), b1, W2, b2, W3, b3]
  loss_grad   binary cross-entropy of sigmoid(logits) vs the 0/1 labels,
              averaged over all batch x output entries, and its gradient by
              back-propagation
              (pinned by central finite differences in float64)
  adam_step   Adam (Kingma & Ba, "Adam: A Method for Stochastic Optimization",
              ICLR 2015, Algorithm 1; pinned by its closed forms: the first
              step and every step under a constant gradient move each
              parameter by -lr g / (|g| + eps))

bf16=True emulates the device arithmetic: every GEMM operand (weights,
activations, back-propagated errors, inputs) is rounded to BF16 (round to
nearest even) and products are accumulated in float32; elementwise work is
float32.  bf16=False runs the same algebra in float64.

Only tests/ use this module, always as the checker.
"""
import numpy as np


def to_bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def layout(n_in, h1, h2, n_out):
    """Offsets of W1, b1, W2, b2, W3, b3 in the parameter vector, and its length."""
    o = {}
    o["W1"] = 0
    o["b1"] = o["W1"] + h1 * n_in
    o["W2"] = o["b1"] + h1
    o["b2"] = o["W2"] + h2 * h1
    o["W3"] = o["b2"] + h2
    o["b3"] = o["W3"] + n_out * h2
    o["n"] = o["b3"] + n_out
    return o


def unpack(p, n_in, h1, h2, n_out):
    o = layout(n_in, h1, h2, n_out)
    W1 = p[o["W1"]:o["b1"]].reshape(h1, n_in)
    b1 = p[o["b1"]:o["W2"]]
    W2 = p[o["W2"]:o["b2"]].reshape(h2, h1)
    b2 = p[o["b2"]:o["W3"]]
    W3 = p[o["W3"]:o["b3"]].reshape(n_out, h2)
    b3 = p[o["b3"]:o["n"]]
    return W1, b1, W2, b2, W3, b3


def _mm(a, b, bf16):
    if bf16:
        return to_bf16(a) @ to_bf16(b)  # float32 accumulation
    return a @ b


def forward(p, X, dims, bf16=True):
    """Activations (A1, A2) and logits Z3 of a batch X (rows = samples)."""
    dt = np.float32 if bf16 else np.float64
    p = np.asarray(p, dt)
    X = np.asarray(X, dt)
    W1, b1, W2, b2, W3, b3 = unpack(p, *dims)
    A1 = np.maximum(_mm(X, W1.T, bf16) + b1, 0)
    A2 = np.maximum(_mm(A1, W2.T, bf16) + b2, 0)
    Z3 = _mm(A2, W3.T, bf16) + b3
    return A1, A2, Z3


def predict(p, X, dims, bf16=True):
    _, _, Z3 = forward(p, X, dims, bf16)
    return 1.0 / (1.0 + np.exp(-Z3))


def loss(p, X, Y, dims, bf16=False):
    """Mean over all batch x output entries of log(1 + e^z) - y z."""
    _, _, Z = forward(p, X, dims, bf16)
    Y = np.asarray(Y, Z.dtype)
    return float(np.sum(np.maximum(Z, 0) - Z * Y + np.log1p(np.exp(-np.abs(Z)))) / Z.size)


def loss_grad(p, X, Y, dims, bf16=True):
    """(loss, gradient vector) of the mean BCE over all entries, by back-propagation."""
    dt = np.float32 if bf16 else np.float64
    p = np.asarray(p, dt)
    X = np.asarray(X, dt)
    Y = np.asarray(Y, dt)
    W1, b1, W2, b2, W3, b3 = unpack(p, *dims)
    A1, A2, Z3 = forward(p, X, dims, bf16)
    P = 1.0 / (1.0 + np.exp(-Z3))
    dZ3 = ((P - Y) / Z3.size).astype(dt)
    lval = float(np.sum(np.maximum(Z3, 0) - Z3 * Y + np.log1p(np.exp(-np.abs(Z3)))) / Z3.size)
    g = np.zeros_like(p)
    o = layout(*dims)
    g[o["W3"]:o["b3"]] = _mm(dZ3.T, A2, bf16).ravel()
    g[o["b3"]:o["n"]] = dZ3.sum(0)
    dA2 = _mm(dZ3, W3, bf16) * (A2 > 0)
    g[o["W2"]:o["b2"]] = _mm(dA2.T, A1, bf16).ravel()
    g[o["b2"]:o["W3"]] = dA2.sum(0)
    dA1 = _mm(dA2, W2, bf16) * (A1 > 0)
    g[o["W1"]:o["b1"]] = _mm(dA1.T, X, bf16).ravel()
    g[o["b1"]:o["W2"]] = dA1.sum(0)
    return lval, g


def adam_step(p, g, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step at step number t (1-based); returns (p, m, v) updated."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mhat = m / (1 - beta1 ** t)
    vhat = v / (1 - beta2 ** t)
    return p - lr * mhat / (np.sqrt(vhat) + eps), m, v
