"""Learned warm start on the GPU: the BF16 MFMA lane layout, the MLP's
forward, back-propagation and Adam kernels against the CPU oracle
(oracle/mlp_oracle.py), and the warm start judged by NASOQ on a small
parametric QP family."""
import numpy as np
import pytest

import mlp_oracle as M
from nasoq_amd import _lib, learned, qp
from nasoq_amd._lib import fptr

pytestmark = pytest.mark.gpu


def block_errors(g, ref, dims):
    o = M.layout(*dims)
    keys = ["W1", "b1", "W2", "b2", "W3", "b3", "n"]
    out = {}
    for a, b in zip(keys[:-1], keys[1:]):
        d = np.linalg.norm(g[o[a]:o[b]] - ref[o[a]:o[b]])
        out[a] = d / max(np.linalg.norm(ref[o[a]:o[b]]), 1e-30)
    return out


def test_bf16_mfma_lane_layout():
    lib = _lib.load()
    rng = np.random.default_rng(0)
    A = rng.uniform(-1, 1, (16, 16)).astype(np.float32)
    B = rng.uniform(-1, 1, (16, 16)).astype(np.float32)
    C = np.zeros((16, 16), np.float32)
    _lib.check(lib.lblg_selftest_mfma_bf16(fptr(A), fptr(B), fptr(C)), "selftest bf16")
    ref = M.to_bf16(A).astype(np.float64) @ M.to_bf16(B).astype(np.float64)
    assert np.max(np.abs(C - ref)) <= 1e-5


def make_net(dims, batch, seed, scale=0.3):
    net = learned.ActiveSetPredictor(dims[0], dims[3], hidden=(dims[1], dims[2]), seed=seed, max_batch=batch)
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(net.num_params) * scale / np.sqrt(max(dims[0], dims[1]) / 16)).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    return net, p


@pytest.mark.parametrize("dims,batch", [((7, 20, 13, 5), 3), ((32, 64, 48, 40), 37), ((32, 256, 256, 200), 128),
                                        ((32, 1024, 1024, 160), 64)])
def test_gradients_vs_oracle(dims, batch):
    print(f"case mlp gradients dims={dims} batch={batch}", flush=True)
    net, p = make_net(dims, batch, seed=sum(dims))
    rng = np.random.default_rng(batch)
    X = rng.standard_normal((batch, dims[0])).astype(np.float32)
    Y = (rng.uniform(size=(batch, dims[3])) < 0.3).astype(np.float32)
    loss, g = net.gradients(X, Y)
    lo, go = M.loss_grad(p, X, Y, dims, bf16=True)
    l64, g64 = M.loss_grad(p.astype(np.float64), X, Y, dims, bf16=False)
    assert abs(loss - lo) <= 1e-4 * abs(lo), (loss, lo)
    for k, e in block_errors(g, go, dims).items():
        assert e <= 2e-3, (k, e)
    # the BF16 arithmetic stays within its own error budget of the exact gradient
    assert np.linalg.norm(g - g64) <= 3e-2 * np.linalg.norm(g64)
    assert abs(loss - l64) <= 1e-2 * abs(l64)
    net.close()


def test_predict_vs_oracle():
    dims = (32, 128, 96, 70)
    net, p = make_net(dims, 50, seed=3)
    X = np.random.default_rng(4).standard_normal((123, dims[0])).astype(np.float32)  # 3 passes of max_batch
    P = net.predict_raw(X)
    assert np.max(np.abs(P - M.predict(p, X, dims, bf16=True))) <= 2e-3
    net.close()


def test_adam_steps_vs_oracle():
    dims = (16, 64, 32, 24)
    net, p0 = make_net(dims, 1, seed=5)
    rng = np.random.default_rng(6)
    X = rng.standard_normal((1, dims[0])).astype(np.float32)
    Y = (rng.uniform(size=(1, dims[3])) < 0.5).astype(np.float32)
    lr = 1e-3
    m = v = np.zeros(net.num_params)
    p = p0.astype(np.float64)
    for t in (1, 2, 3):
        _, g = net.gradients(X, Y)  # the gradient train() computes at these parameters (one sample)
        net.train_raw(X, Y, batch=1, epochs=1, lr=lr, seed=t)
        p_ref, m, v = M.adam_step(p, g.astype(np.float64), m, v, t, lr)
        got = net.params()
        assert np.max(np.abs(got - p_ref)) <= 1e-6 + 1e-6 * np.abs(p_ref).max(), t
        p = got.astype(np.float64)
    net.close()


def test_training_fits_a_learnable_rule():
    # labels = sign pattern of a fixed linear map of the features
    dims = (32, 256, 256, 64)
    rng = np.random.default_rng(7)
    W = rng.standard_normal((32, 64))
    X = rng.standard_normal((2048, 32)).astype(np.float32)
    Y = (X @ W > 0).astype(np.float32)
    net = learned.ActiveSetPredictor(32, 64, hidden=(256, 256), seed=1, max_batch=128)
    p0 = net.params()
    l0 = M.loss(p0, X[:256], Y[:256], dims, bf16=True)
    last = net.train_raw(X, Y, batch=128, epochs=20, lr=2e-3, seed=2)
    assert last <= 0.25 * l0, (last, l0)
    Xt = rng.standard_normal((512, 32)).astype(np.float32)
    acc = np.mean((net.predict_raw(Xt) > 0.5) == (Xt @ W > 0))
    assert acc >= 0.9, acc
    net.close()


def _family_solutions(fam, S, thetas):
    out = []
    for th in thetas:
        st, x, z, act, it, nf = learned.solve_cold(S, fam.q(th))
        assert st == qp.OPTIMAL
        out.append((x, z, act, it))
    return out


def test_predicted_active_set_edge_cases():
    # the exact active set: no loop iterations after the single KKT solve;
    # every constraint predicted, or none: NASOQ still ends at the same optimum
    fam = learned.ParametricQpFamily(grid=8, seed=3)
    S = fam.solver()
    for th in fam.thetas(3, 4):
        q = fam.q(th)
        st, x, z, act, it, nf = learned.solve_cold(S, q)
        assert st == qp.OPTIMAL and it > 0
        for pred, exact in ((act, True), (np.arange(fam.mi, dtype=np.int32), False), (np.zeros(0, np.int32), False)):
            stw, xw, zw, actw, itw, nfw, fb = learned.solve_warm(S, q, pred)
            assert stw == qp.OPTIMAL
            assert np.linalg.norm(xw - x) <= 1e-8 * max(1.0, np.linalg.norm(x))
            assert max(fam.kkt_residuals(xw, zw, th)) <= 1e-7
            if exact:
                assert itw <= 2 and not fb  # the KKT solve of the set (a prune at most), no loop
    S.close()


def test_learned_warm_start_halves_iterations():
    # plain NASOQ labels a parametric family; the MLP trained on those labels
    # (on the GPU) predicts active sets that NASOQ turns into the same optimum
    # with at least 2x fewer active-set iterations (pruning solves counted)
    fam = learned.ParametricQpFamily(grid=12, seed=0)
    S = fam.solver()
    train = fam.thetas(240, 1)
    Y, iters, secs = learned.generate(fam, train, solver=S)
    print(f"case learned grid=12 labels={len(train)} cold iters {iters.mean():.1f} ({secs:.1f}s)", flush=True)
    net = learned.ActiveSetPredictor(fam.n_theta, fam.mi, hidden=(256, 256), seed=0, max_batch=64)
    loss = net.fit(train, Y, epochs=60, batch=32, lr=1e-3)
    assert loss < 0.1, loss
    test = fam.thetas(20, 2)
    res = learned.evaluate(fam, net, test, solver=S)
    print(f"cold {res['cold_iter'].sum()} warm {res['warm_iter'].sum()} fallbacks {res['fallback'].sum()}", flush=True)
    assert res["dx"].max() <= 1e-8 and res["kkt"].max() <= 1e-7
    assert res["warm_iter"].sum() * 2 <= res["cold_iter"].sum()
    # the same network untrained gives (nearly) nothing: the gain is learned
    raw = learned.ActiveSetPredictor(fam.n_theta, fam.mi, hidden=(256, 256), seed=0, max_batch=64)
    raw.mu, raw.sd = net.mu, net.sd
    res0 = learned.evaluate(fam, raw, test[:8], solver=S)
    assert res["warm_iter"][:8].sum() < res0["warm_iter"].sum()
    for m in (net, raw, S):
        m.close()
