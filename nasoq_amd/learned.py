"""Learned warm start for a parametric QP family.

A family fixes H, the inequality rows B x <= d and a linear map U; each
member has q = U theta for a parameter theta.  Plain NASOQ (qp.Nasoq, every
KKT factorization and solve on the GPU) solves sampled members cold; the
active sets it returns label an MLP theta -> P(constraint active) that is
trained and evaluated by the library's own HIP kernels (lblg_mlp_*: BF16
MFMA products with FP32 accumulation, back-propagation, Adam).  A predicted
active set is handed to Nasoq.solve_predicted: the predicted constraints are
made equalities, those with negative multipliers dropped, and NASOQ's dual
active-set loop (src/nasoq.cpp:478-610 in the reference) continues from
that dual-feasible point; its KKT optimality check (check_solve_status,
src/nasoq.cpp:712-728) judges the result and anything short of OPTIMAL falls
back to a cold start.  The solution returned is therefore always NASOQ's own.

No reference counterpart: NASOQ has no learned component.
"""
import ctypes
import time

import numpy as np

from . import _lib, qp
from ._lib import check, fptr
from .problems import laplacian_2d, random_rows


class ParametricQpFamily:
    """min 1/2 x'Hx + (U theta)'x  s.t.  B x <= d, with H the 5-point
    Laplacian of a grid x grid mesh + 1e-2 I, round(ineq_ratio n) rows of
    nnz_row N(0,1) entries at uniform random columns, d ~ U[0, 0.5] and
    U ~ N(0, 1/n) of n x n_theta; all from one seeded generator."""

    def __init__(self, grid=100, ineq_ratio=0.4, nnz_row=6, n_theta=32, seed=0):
        rng = np.random.default_rng(seed)
        self.H = laplacian_2d(grid, grid, 1e-2)
        self.n = self.H.shape[0]
        self.mi = int(round(ineq_ratio * self.n))
        self.B = random_rows(rng, self.mi, self.n, nnz_row).tocsc()
        self.d = rng.uniform(0.0, 0.5, self.mi)
        self.U = rng.standard_normal((self.n, n_theta)) / np.sqrt(self.n)
        self.n_theta = n_theta
        self.separate_rows = None  # solver(): the Nasoq default unless set

    def q(self, theta):
        return self.U @ np.asarray(theta, dtype=np.float64)

    def thetas(self, count, seed):
        return np.random.default_rng(seed).standard_normal((count, self.n_theta))

    def solver(self, eps_abs=1e-8, autotune=False, separate_rows=None):
        """A Nasoq instance of the family (analysis and device buffers kept
        across members; set_linear swaps q); autotune: pick the KKT's device
        schedule once for the whole family; separate_rows: the KKT's
        inequality rows as one-column supernodes (1, the reference's) or in
        supernodes with other columns (0); None keeps the family's setting."""
        m = qp.Nasoq(self.H, np.zeros(self.n), B=self.B, b=self.d)
        m.eps_abs = eps_abs
        m.variant = qp.FIXED
        sep = self.separate_rows if separate_rows is None else separate_rows
        if sep is not None:
            m.separate_rows = sep
        if autotune:
            m.autotune()
        return m

    def kkt_residuals(self, x, z, theta):
        """(stationarity, primal infeasibility, dual infeasibility,
        complementarity) infinity norms of a candidate solution, computed
        from the problem data on the host."""
        Hs = self.H
        r = Hs @ x + self.q(theta) + self.B.T @ z
        s = self.B @ x - self.d
        return (float(np.abs(r).max()), float(max(s.max(), 0.0)), float(max(-z.min(), 0.0)) if len(z) else 0.0,
                float(np.abs(z * s).max()) if len(z) else 0.0)


def _factorizations(solver):
    return int(getattr(solver, "factorizations", 0))


def solve_cold(solver, q):
    """Plain NASOQ: (status, x, z, active set, active-set iterations, KKT
    factorizations)."""
    solver.set_linear(q=q)
    f0 = _factorizations(solver)
    st = solver.solve()
    return (st, solver.primal_vars.copy(), solver.dual_vars.copy(), np.array(solver.active_set, dtype=np.int32),
            solver.num_iter, _factorizations(solver) - f0)


def solve_warm(solver, q, predicted):
    """Warm start from a predicted active set, falling back to a cold start
    unless NASOQ's check reports OPTIMAL.  Returns (status, x, z, active set,
    active-set iterations, KKT factorizations, fell_back); iterations and
    factorizations count everything spent, the prediction's pruning solves
    (one factorization each) and a fallback included."""
    solver.set_linear(q=q)
    f0 = _factorizations(solver)
    st, prunes = solver.solve_predicted(predicted)
    iters = solver.num_iter + prunes
    fell_back = False
    if st != qp.OPTIMAL:
        fell_back = True
        st = solver.solve()
        iters += solver.num_iter
    return (st, solver.primal_vars.copy(), solver.dual_vars.copy(), np.array(solver.active_set, dtype=np.int32),
            iters, _factorizations(solver) - f0, fell_back)


def _in_threads(family, count, threads, solver, work):
    """Run work(solver, k) for k < count: on `solver` alone (threads <= 1) or
    on `threads` host threads, each with its own solver of the family (its
    own engine handle and HIP stream; ctypes releases the interpreter lock
    in the library, so the members' KKT work overlaps on the GPU)."""
    if threads <= 1:
        own = solver is None
        solver = solver or family.solver()
        try:
            for k in range(count):
                work(solver, k)
        finally:
            if own:
                solver.close()
        return
    from concurrent.futures import ThreadPoolExecutor
    import threading
    nxt = iter(range(count))
    lock = threading.Lock()

    def run():
        S = family.solver()
        try:
            while True:
                with lock:
                    k = next(nxt, None)
                if k is None:
                    return
                work(S, k)
        finally:
            S.close()

    with ThreadPoolExecutor(max_workers=threads) as ex:
        for f in [ex.submit(run) for _ in range(threads)]:
            f.result()


def generate(family, thetas, solver=None, log=None, threads=1):
    """Label parameters with plain NASOQ's active sets: returns (Y (count x mi)
    0/1 float32, iterations per solve, seconds).  threads > 1: members are
    solved concurrently, one solver per thread."""
    Y = np.zeros((len(thetas), family.mi), dtype=np.float32)
    iters = np.zeros(len(thetas), dtype=np.int64)
    done = [0]

    def work(S, k):
        st, _, _, act, it, _ = solve_cold(S, family.q(thetas[k]))
        if st != qp.OPTIMAL:
            raise RuntimeError(f"plain NASOQ returned status {st} on training member {k}")
        Y[k, act] = 1.0
        iters[k] = it
        done[0] += 1
        if log and (threads > 1 or done[0] % max(1, len(thetas) // 10) == 0):
            log(f"labelled {done[0]}/{len(thetas)} ({it} active-set iterations)")

    t0 = time.perf_counter()
    _in_threads(family, len(thetas), threads, solver, work)
    return Y, iters, time.perf_counter() - t0


class ActiveSetPredictor:
    """MLP n_theta -> h1 -> h2 -> mi (ReLU, sigmoid outputs) on the GPU."""

    def __init__(self, n_in, n_out, hidden=(1024, 1024), seed=0, max_batch=256):
        self.lib = _lib.load()
        self.dims = (n_in, hidden[0], hidden[1], n_out)
        self.max_batch = max_batch
        self.h = ctypes.c_void_p()
        check(self.lib.lblg_mlp_create(ctypes.byref(self.h), n_in, hidden[0], hidden[1], n_out, seed, max_batch),
              "lblg_mlp_create")
        self.num_params = int(self.lib.lblg_mlp_num_params(self.h))
        # feature scaling fixed at the first fit (theta is already N(0, I) for the family)
        self.mu = np.zeros(n_in, dtype=np.float32)
        self.sd = np.ones(n_in, dtype=np.float32)

    def close(self):
        if self.h:
            self.lib.lblg_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _features(self, thetas):
        return np.ascontiguousarray((np.asarray(thetas, dtype=np.float32) - self.mu) / self.sd, dtype=np.float32)

    def params(self):
        p = np.zeros(self.num_params, dtype=np.float32)
        check(self.lib.lblg_mlp_get_params(self.h, fptr(p)), "lblg_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (self.num_params,)
        check(self.lib.lblg_mlp_set_params(self.h, fptr(p)), "lblg_mlp_set_params")

    def gradients(self, X, Y):
        """(loss, gradient) of the mean BCE on one batch of raw features X."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        assert X.shape[0] == Y.shape[0] <= self.max_batch
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = ctypes.c_float()
        check(self.lib.lblg_mlp_gradients(self.h, fptr(X), fptr(Y), X.shape[0], fptr(g), ctypes.byref(loss)),
              "lblg_mlp_gradients")
        return loss.value, g

    def train_raw(self, X, Y, batch, epochs, lr, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        assert X.shape[0] == Y.shape[0] and batch <= self.max_batch
        loss = ctypes.c_float()
        check(self.lib.lblg_mlp_train(self.h, fptr(X), fptr(Y), X.shape[0], batch, epochs, lr, seed,
                                      ctypes.byref(loss)), "lblg_mlp_train")
        return loss.value

    def fit(self, thetas, Y, epochs=30, batch=128, lr=1e-3, seed=0):
        th = np.asarray(thetas, dtype=np.float32)
        self.mu = th.mean(0)
        self.sd = th.std(0) + 1e-6
        return self.train_raw(self._features(th), Y, min(batch, len(th)), epochs, lr, seed)

    def predict_raw(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        P = np.zeros((X.shape[0], self.dims[3]), dtype=np.float32)
        check(self.lib.lblg_mlp_predict(self.h, fptr(X), X.shape[0], fptr(P)), "lblg_mlp_predict")
        return P

    def predict_proba(self, thetas):
        return self.predict_raw(self._features(thetas))

    def predict_active(self, thetas, threshold=0.5):
        P = self.predict_proba(thetas)
        return [np.flatnonzero(p > threshold).astype(np.int32) for p in P]


def evaluate(family, predictor, thetas, solver=None, threshold=0.5, log=None, threads=1):
    """Cold vs warm on held-out parameters.  Returns a dict with per-member
    iterations and KKT factorizations of both, fallbacks, the largest relative
    difference of the two solutions, the largest KKT residual of the warm
    solutions, and the wall time of each method (all members cold, then all
    warm; threads > 1 solves members concurrently in both phases)."""
    preds = predictor.predict_active(thetas, threshold)
    cnt = len(thetas)
    cold = [None] * cnt
    warm = [None] * cnt

    def work_cold(S, k):
        cold[k] = solve_cold(S, family.q(thetas[k]))
        if log and threads > 1:
            log(f"member {k} cold: {cold[k][4]} iterations")

    def work_warm(S, k):
        warm[k] = solve_warm(S, family.q(thetas[k]), preds[k])
        if log and threads > 1:
            log(f"member {k} warm: {warm[k][4]} iterations")

    t0 = time.perf_counter()
    _in_threads(family, cnt, threads, solver, work_cold)
    t1 = time.perf_counter()
    _in_threads(family, cnt, threads, solver, work_warm)
    t2 = time.perf_counter()
    out = {k: [] for k in ("cold_iter", "warm_iter", "cold_fact", "warm_fact", "fallback", "dx", "kkt",
                           "pred_size", "true_size", "hits")}
    for k in range(cnt):
        stc, xc, zc, actc, itc, svc = cold[k]
        stw, xw, zw, actw, itw, svw, fb = warm[k]
        if stc != qp.OPTIMAL or stw != qp.OPTIMAL:
            raise RuntimeError(f"member {k}: cold status {stc}, warm status {stw}")
        out["cold_iter"].append(itc)
        out["warm_iter"].append(itw)
        out["cold_fact"].append(svc)
        out["warm_fact"].append(svw)
        out["fallback"].append(fb)
        out["dx"].append(float(np.linalg.norm(xw - xc) / max(1.0, np.linalg.norm(xc))))
        out["kkt"].append(max(family.kkt_residuals(xw, zw, thetas[k])))
        out["pred_size"].append(len(preds[k]))
        out["true_size"].append(len(actc))
        out["hits"].append(len(np.intersect1d(preds[k], actc)))
    if log:
        log(f"evaluated {cnt} members cold and warm")
    res = {k: np.array(v) for k, v in out.items()}
    res["cold_seconds"] = t1 - t0
    res["warm_seconds"] = t2 - t1
    return res


def label_for(family, thetas, budget_s, threads):
    """Label members with plain NASOQ on `threads` concurrent solvers until
    `budget_s` seconds have passed (members started by then finish).
    Returns (thetas, Y, iterations, seconds) of the labelled members."""
    import threading
    count = len(thetas)
    Y = np.zeros((count, family.mi), np.float32)
    iters = np.full(count, -1, np.int64)
    t0 = time.perf_counter()
    lock = threading.Lock()
    nxt = [0]

    def run():
        S = family.solver()
        try:
            while True:
                with lock:
                    k = nxt[0]
                    if k >= count or time.perf_counter() - t0 > budget_s:
                        return
                    nxt[0] += 1
                st, _, _, act, it, _ = solve_cold(S, family.q(thetas[k]))
                if st != qp.OPTIMAL:
                    raise RuntimeError(f"plain NASOQ returned status {st} on member {k}")
                Y[k, act] = 1.0
                iters[k] = it
                print(f"labelled member {k}: {it} iterations, {len(act)} active, "
                      f"t={time.perf_counter() - t0:.0f}s", flush=True)
        finally:
            S.close()

    ths = [threading.Thread(target=run) for _ in range(max(1, threads))]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    done = iters >= 0
    return thetas[done], Y[done], iters[done], time.perf_counter() - t0


def _main(argv=None):
    """Labelling and evaluation in separate runs (a large family's labels can
    take longer than one run may):
      python -m nasoq_amd.learned label --grid 100 --count 200 --budget 850 --out labels.npz
      python -m nasoq_amd.learned evaluate --grid 100 --labels labels.npz --test 12
    Both take --threads and --separate-rows; evaluate prints one JSON line."""
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m nasoq_amd.learned")
    ap.add_argument("mode", choices=["label", "evaluate"])
    ap.add_argument("--grid", type=int, default=100)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--separate-rows", type=int, default=1, choices=[0, 1])
    ap.add_argument("--count", type=int, default=200, help="label: members to label at most")
    ap.add_argument("--skip", type=int, default=0, help="label: start at this training member (continue a run)")
    ap.add_argument("--budget", type=float, default=850.0, help="label: seconds after which no member is started")
    ap.add_argument("--out", default="labels.npz")
    ap.add_argument("--labels", default="labels.npz",
                    help="evaluate: labels written by the label mode (several files: comma-separated)")
    ap.add_argument("--test", type=int, default=12, help="evaluate: held-out members")
    ap.add_argument("--epochs", type=int, default=60)
    a = ap.parse_args(argv)
    fam = ParametricQpFamily(grid=a.grid, seed=0)
    fam.separate_rows = a.separate_rows
    if a.mode == "label":
        th, Y, it, secs = label_for(fam, fam.thetas(a.skip + a.count, 1)[a.skip:], a.budget, a.threads)
        np.savez(a.out, thetas=th, Y=Y, iters=it, seconds=secs)
        print(f"saved {len(th)} labelled members in {secs:.0f}s", flush=True)
        return
    parts = [np.load(f) for f in a.labels.split(",")]
    d = {"thetas": np.concatenate([q["thetas"] for q in parts]), "Y": np.concatenate([q["Y"] for q in parts]),
         "iters": np.concatenate([q["iters"] for q in parts]), "seconds": sum(float(q["seconds"]) for q in parts)}
    net = ActiveSetPredictor(fam.n_theta, fam.mi, hidden=(1024, 1024), seed=0, max_batch=128)
    t0 = time.perf_counter()
    loss = net.fit(d["thetas"], d["Y"], epochs=a.epochs, batch=16, lr=1e-3)
    t_train = time.perf_counter() - t0
    res = evaluate(fam, net, fam.thetas(a.test, 2), threads=a.threads, log=lambda m: print(m, flush=True))
    net.close()
    out = {"grid": a.grid, "n": fam.n, "m": fam.mi, "separate_rows": a.separate_rows,
           "train_solves": int(len(d["thetas"])), "label_seconds": float(d["seconds"]),
           "train_iterations_mean": float(d["iters"].mean()), "train_epochs": a.epochs, "train_seconds": t_train,
           "final_train_loss": float(loss), "test_members": a.test,
           "cold_iterations_mean": float(res["cold_iter"].mean()), "warm_iterations_mean": float(res["warm_iter"].mean()),
           "iteration_ratio": float(res["cold_iter"].sum() / max(res["warm_iter"].sum(), 1)),
           "cold_seconds": res["cold_seconds"], "warm_seconds": res["warm_seconds"],
           "fallbacks": int(res["fallback"].sum()), "max_rel_solution_difference": float(res["dx"].max()),
           "max_kkt_residual_warm": float(res["kkt"].max()), "predicted_active_mean": float(res["pred_size"].mean()),
           "true_active_mean": float(res["true_size"].mean()), "correctly_predicted_mean": float(res["hits"].mean()),
           "cold_iter": res["cold_iter"].tolist(), "warm_iter": res["warm_iter"].tolist()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    import torch  # noqa: F401  (the HIP runtime of torch first, as in every entry point)
    _main()
