#!/usr/bin/env python
"""Benchmark: KKT solves/s at 1e-10 backward error, plain FP64 LBL on one MI355X.

Workload (default): K = [H A'; A 0], H = 3-D 7-point Laplacian on a 64^3
grid + 1e-3 I (n = 262,144), A = 16,384 x n with 8 N(0,1) entries per row at
uniform random columns, explicit zero (2,2) diagonal, 4 right-hand sides
U[-1,1], one seeded generator, this engine's nested-dissection ordering.

One step ("one KKT solve") = numeric factorization (from values already in
HBM) + solve of the 4 right-hand sides + GMRES refinement until every
column's backward error ||b-Kx||_1/(||K||_inf ||x||_1 + ||b||_1) <= 1e-10.
The symbolic analysis is done once, outside the timed region.

Prints one JSON line (rank 0).  Usage:
    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config kkt64]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "KKT solves/s at 1e-10 residual: mixed/batched/learned vs plain FP64 on GPU"
CONFIGS = {  # name: (grid, constraint rows, nnz per row)
    "kkt96": (96, 65536, 8),
    "kkt64": (64, 16384, 8),
    "kkt48": (48, 6912, 8),
    "kkt32": (32, 2048, 8),
    "kkt16": (16, 256, 8),
}
FP64_MFMA_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def schur_flops(sup, pi, big=None):
    """Algorithmic flops of the Schur-complement updates: sum_s m_s (m_s + 1) w_s
    over all supernodes, or over those with m_s >= 256 (big=True: the 128-row
    tile kernel) / m_s < 256 (big=False: the 64-row tile kernel)."""
    w = np.diff(sup).astype(np.float64)
    r = np.diff(pi).astype(np.float64)
    m = r - w
    sel = np.ones_like(m, dtype=bool) if big is None else ((m >= 256) if big else (m < 256))
    return float(np.sum((m * (m + 1.0) * w)[sel]))


def trailing_flops(sup, small_w=64):
    """Flops of the diagonal-block updates of wide supernodes (w > small_w): ~w^3/3 each."""
    w = np.diff(sup).astype(np.float64)
    w = w[w > small_w]
    return float(np.sum(w ** 3 / 3.0))


def committed_traffic(kernel):
    """HBM bytes per launch of `kernel` at the 64^3 KKT, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/kkt64_hbm_traffic.txt,
    a line 'traffic_bytes_per_launch <kernel> <bytes>'), or None."""
    path = os.path.join(ROOT, "profiles", "kkt64_hbm_traffic.txt")
    try:
        with open(path) as f:
            for ln in f:
                fld = ln.split()
                if len(fld) == 3 and fld[0] == "traffic_bytes_per_launch" and fld[1] == kernel:
                    return float(fld[2])
    except OSError:
        pass
    return None


def cpu_info():
    """(model name, usable cores) of the host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cores = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        cores = min(cores, int(omp))
    return model, max(cores, 1)


def cpu_run(grid, m, reps, threads):
    """The BLAS-3 CPU restatement (oracle/cpu_lbl.c) on the KKT construction at
    grid^3: symbolic analysis once, then one warm-up and `reps` timed runs of
    factor and of the 4-RHS solve with refinement (x += M^-1 (b - K x)) to a
    backward error <= 1e-10; medians of the timed runs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from nasoq_amd import lbl
    from nasoq_amd import problems as P
    if O.cpu_lib() is None:
        return None
    n, Ap, Ai, Ax, B = P.kkt_config(grid, m, 8, seed=0, nrhs=4)
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = 1e-9
    s.cost_param = threads
    t = time.perf_counter()
    s.symbolic_analysis()
    t_sym = time.perf_counter() - t
    fl, _ = s.flops()
    s_xsize = s.sym_scalar("xsize")
    args = (n, Ap, Ai, Ax, 1e-9, s.sym_ints("Perm"), s.sym_ints("super"), s.sym_longs("pi"), s.sym_ints("s"),
            s.sym_longs("p"), s.sym_ints("sParent"), threads)
    tf, tsv, worst, rounds = [], [], 0.0, []
    for rep in range(reps + 1):
        t0 = time.perf_counter()
        C = O.CpuFactor(*args)
        t1 = time.perf_counter()
        X = C.solve(B)
        k = 0
        while True:
            e = max(O.backward_error(n, Ap, Ai, Ax, X[:, c], B[:, c]) for c in range(4))
            if e <= 1e-10 or k >= 20:
                break
            R = np.stack([B[:, c] - O.spmv(n, Ap, Ai, Ax, X[:, c]) for c in range(4)], axis=1)
            X = X + C.solve(R)
            k += 1
        t2 = time.perf_counter()
        if rep > 0:  # the first run is the warm-up
            tf.append(t1 - t0)
            tsv.append(t2 - t1)
            worst = max(worst, e)
            rounds.append(k)
        del C
    s.close()
    return {"grid": grid, "N": n, "sum_colcount_sq": fl, "xsize": float(s_xsize), "symbolic_s": t_sym, "factor_s_median": float(np.median(tf)),
            "solve_refine_s_median": float(np.median(tsv)), "factor_s_runs": tf, "solve_refine_s_runs": tsv,
            "worst_backward_error": worst, "refine_rounds": rounds, "blas": O.cpu_lib().cpu_lbl_blas_name().decode()}


def cpu_baseline(config, full_fl, full_xsize, grid=48, reps=5):
    """The reference's CPU numeric path restated on BLAS-3 (oracle/cpu_lbl.c:
    supernodal left-looking LDL^T with DGEMM updates, DSYTRF, DTRSM, tree
    levels in parallel) on all cores of the host, on a bounded sample of the
    same workload (the KKT construction at grid^3, constraint rows scaled
    with the volume), one warm-up and the median of `reps` runs, scaled to
    the full configuration phase by phase: the factor by the ratio of
    sum c_j^2 (its flops), the solves by the ratio of the factor's stored
    entries (their bytes).  A CPU restatement, not the reference binary
    (METIS and OpenBLAS are not in the image)."""
    model, cores = cpu_info()
    m = int(round(CONFIGS[config][1] * (grid / CONFIGS[config][0]) ** 3))
    r = cpu_run(grid, m, reps, cores)
    if r is None:
        return {"value": None, "unit": "KKT solves/s", "cores": cores, "kind": "port",
                "sample": "no BLAS library found for the CPU restatement"}
    t_step = (r["factor_s_median"] * (full_fl / r["sum_colcount_sq"]) +
              r["solve_refine_s_median"] * (full_xsize / r["xsize"]))
    return {
        "value": 1.0 / t_step,
        "unit": "KKT solves/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"CPU restatement, not the reference binary: BLAS-3 supernodal LDL^T (DGEMM / DSYTRF / DTRSM, "
                   f"tree levels in parallel) on {cores} threads of {model}, {r['blas'][:60]}; KKT construction at "
                   f"grid {grid}^3 (N={r['N']}, sum c^2={r['sum_colcount_sq']:.3e}): symbolic {r['symbolic_s']:.2f} s "
                   f"once, factor {r['factor_s_median']:.2f} s and 4-RHS solve+refine to 1e-10 "
                   f"{r['solve_refine_s_median']:.2f} s (medians of {reps} after a warm-up, worst bwd "
                   f"{r['worst_backward_error']:.1e}); scaled to the full configuration phase by phase, the factor by "
                   f"sum c^2 {r['sum_colcount_sq']:.3e} -> {full_fl:.3e}, the solves by stored entries "
                   f"{r['xsize']:.3e} -> {full_xsize:.3e}"),
        "measured": r,
        "cpu_model": model,
    }


KERNEL_FAMILIES = ["bk_small", "bk_panel", "bk_panel_group", "trailing", "offdiag", "schur", "scale_ld", "schur_big",
                   "fwd_tri", "fwd_gemv", "fwd_chain", "bwd_chain", "bwd_gemv", "bwd_tri", "solve_small",
                   "solve_batch", "spmv"]


def kernel_time(lib, h, name, steps):
    """(ms per step, launches per step) of a kernel family since timing was enabled."""
    tot, cnt = ctypes.c_double(), ctypes.c_int64()
    lib.lblg_kernel_time(h, name.encode(), ctypes.byref(tot), ctypes.byref(cnt))
    return tot.value / steps, cnt.value // max(steps, 1)


def kernel_breakdown(lib, h, step):
    """Per-family kernel time of one step run with every launch timed; timing is off afterwards."""
    from nasoq_amd import _lib
    _lib.check(lib.lblg_set_kernel_timing(h, 1), "timing")
    step()
    kt = {name: kernel_time(lib, h, name, 1) for name in KERNEL_FAMILIES}
    _lib.check(lib.lblg_set_kernel_timing(h, 0), "timing")
    return kt


def mixed_leg(s, lib, step, bwd, dX, x64, args, torch):
    """The same inputs with the FP32 factor (FP32 MFMA updates, FP32 sweeps)
    and FP64 refinement (residual, GMRES, solution update) to the same 1e-10
    backward error; forward error per right-hand side against the FP64
    solution of the plain path (itself refined to 1e-10)."""
    from nasoq_amd import _lib
    s.precision = 1
    s.numerical_factorization()  # re-creates the value buffers in FP32
    for _ in range(args.warmup):
        step()
    _lib.check(lib.lblg_synchronize(s.h), "sync")
    torch.cuda.synchronize()
    worst, rounds = 0.0, []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        worst = max(worst, float(bwd.max()))
        rounds.append(int(lib.lblg_num_ref_iter(s.h)))
    _lib.check(lib.lblg_synchronize(s.h), "sync")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    assert worst <= 1e-10, f"FP32-factor backward error {worst} above 1e-10"
    x32 = dX.cpu().numpy()
    fwd = [float(np.linalg.norm(x32[c] - x64[c]) / max(np.linalg.norm(x64[c]), 1e-300)) for c in range(x64.shape[0])]
    kt = {k: v[0] for k, v in kernel_breakdown(lib, s.h, step).items()}  # one extra step, outside the timing
    print(f"[bench] mixed FP32 factor: {dt / args.steps * 1e3:.1f} ms/step rounds={rounds} worst={worst:.2e} "
          f"fwd={max(fwd):.2e}", file=sys.stderr, flush=True)
    s.precision = 0
    return {"variant": "FP32 factor (FP32 MFMA updates, FP32 sweeps) + FP64 GMRES refinement",
            "value": args.steps / dt, "unit": "KKT solves/s", "ms_per_step": dt / args.steps * 1e3,
            "dtype": "f32 factor / f64 refinement", "worst_backward_error": worst, "refine_rounds": rounds,
            "forward_error_vs_fp64_per_rhs": fwd, "kernel_ms_per_step": kt}


def precision_leg(args, grid=96, ill=False):
    """The mixed-precision configuration: the KKT of a grid^3 Laplacian (+1e-3 I)
    with grid^3 / 13.5 constraint rows of 8 N(0,1) entries (96^3: 65,536
    rows), 4 right-hand sides, factored in FP64, FP32 and BF16 (FP32 storage,
    three-term BF16 MFMA updates) on the same inputs and refined in FP64 to a
    backward error <= 1e-10 on K; ill=True scales H's diagonal by 10^U[-3,3].
    Per precision: first factor untimed, one untimed step, then
    args.mixed_steps timed steps; the forward error of every right-hand side
    against the FP64 solution."""
    from nasoq_amd import _lib, lbl
    from nasoq_amd import problems as P
    import torch

    m = CONFIGS["kkt96"][1] if grid == 96 else int(round(65536 * (grid / 96.0) ** 3))
    t = time.perf_counter()
    n, Ap, Ai, Ax, B = P.kkt_config(grid, m, 8, shift=1e-3, seed=0, nrhs=4, ill=ill)
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = 1e-9
    s.cost_param = 16
    s.target_bwd_err = 1e-10
    s.max_refine_rounds = 20
    assert s.symbolic_analysis() == 1
    fl, _ = s.flops()
    t_sym = time.perf_counter() - t
    print(f"[bench] precision leg {grid}^3{' ill' if ill else ''}: N={n}, generated and analysed in {t_sym:.1f}s, "
          f"xsize={s.sym_scalar('xsize'):.3e} sum c^2={fl:.3e}", file=sys.stderr, flush=True)
    lib = s.lib
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda")
    dX = torch.zeros_like(dB)
    torch.cuda.synchronize()
    bwd = np.zeros(4)

    def step():
        _lib.check(lib.lblg_refactor_resident(s.h), "refactor")
        _lib.check(lib.lblg_solve_device(s.h, 4, ctypes.c_void_p(dB.data_ptr()), n,
                                         ctypes.c_void_p(dX.data_ptr()), n, _lib.dptr(bwd)), "solve")

    # the ill-conditioned variant's H is indefinite: the FP32 / BF16
    # preconditioners take GMRES(30) cycles (at most 12) there, one timed step
    steps = 1 if ill else args.mixed_steps
    res, x64 = {}, None
    for name, prec in (("fp64", 0), ("fp32", 1), ("bf16", 2)):
        s.precision = prec
        s.max_inner_iter = 30 if (ill and prec != 0) else 2
        s.max_refine_rounds = 12 if ill else 20
        s.numerical_factorization()
        if not ill:
            step()
        _lib.check(lib.lblg_synchronize(s.h), "sync")
        torch.cuda.synchronize()
        worst, rounds = 0.0, []
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
            worst = max(worst, float(bwd.max()))
            rounds.append(int(lib.lblg_num_ref_iter(s.h)))
        _lib.check(lib.lblg_synchronize(s.h), "sync")
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        reached = worst <= 1e-10
        assert reached or ill, f"{name} backward error {worst} above 1e-10"
        X = dX.cpu().numpy().copy()
        if x64 is None:
            x64 = X
        fwd = [float(np.linalg.norm(X[c] - x64[c]) / max(np.linalg.norm(x64[c]), 1e-300)) for c in range(4)]
        res[name] = {"value": 1.0 / dt if reached else None, "unit": "KKT solves/s", "ms_per_step": dt * 1e3,
                     "reached_1e-10": reached, "worst_backward_error": worst, "refine_rounds": rounds,
                     "gmres_restart": int(s.max_inner_iter), "inertia": list(s.inertia()),
                     "forward_error_vs_fp64_per_rhs": fwd}
        print(f"[bench] {grid}^3{' ill' if ill else ''} {name}: {dt * 1e3:.0f} ms/step rounds={rounds} "
              f"worst={worst:.1e} fwd={max(fwd):.1e} inertia={res[name]['inertia']}", file=sys.stderr, flush=True)
    s.close()
    del dB, dX
    torch.cuda.empty_cache()
    base = res["fp64"]["value"]
    sp = lambda k: res[k]["value"] / base if res[k]["value"] else None
    out = {"workload": f"KKT of 3-D Laplacian {grid}^3 + 1e-3 I{' with diagonal x 10^U[-3,3]' if ill else ''}, "
                       f"{m} constraint rows of 8 N(0,1), 4 RHS, refined to 1e-10 on K",
           "N": n, "sum_colcount_sq": fl, "steps": steps,
           "bf16_form": "FP32 storage of S K S (diagonal scaling), BF16 MFMA operands hi+lo (3 terms), FP32 accumulation",
           **res, "speedup_fp32_vs_fp64": sp("fp32"), "speedup_bf16_vs_fp64": sp("bf16")}
    return out


def bench_batch(args):
    """--config batch1024: the batched leg alone, as its own JSON line."""
    print(json.dumps(batch_leg(args)[1]))


def batch_leg(args):
    """Batched form: the KKT systems [H A'; A 0] of a batch of QPs of uneven
    size (2-D Laplacian r x c, r, c ~ U{16..160}, + 1e-2 I; 1..64 equality
    rows of 4 N(0,1) entries) factored and solved to a 1e-10 backward error
    (checked per system) in one block-diagonal handle, FP64 and with the FP32
    factor, against the plain path solving the same systems one by one."""
    from nasoq_amd import _lib, lbl
    from nasoq_amd import problems as P
    import torch

    t = time.perf_counter()
    inst = P.batch_qp_kkts(args.batch, seed=0)
    sizes = np.array([x[0] for x in inst])
    rhs = np.concatenate([x[4] for x in inst])
    print(f"[bench] batch of {len(inst)} KKTs, N={sizes.sum()} (n {sizes.min()}..{sizes.max()}) generated in "
          f"{time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
    b = lbl.BatchSolverSettings([(n, Ap, Ai, Ax) for n, Ap, Ai, Ax, _ in inst])
    b.reg_diag = 1e-9
    b.target_bwd_err = 1e-10  # the worst system's backward error (per-system check on the device)
    b.max_refine_rounds = 20
    t = time.perf_counter()
    assert b.symbolic_analysis() == 1
    fl, _ = b.flops()
    print(f"[bench] batch symbolic {time.perf_counter() - t:.1f}s nsuper={len(b.sym_ints('super')) - 1} "
          f"xsize={b.sym_scalar('xsize'):.3e} sum c^2={fl:.3e}", file=sys.stderr, flush=True)
    lib = b.lib
    N = b.n
    dB = torch.from_numpy(rhs.copy()).to("cuda")
    dX = torch.zeros_like(dB)
    bwd = np.zeros(1)

    schedules = {}
    kts = {}

    def run(prec):
        b.precision = prec
        b.numerical_factorization()
        schedules[prec] = b.tuning()
        if args.autotune:
            t = time.perf_counter()
            schedules[prec] = b.autotune(nrhs=1, reps=1)
            print(f"[bench] batched autotune (precision {prec}) {time.perf_counter() - t:.1f}s: {schedules[prec]}",
                  file=sys.stderr, flush=True)

        def step():
            _lib.check(lib.lblg_refactor_resident(b.h), "refactor")
            _lib.check(lib.lblg_solve_device(b.h, 1, ctypes.c_void_p(dB.data_ptr()), N,
                                             ctypes.c_void_p(dX.data_ptr()), N, _lib.dptr(bwd)), "solve")
        for _ in range(args.warmup):
            step()
        _lib.check(lib.lblg_synchronize(b.h), "sync")
        torch.cuda.synchronize()
        rounds = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
            rounds.append(int(lib.lblg_num_ref_iter(b.h)))
        _lib.check(lib.lblg_synchronize(b.h), "sync")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        kts[prec] = kernel_breakdown(lib, b.h, step)
        X = dX.cpu().numpy()
        err = b.backward_errors(rhs, X)[:, 0]
        assert err.max() <= 1e-10, f"batched (precision {prec}) worst per-system backward error {err.max()}"
        print(f"[bench] batched precision={prec}: {dt / args.steps * 1e3:.1f} ms per batch, rounds={rounds}, "
              f"worst per-system bwd {err.max():.2e}", file=sys.stderr, flush=True)
        return dt / args.steps, err.max(), rounds, X

    t64, e64, r64, X64 = run(0)
    t32, e32, r32, X32 = run(1)
    fwd32 = float(max(np.linalg.norm(X32[o:o + n] - X64[o:o + n]) / max(np.linalg.norm(X64[o:o + n]), 1e-300)
                      for o, n in zip(b.offsets[:-1], sizes)))
    b.close()

    # the plain path on the same systems, one handle after another (analysis and
    # first factor untimed, as in the batched legs)
    t_one, worst_one = 0.0, 0.0
    for k, (n, Ap, Ai, Ax, bk) in enumerate(inst):
        s = lbl.SolverSettings(n, Ap, Ai, Ax)
        s.reg_diag = 1e-9
        s.target_bwd_err = 1e-10
        s.max_refine_rounds = 20
        assert s.symbolic_analysis() == 1
        s.numerical_factorization()
        o = int(np.sum(sizes[:k]))
        pb = ctypes.c_void_p(dB.data_ptr() + 8 * o)
        px = ctypes.c_void_p(dX.data_ptr() + 8 * o)
        one = np.zeros(1)
        _lib.check(s.lib.lblg_solve_device(s.h, 1, pb, n, px, n, _lib.dptr(one)), "solve")  # warm
        _lib.check(s.lib.lblg_synchronize(s.h), "sync")
        t0 = time.perf_counter()
        _lib.check(s.lib.lblg_refactor_resident(s.h), "refactor")
        _lib.check(s.lib.lblg_solve_device(s.h, 1, pb, n, px, n, _lib.dptr(one)), "solve")
        _lib.check(s.lib.lblg_synchronize(s.h), "sync")
        t_one += time.perf_counter() - t0
        worst_one = max(worst_one, float(one[0]))
        s.close()
    print(f"[bench] one by one (plain FP64): {t_one * 1e3:.1f} ms for the batch, worst bwd {worst_one:.2e}",
          file=sys.stderr, flush=True)
    nb = len(inst)
    out = {
        "metric": METRIC,
        "value": nb / t64,
        "unit": "KKT solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": f"batched FP64 LBL factor+solve+refine of {nb} QP KKT systems [H A'; A 0], "
                               "H = 2-D Laplacian r x c (r, c ~ U{16..160}) + 1e-2 I, 1..64 equality rows",
                   "systems": nb, "N_total": int(sizes.sum()), "n_min": int(sizes.min()), "n_max": int(sizes.max()),
                   "nrhs": 1, "backward_error": 1e-10, "parallelism": "single GPU", "sum_colcount_sq": fl},
        "worst_backward_error": float(e64),
        "refine_rounds": r64,
        "batched_fp32_factor": {"value": nb / t32, "ms_per_step": t32 * 1e3, "worst_backward_error": float(e32),
                                "refine_rounds": r32, "forward_error_vs_fp64_max": fwd32},
        "one_by_one_plain_fp64": {"value": nb / t_one, "ms_for_batch": t_one * 1e3,
                                  "worst_backward_error": worst_one},
        "schedule": dict(schedules[0], autotuned=bool(args.autotune)),
        "kernel_ms_per_step": {k: v[0] for k, v in kts[0].items()},
        "kernel_launches_per_step": {k: v[1] for k, v in kts[0].items()},
        "kernel_ms_note": "one extra untimed FP64 step after the timed ones with every launch bracketed by HIP "
                          "events (families of the factor and the sweeps; the rest of a step is assembly, "
                          "refinement and launch gaps)",
        "speedup_batched_vs_one_by_one": t_one / t64,
        "speedup_batched_fp32_vs_one_by_one": t_one / t32,
    }
    summary = {"workload": out["config"]["workload"], "systems": nb, "N_total": int(sizes.sum()),
               "value": nb / t64, "unit": "KKT solves/s", "ms_per_batch": t64 * 1e3,
               "worst_backward_error": float(e64), "fp32_factor_value": nb / t32,
               "fp32_factor_worst_backward_error": float(e32), "one_by_one_plain_fp64_value": nb / t_one,
               "speedup_vs_one_by_one": t_one / t64}
    return summary, out


def qp_leg(args):
    """The plain NASOQ configuration on the GPU: min 1/2 x'Hx + q'x, Ax = b,
    x >= -0.5 with H the 40 x 40 Laplacian + 1e-2 I and 100 equality rows
    (eps 1e-8); one cold solve after an untimed one (analysis and device
    setup), wall time and active-set iterations."""
    from nasoq_amd import problems as P
    from nasoq_amd import qp
    import scipy.sparse as sp

    H, q, A, b, C, d = P.qp_config()
    m = qp.Nasoq(sp.tril(H, format="csc"), q, A, b, C, d)
    m.eps_abs = 1e-8
    m.variant = qp.FIXED
    assert m.solve() == qp.OPTIMAL
    f0 = int(m.factorizations)
    t0 = time.perf_counter()
    st = m.solve()
    dt = time.perf_counter() - t0
    assert st == qp.OPTIMAL
    out = {"workload": "QP: H = 2-D Laplacian 40x40 + 1e-2 I (n=1600), 100 equality rows of 5 N(0,1), "
                       "x >= -0.5, eps 1e-8", "status": "optimal", "seconds": dt, "active_set_iterations": m.num_iter,
           "ms_per_iteration": dt / max(m.num_iter, 1) * 1e3, "active_constraints": len(m.active_set),
           "objective": m.objective, "kkt_factorizations": int(m.factorizations) - f0}
    # the same solve with the KKT schedule picked by the autotuner
    t0 = time.perf_counter()
    sched = m.autotune()
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    st = m.solve()
    dt2 = time.perf_counter() - t0
    assert st == qp.OPTIMAL
    out["autotuned"] = {"seconds": dt2, "ms_per_iteration": dt2 / max(m.num_iter, 1) * 1e3, "tune_seconds": t_tune,
                        "schedule": sched}
    m.close()
    print(f"[bench] qp configuration: {dt:.2f} s, {out['active_set_iterations']} iterations "
          f"({out['ms_per_iteration']:.2f} ms each), {out['active_constraints']} active; autotuned schedule "
          f"{dt2:.2f} s (tuning {t_tune:.1f} s)", file=sys.stderr, flush=True)
    return out


def qp_batch_leg(args):
    """The batched QP configuration: `args.qp_batch` independent box-constrained
    QPs (H = 2-D Laplacian r x c + 1e-2 I with r, c ~ U{16..160}, 1..64
    equality rows of 4 N(0,1), |x| <= 1), each solved to NASOQ's own KKT test
    by its dual active-set loop with the KKT factor in FP32 (FP64
    refinement) and, beside it, in FP64.  Every iteration's KKT factor is
    the SoMod update along the tree paths of the edited rows; `args.qp_threads`
    host threads run the problems concurrently, one engine handle and HIP
    stream each.  Throughput is whole-batch wall time."""
    from nasoq_amd import problems as P
    from nasoq_amd import qp
    probs = P.box_qp_batch(args.qp_batch, seed=0)
    sizes = [p[0].shape[0] for p in probs]
    settings = {"eps_abs": 1e-6, "variant": qp.FIXED}
    qp.solve_many(probs[:2], settings, threads=2, keep=())  # warm-up: library load, first launches
    out = {"workload": f"{len(probs)} box-constrained QPs, H = 2-D Laplacian r x c + 1e-2 I (r, c ~ U{{16..160}}), "
                       f"1..64 equality rows of 4 N(0,1), |x| <= 1, eps 1e-6",
           "qps": len(probs), "n_total": int(sum(sizes)), "n_range": [int(min(sizes)), int(max(sizes))],
           "threads": args.qp_threads}
    objs = {}
    legs = ((1, "fp32_factor"), (0, "fp64")) if args.qp_batch_fp64 else ((1, "fp32_factor"),)
    for prec, name in legs:
        t0 = time.perf_counter()
        log = (lambda m: print(f"[bench] qp batch {name}: {m}", file=sys.stderr, flush=True)) if len(probs) > 64 else None
        res = qp.solve_many(probs, dict(settings, precision=prec), threads=args.qp_threads, keep=(), log=log)
        dt = time.perf_counter() - t0
        fac = sum(r["factorizations"] for r in res)
        full = sum(r["factorizations"] * r["supernodes"] for r in res)
        out[name] = {"seconds": dt, "qps_per_s": len(res) / dt, "optimal": sum(r["status"] == qp.OPTIMAL for r in res),
                     "active_set_iterations": int(sum(r["num_iter"] for r in res)),
                     "kkt_solves": int(sum(r["kkt_solves"] for r in res)),
                     "kkt_solves_per_s": sum(r["kkt_solves"] for r in res) / dt, "kkt_factorizations": int(fac),
                     "refactored_fraction": sum(r["refactored_supernodes"] for r in res) / max(full, 1)}
        objs[name] = np.array([r["objective"] for r in res])
        print(f"[bench] qp batch {name}: {len(res)} QPs in {dt:.1f} s, {out[name]['optimal']} optimal, "
              f"{out[name]['active_set_iterations']} iterations, {out[name]['kkt_solves_per_s']:.0f} KKT solves/s, "
              f"refactored fraction {out[name]['refactored_fraction']:.3f}", file=sys.stderr, flush=True)
    if "fp64" in objs:
        rel = np.abs(objs["fp32_factor"] - objs["fp64"]) / np.maximum(1.0, np.abs(objs["fp64"]))
        out["objective_rel_diff_fp32_vs_fp64_max"] = float(rel.max())
    return out


def learned_leg(args):
    """Learned warm start on a parametric QP family: min 1/2 x'Hx + (U theta)'x
    s.t. Bx <= d with H the 5-point Laplacian of a g x g grid + 1e-2 I,
    round(0.4 n) inequality rows of 6 N(0,1) entries, d ~ U[0, 0.5] and
    U ~ N(0, 1/n) of n x 32.  Plain NASOQ (every KKT factorization on the GPU)
    labels the training parameters with its active sets; the MLP
    32 -> 1024 -> 1024 -> m (BF16 MFMA, FP32 accumulation) is trained on them
    by the library's HIP kernels; held-out members are solved cold and warm
    (prediction -> Nasoq.solve_predicted, fallback to a cold start unless
    NASOQ's check says OPTIMAL) and the two solutions compared."""
    from nasoq_amd import learned

    fam = learned.ParametricQpFamily(grid=args.learned_grid, seed=0)
    fam.separate_rows = args.learned_separate_rows
    S = fam.solver()
    train = fam.thetas(args.learned_train, 1)
    log = lambda m: print(f"[bench] learned: {m}", file=sys.stderr, flush=True)
    log(f"family n={fam.n} m={fam.mi}, labelling {len(train)} members with plain NASOQ")
    Y, iters, t_label = learned.generate(fam, train, solver=S, log=log, threads=args.learned_threads)
    net = learned.ActiveSetPredictor(fam.n_theta, fam.mi, hidden=(1024, 1024), seed=0, max_batch=128)
    t0 = time.perf_counter()
    loss = net.fit(train, Y, epochs=args.learned_epochs, batch=64, lr=1e-3)
    t_train = time.perf_counter() - t0
    test = fam.thetas(args.learned_test, 2)
    t0 = time.perf_counter()
    net.predict_proba(test)
    t_pred = time.perf_counter() - t0
    res = learned.evaluate(fam, net, test, solver=S, threads=args.learned_threads)
    net.close()
    S.close()
    ci, wi = res["cold_iter"].sum(), res["warm_iter"].sum()
    out = {"workload": f"parametric QP family, H = 2-D Laplacian {args.learned_grid}x{args.learned_grid} + 1e-2 I "
                       f"(n={fam.n}), {fam.mi} inequality rows of 6 N(0,1), d ~ U[0,0.5], q = U theta (32 params)",
           "separate_rows": args.learned_separate_rows,
           "mlp": f"32 -> 1024 -> 1024 -> {fam.mi}, BF16 weights, FP32 accumulation",
           "train_solves": len(train), "label_seconds": t_label, "label_threads": args.learned_threads,
           "train_iterations_mean": float(iters.mean()), "train_epochs": args.learned_epochs,
           "train_seconds": t_train, "final_train_loss": loss,
           "train_samples_per_s": len(train) * args.learned_epochs / t_train,
           "predict_ms_per_member": t_pred / len(test) * 1e3,
           "test_members": len(test), "cold_iterations_mean": float(res["cold_iter"].mean()),
           "warm_iterations_mean": float(res["warm_iter"].mean()), "iteration_ratio": float(ci / max(wi, 1)),
           "cold_factorizations_mean": float(res["cold_fact"].mean()),
           "warm_factorizations_mean": float(res["warm_fact"].mean()),
           "cold_solves_per_s": len(test) / res["cold_seconds"], "warm_solves_per_s": len(test) / res["warm_seconds"],
           "fallbacks": int(res["fallback"].sum()), "max_rel_solution_difference": float(res["dx"].max()),
           "max_kkt_residual_warm": float(res["kkt"].max()),
           "predicted_active_mean": float(res["pred_size"].mean()), "true_active_mean": float(res["true_size"].mean()),
           "correctly_predicted_mean": float(res["hits"].mean())}
    log(f"iterations cold {out['cold_iterations_mean']:.1f} warm {out['warm_iterations_mean']:.1f} "
        f"(x{out['iteration_ratio']:.2f}), solves/s cold {out['cold_solves_per_s']:.1f} warm "
        f"{out['warm_solves_per_s']:.1f}, fallbacks {out['fallbacks']}, max dx {out['max_rel_solution_difference']:.1e}")
    assert out["max_kkt_residual_warm"] <= 1e-6 and out["max_rel_solution_difference"] <= 1e-6
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="kkt64", choices=sorted(CONFIGS) + ["batch1024", "learned", "cpu64"])
    ap.add_argument("--batch", type=int, default=1024, help="instances of the batch1024 configuration")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32-factor + FP64-refinement leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched leg (1024 QP KKT systems)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start leg")
    ap.add_argument("--no-qp", action="store_true", help="skip the plain NASOQ QP configuration leg")
    ap.add_argument("--qp-batch", type=int, default=16, help="box-constrained QPs of the batched QP leg (0: skip)")
    ap.add_argument("--qp-batch-fp64", action="store_true", help="the batched QP leg also with the FP64 KKT factor")
    ap.add_argument("--qp-threads", type=int, default=16, help="host threads of the batched QP leg")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the 96^3 precision leg (FP64 / FP32 / BF16 and the ill-conditioned variant)")
    ap.add_argument("--mixed-steps", type=int, default=2, help="timed steps per precision of the 96^3 leg")
    ap.add_argument("--precision-grid", type=int, default=96, help="grid of the precision leg")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the device schedule with lblg_autotune before timing (plain and batched legs)")
    ap.add_argument("--learned-grid", type=int, default=12, help="grid side of the learned leg's QP family")
    ap.add_argument("--learned-train", type=int, default=400, help="training members (plain NASOQ solves)")
    ap.add_argument("--learned-test", type=int, default=40, help="held-out members solved cold and warm")
    ap.add_argument("--learned-epochs", type=int, default=60)
    ap.add_argument("--learned-separate-rows", type=int, default=1, choices=[0, 1],
                    help="the family's KKT inequality rows as one-column supernodes (1, the reference's) or not (0)")
    ap.add_argument("--learned-threads", type=int, default=1,
                    help="host threads labelling and evaluating members concurrently (one solver each)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (reference traffic for HBM counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    from nasoq_amd import _lib, lbl
    from nasoq_amd import problems as P
    import torch

    if args.config == "batch1024":
        bench_batch(args)
        return
    if args.config == "learned":
        print(json.dumps(learned_leg(args)))
        return
    if args.config == "cpu64":  # the CPU restatement at the full 64^3 configuration
        model, cores = cpu_info()
        r = cpu_run(64, CONFIGS["kkt64"][1], 5, cores)
        print(json.dumps({"workload": "CPU restatement (not the reference binary) of the 64^3 KKT configuration",
                          "cpu_model": model, "cores": cores, **(r or {}),
                          "kkt_solves_per_s": 1.0 / (r["factor_s_median"] + r["solve_refine_s_median"]) if r else None}))
        return
    grid, m, k = CONFIGS[args.config]
    t = time.perf_counter()
    n, Ap, Ai, Ax, B = P.kkt_config(grid, m, k, shift=1e-3, seed=0, nrhs=4)
    t_gen = time.perf_counter() - t
    nrhs = B.shape[1]
    print(f"[bench] config={args.config} N={n} nnz_lower={Ap[-1]} generated in {t_gen:.1f}s", file=sys.stderr,
          flush=True)
    s = lbl.SolverSettings(n, Ap, Ai, Ax)
    s.reg_diag = 1e-9
    s.cost_param = 16
    s.target_bwd_err = 1e-10
    s.max_refine_rounds = 20
    t = time.perf_counter()
    assert s.symbolic_analysis() == 1
    t_sym = time.perf_counter() - t
    fl, lnz = s.flops()
    sup, pi = s.sym_ints("super"), s.sym_longs("pi")
    print(f"[bench] symbolic {t_sym:.1f}s nsuper={len(sup) - 1} maxw={s.sym_scalar('max_sup_wid')} "
          f"xsize={s.sym_scalar('xsize'):.3e} sum c^2={fl:.3e}", file=sys.stderr, flush=True)
    t = time.perf_counter()
    s.numerical_factorization()
    t_first = time.perf_counter() - t
    print(f"[bench] first factor (incl. device setup) {t_first:.2f}s", file=sys.stderr, flush=True)
    schedule = s.tuning()
    if args.autotune:
        t = time.perf_counter()
        schedule = s.autotune(nrhs=nrhs, reps=2)
        print(f"[bench] autotune {time.perf_counter() - t:.1f}s: {schedule}", file=sys.stderr, flush=True)

    lib = s.lib
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda")  # nrhs x n rows = column-major n x nrhs
    dX = torch.zeros_like(dB)
    torch.cuda.synchronize()
    bwd = np.zeros(nrhs)

    def step():
        _lib.check(lib.lblg_refactor_resident(s.h), "refactor")
        _lib.check(lib.lblg_solve_device(s.h, nrhs, ctypes.c_void_p(dB.data_ptr()), n,
                                         ctypes.c_void_p(dX.data_ptr()), n, _lib.dptr(bwd)), "solve")

    for w in range(args.warmup):
        step()
        print(f"[bench] warmup {w} bwd={bwd.max():.2e} rounds={lib.lblg_num_ref_iter(s.h)}", file=sys.stderr,
              flush=True)
    # per-kernel breakdown: one more untimed step with every launch bracketed
    # by HIP events; it also names the dominant MFMA update family
    kt = kernel_breakdown(lib, s.h, step)
    print("[bench] per-step kernel ms (extra step, every launch timed): " +
          ", ".join(f"{k}={v[0]:.2f}({v[1]})" for k, v in kt.items()), file=sys.stderr, flush=True)
    dom = "schur_big" if kt["schur_big"][0] >= kt["trailing"][0] else "trailing"
    _lib.check(lib.lblg_synchronize(s.h), "sync")
    # the timed steps carry HIP events around the dominant family only, so
    # the launch rate is the untimed one
    _lib.check(lib.lblg_set_kernel_timing_only(s.h, dom.encode()), "timing")
    torch.cuda.synchronize()
    worst = 0.0
    rounds = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        worst = max(worst, float(bwd.max()))
        rounds.append(int(lib.lblg_num_ref_iter(s.h)))
    _lib.check(lib.lblg_synchronize(s.h), "sync")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    assert worst <= 1e-10, f"backward error {worst} above 1e-10"
    ms_step = dt / args.steps * 1e3
    rk = {dom: kernel_time(lib, s.h, dom, args.steps)}
    _lib.check(lib.lblg_set_kernel_timing(s.h, 0), "timing")

    # roofline of the dominant MFMA kernel (128-row-tile Schur-complement
    # updates, or the trailing updates of the wide diagonal blocks)
    sf = schur_flops(sup, pi, big=True)
    tf = trailing_flops(sup)
    work = sf if dom == "schur_big" else tf
    launches = max(rk[dom][1], 1)
    avg_ms = rk[dom][0] / launches
    achieved = (work / launches) / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    roof = {"bound": "mfma", "kernel": "k_" + dom, "achieved": achieved, "peak": FP64_MFMA_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_MFMA_PEAK_TFLOPS,
            "traffic": committed_traffic("k_" + dom) if args.config == "kkt64" else None,
            "avg_launch_ms": avg_ms, "launches_per_step": launches,
            "algorithmic_flops_per_launch": work / launches}

    out = {
        "metric": METRIC,
        "value": args.steps / dt,
        "unit": "KKT solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": f"plain FP64 LBL factor+solve+refine, KKT of 3-D Laplacian {grid}^3 "
                               f"+ {m} random constraint rows", "grid": grid, "n": grid ** 3, "m": m, "N": n,
                   "nnz_lower": int(Ap[-1]), "nrhs": nrhs, "backward_error": 1e-10,
                   "parallelism": "single GPU", "sum_colcount_sq": fl, "xsize": s.sym_scalar("xsize")},
        "worst_backward_error": worst,
        "refine_rounds": rounds,
        "kernel_ms_per_step": {k: v[0] for k, v in kt.items()},
        "kernel_ms_note": "one extra untimed step before the timed ones with every launch bracketed by HIP "
                          "events; the timed steps carry events around the roofline kernel's launches only",
        "roofline": roof,
        "schedule": dict(schedule, autotuned=bool(args.autotune)),
        "graph_captures_launches": list(s.graph_stats()),
    }
    # HBM rooflines of the solve sweeps and the SpMV, from the event-timed step:
    # a sweep reads the stored trapezoids (8 B per value of L, 4 B per row
    # label) once for all right-hand sides and moves the n x nrhs blocks
    # about three times; the SpMV reads the full CSR (8 + 4 B per entry, row
    # pointers), gathers x and writes y (n x nrhs each) through the permutation
    xs, ss = float(s.sym_scalar("xsize")), float(s.sym_scalar("ssize"))
    sweep_bytes = 8.0 * xs + 4.0 * ss + 3 * 8.0 * n * nrhs
    nnzf = 2.0 * Ap[-1] - n
    spmv_bytes = 12.0 * nnzf + 4.0 * (n + 1) + 2 * 8.0 * n * nrhs + 8.0 * n
    fwd_ms = sum(kt[k][0] for k in ("fwd_tri", "fwd_gemv", "fwd_chain"))
    bwd_ms = sum(kt[k][0] for k in ("bwd_chain", "bwd_gemv", "bwd_tri"))
    sp_ms = kt["spmv"][0] / max(kt["spmv"][1], 1)
    hb = lambda b, ms: {"achieved": b / (ms * 1e-3) / 1e9 if ms > 0 else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": (b / (ms * 1e-3) / 1e9) / HBM_PEAK_GBS if ms > 0 else None,
                        "algorithmic_bytes": b, "ms": ms}
    out["hbm_rooflines"] = {"forward_sweep": hb(sweep_bytes, fwd_ms), "backward_sweep": hb(sweep_bytes, bwd_ms),
                            "spmv": hb(spmv_bytes, sp_ms),
                            "note": "one solve of the event-timed step; sweep = all kernels of that direction"}
    dval = np.zeros(2 * n)
    _lib.check(lib.lblg_get_factor(s.h, None, _lib.dptr(dval), None, None), "get_factor")
    out["pivots_2x2"] = int(np.count_nonzero(dval[n:]))
    out["inertia"] = list(s.inertia())
    x64 = dX.cpu().numpy().copy()
    a2, a3 = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(lib.lblg_pivot_stats(s.h, ctypes.byref(a2), ctypes.byref(a3)), "pivot_stats")
    p1, p2 = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(lib.lblg_panel_stats(s.h, ctypes.byref(p1), ctypes.byref(p2)), "panel_stats")
    out["wide_panel_pivoting"] = {"second_column_steps": a2.value, "interchanges": a3.value,
                                  "panels": p1.value, "panels_with_pivot_search": p2.value}
    if not args.no_mixed:
        out["mixed"] = mixed_leg(s, lib, step, bwd, dX, x64, args, torch)
    s.close()
    del dB, dX
    torch.cuda.empty_cache()
    if not args.no_precision and args.config == "kkt64":
        print("[bench] precision leg ...", file=sys.stderr, flush=True)
        out["precision_96"] = precision_leg(args, grid=args.precision_grid)
        out["precision_96_ill"] = precision_leg(args, grid=args.precision_grid, ill=True)
    if not args.no_batched and args.config == "kkt64":
        print("[bench] batched leg ...", file=sys.stderr, flush=True)
        out["batched"] = batch_leg(args)[0]
    if not args.no_learned and args.config == "kkt64":
        print("[bench] learned warm-start leg ...", file=sys.stderr, flush=True)
        out["learned"] = learned_leg(args)
    if not args.no_qp and args.config == "kkt64":
        print("[bench] plain NASOQ QP configuration ...", file=sys.stderr, flush=True)
        out["qp_configuration"] = qp_leg(args)
    if args.qp_batch > 0 and args.config == "kkt64":
        print("[bench] batched box-constrained QPs ...", file=sys.stderr, flush=True)
        out["qp_batch"] = qp_batch_leg(args)
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.config, fl, xs)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
