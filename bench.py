"""Benchmark of the Cholesky path on one MI355X.

A step is one FP64 pposv (ppotrf + ppotrs, uplo L; --uplo U for the upper
triangle) of the synthetic SPD system
A = G G^T / n + 0.1 I (G: n x n seeded normals, seed 1), B: n x nrhs normals,
at n = 32768, nrhs = 64, with A and B already resident in HBM.  Before each
step A and B are restored from pristine device copies (outside the timed
call); the timed region is the pposv call itself, which returns synchronized.
value = (n^3/3 + 2 n^2 nrhs) / time, in TFLOP/s.

roofline: the dominant kernel is the MFMA trailing update; its launches are
bracketed by HIP events on the handle's stream during the timed steps, and
achieved = (sum of its algorithmic flops, m (m + 1) kb per launch) / (sum of
its durations).  solve_phase: the two triangular-solve sweeps of the step,
their algorithmic HBM bytes (the lower factor read once per sweep, n^2 * 8)
over their event-timed duration.

mixed: the same n and nrhs with seed 2 and delta in {4e-2, 4e-4, 4e-6}
(condition numbers about 1e2, 1e4, 1e6), solved by pposv_mixed (FP32 factor
on the matrix cores, FP64 residual and refinement on the GPU, tol 1e-14) with
each trailing-update mode (FP32 MFMA, bf16x3 and bf16x6 split bfloat16
products, fp16x3 split half-precision products of row-scaled panels, all
with FP32 accumulation) and, on the same matrix, by the FP64 pposv;
per delta and mode: refinement iterations,
time of each, their ratio, the backward error max_j ||b_j - A x_j||_inf /
(||A||_inf ||x_j||_inf) (FP64, on the GPU) and the forward error
||x - x_fp64||_inf / ||x_fp64||_inf against the FP64 pposv solution.

bottom_right: ppotrf_bottom_right (uplo U, A = U U^T) of the headline matrix
beside ppotrf of the same matrix, best of two calls each after a sizing call:
the cost of the bottom-right entry's two anti-transpose passes.

batched: 4096 SPD systems with orders uniform in [32, 512] (seed 3), A_i =
G_i G_i^T / n_i + 0.5 I, one right-hand side each, FP64 and FP32: the fused
potrf + potrs + logdet (tuned schedule: the persistent kernel, with the
members above a tuned order factored on the blocked path whose trailing
updates are gemmVbatched launches; replayed from a captured hipGraph)
against a loop of per-matrix pposv + diag_logdet calls on the same GPU;
inputs restored from a device copy before every timed launch.

gp: 16 Gaussian-process likelihoods (n = 4096 points uniform in [0, 1]^3,
seed 4, RBF l = 0.2, s = 1, jitter 1e-2, y ~ N(0, 1)): forward Cholesky,
solve and log-determinant plus the hand-written backward through
torch.autograd, timed over the 16 problems (K formed by torch outside the
timed region), once with the fused likelihood Function (backward: the
closed form (K^-1 - alpha alpha^T) / 2 with K^-1 from potri) and once
composed from the Cholesky, solve and log-determinant Functions and their
adjoints; flops counted as each algorithm runs them.  Problem 0's
dLoss/dK and dLoss/dl are compared with the FP64 CPU closed form.

cpu_baseline: the reference's CPU path at a 1 x 1 grid is one LAPACK dpotrf_
and two BLAS dtrsm_ calls (src/Cholesky/potrf.cpp:46-49,
src/BLAS/trsm.cpp:135-141); it is timed here as OpenBLAS dpotrf + dpotrs
(scipy's bundled OpenBLAS) on the host cores: once on the headline
workload (n = 32768, nrhs = 64, about 10 s), and on the reference's own
CPU-handle case (n = 2048, nrhs = 16, A = G G^T / n + I, seed 0) with the
backward error of its solution and of this library's GPU pposv on the same
inputs.  The host's CPU model, logical CPU count and OpenBLAS thread count
are recorded.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TFLOP/s, % peak: FP64 vs mixed-prec pposv n=32768 at 1e-14 bwd err, 1 GPU"
FP64_PEAK = 78.6  # TFLOP/s, MI355X dense FP64 matrix
FP32_PEAK = 157.3  # TFLOP/s, MI355X dense FP32 matrix
FACTOR_NAMES = {0: "fp32", 1: "bf16x3", 2: "bf16x6", 3: "fp16x3"}  # pposv_mixed factor argument


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """CPU model and logical CPU count of this host (/proc/cpuinfo, os.cpu_count())."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count()}


def openblas_threads():
    """(threads, version) of the OpenBLAS that scipy's LAPACK wrappers call."""
    try:
        from threadpoolctl import threadpool_info

        for info in threadpool_info():
            if info.get("internal_api") == "openblas":
                return info.get("num_threads"), info.get("version")
    except Exception:
        pass
    return None, None


def host_spd(h, dd, n, seed, delta):
    """A = G G^T / n + delta I of the project's seed convention, formed on the GPU, as a host Fortran array."""
    import ctypes

    import numpy as np

    a_dev = dd.DistributedMatrix(h, n, n, 512, 512, np.float64)
    g_dev = dd.DistributedMatrix(h, n, n, 512, 512, np.float64)
    dd.check_status(dd.lib.ddla_d_spd_generate(h.ptr, n, seed, delta, ctypes.c_void_p(a_dev.data_ptr), n,
                                               ctypes.c_void_p(g_dev.data_ptr)))
    g_dev.free()
    a = a_dev.to_global_matrix()
    a_dev.free()
    return a


def host_normals(h, dd, seed, stream, rows, cols):
    import ctypes

    import numpy as np

    b_dev = dd.DistributedMatrix(h, rows, cols, 512, 512, np.float64)
    dd.check_status(dd.lib.ddla_c_fill_normal(h.ptr, seed, stream, 0, rows * cols, ctypes.c_void_p(b_dev.data_ptr)))
    b = b_dev.to_global_matrix()
    b_dev.free()
    return b


def cpu_baseline(h, dd, n_cpu, nrhs, reps):
    """The reference's CPU path (LAPACK dpotrf + dpotrs in OpenBLAS) on the host cores: the headline
    workload (FP64 pposv, n_cpu = 32768 by default, nrhs 64, A = G G^T / n + 0.1 I, seed 1), and the
    reference's own CPU-handle case (n = 2048, nrhs 16, A = G G^T / n + I, seed 0) with its backward
    error ||A X - B|| / (||A|| ||X|| n eps) beside that of this library's GPU pposv on the same inputs."""
    import numpy as np
    from scipy.linalg import lapack

    threads, version = openblas_threads()

    def timed(a, b, reps):
        times, x = [], None
        for _ in range(reps):
            work = np.array(a, order="F", copy=True)
            t0 = time.perf_counter()
            c, info = lapack.dpotrf(work, lower=1, clean=0, overwrite_a=1)
            x, info2 = lapack.dpotrs(c, b, lower=1)
            times.append(time.perf_counter() - t0)
            assert info == 0 and info2 == 0
        return sorted(times)[len(times) // 2], x

    a = host_spd(h, dd, n_cpu, 1, 0.1)
    b = np.asfortranarray(host_normals(h, dd, 1, 1, n_cpu, nrhs))
    t, _ = timed(a, b, reps)
    del a
    flops = n_cpu ** 3 / 3 + 2.0 * n_cpu ** 2 * nrhs

    # the reference's CPU-handle case: GPU pposv and the host LAPACK solve on the same inputs
    n1, r1 = 2048, 16
    a1 = host_spd(h, dd, n1, 0, 1.0)
    b1 = np.asfortranarray(host_normals(h, dd, 0, 1, n1, r1))
    t1, x_cpu = timed(a1, b1, 5)
    A1 = dd.DistributedMatrix.from_global_matrix(h, a1, mb=128, nb=128)
    B1 = dd.DistributedMatrix.from_global_matrix(h, b1, mb=128, nb=128)
    X1, info = dd.pposv(A1, B1, "L")
    assert info == 0
    x_gpu = X1.to_global_matrix()
    eps = float(np.finfo(np.float64).eps)

    def berr(x):
        r = a1 @ x - b1
        return float(np.linalg.norm(r, np.inf) / (np.linalg.norm(a1, np.inf) * np.linalg.norm(x, np.inf) * n1 * eps))

    flops1 = n1 ** 3 / 3 + 2.0 * n1 ** 2 * r1
    cpu = host_cpu()
    return {
        "value": round(flops / t / 1e12, 4),
        "unit": "TFLOP/s",
        "cores": threads or cpu["nproc"],
        "kind": "reference",
        "impl": f"OpenBLAS {version or '?'} dpotrf+dpotrs (scipy build), the LAPACK calls of the reference's "
                "1x1 CPU handle",
        "sample": f"FP64 pposv n={n_cpu} nrhs={nrhs} (the headline workload: A = G G^T/n + 0.1 I, seed 1); "
                  f"median of {reps}, {t:.2f} s each",
        "ms": round(t * 1e3, 1),
        "host": {"cpu_model": cpu["model"], "nproc": cpu["nproc"], "openblas_threads": threads},
        "cpu_handle_case": {
            "workload": f"pposv n={n1} nrhs={r1}, A = G G^T/n + I, seed 0 (the reference's CPU-handle case)",
            "cpu_ms": round(t1 * 1e3, 3),
            "cpu_gflops": round(flops1 / t1 / 1e9, 1),
            "cpu_backward_error_n_eps": berr(x_cpu),
            "gpu_backward_error_n_eps": berr(x_gpu),
            "max_rel_diff_gpu_vs_cpu": float(np.abs(x_gpu - x_cpu).max() / np.abs(x_cpu).max()),
        },
    }


def backward_error(a0, b0, x):
    """max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf); a0 symmetric (n, n), b0 and x (nrhs, n)."""
    r = b0 - x @ a0  # row j: (A x_j)^T, A symmetric
    anorm = a0.abs().sum(1).max()
    return (r.abs().amax(1) / (anorm * x.abs().amax(1))).max().item()


def mixed_leg(h, dd, args, a0, a, b0, b, g, A, B, times_fp64):
    import ctypes

    import torch

    n, nrhs = args.n, args.nrhs
    rows = []
    for delta in args.deltas:
        dd.check_status(dd.lib.ddla_d_spd_generate(h.ptr, n, 2, delta, ctypes.c_void_p(a0.data_ptr()), n,
                                                   ctypes.c_void_p(g.data_ptr())))
        dd.check_status(dd.lib.ddla_c_fill_normal(h.ptr, 2, 1, 0, n * nrhs, ctypes.c_void_p(b0.data_ptr())))
        h.synchronize()

        def restore():
            a.copy_(a0)
            b.copy_(b0)
            torch.cuda.synchronize()

        restore()
        t0 = time.perf_counter()
        _, info = dd.pposv(A, B, "L")
        t64 = time.perf_counter() - t0
        assert info == 0, info
        x64 = b.clone()
        for factor in args.factors:
            fname = FACTOR_NAMES[factor]
            for _ in range(args.warmup):
                restore()
                dd.pposv_mixed(A, B, "L", factor=factor, tol=args.tol, max_iter=30)
            h.profile(True)
            ts, its = [], []
            for s in range(args.steps):
                restore()
                t0 = time.perf_counter()
                _, it, info = dd.pposv_mixed(A, B, "L", factor=factor, tol=args.tol, max_iter=30)
                ts.append(time.perf_counter() - t0)
                its.append(it)
                assert info == 0, info
                log(f"mixed delta={delta:g} factor={fname} step {s}: {ts[-1]:.3f} s iter={it}")
            prof = h.profile_read()
            h.profile(False)
            t = sum(ts) / len(ts)
            fwd = ((b - x64).abs().max() / x64.abs().max()).item()
            cnt, ms, fl, _ = prof["trailing_update"]
            rows.append({
                "delta": delta,
                "factor": fname,
                "iter": its[-1],
                "ms": round(t * 1e3, 3),
                "fp64_pposv_ms": round(t64 * 1e3, 3),
                "time_ratio_vs_fp64": round(t / min(t64, times_fp64), 4),
                "fp64_equiv_tflops": round((n ** 3 / 3 + 2.0 * n ** 2 * nrhs) / t / 1e12, 3),
                "backward_error": backward_error(a0, b0, b),
                "forward_error_vs_fp64": fwd,
                "trailing_update_tflops": round(fl / (ms / 1e3) / 1e12, 3) if ms > 0 else None,
                "breakdown_ms": {k: round(v[1] / args.steps, 3) for k, v in prof.items() if v[0]},
            })
    return rows


def bottom_right_leg(h, dd, args, a0, a, A, reps=2):
    """ppotrf_bottom_right (uplo U, A = U U^T) at the headline order beside
    ppotrf on the same matrix: the cost of the two anti-transpose passes."""
    import torch

    n = args.n

    def timed(fn):
        a.copy_(a0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, info = fn(A, "U")
        dt = time.perf_counter() - t0
        assert info == 0, info
        return dt

    timed(dd.ppotrf_bottom_right)  # first call sizes the scratch
    t_br = min(timed(dd.ppotrf_bottom_right) for _ in range(reps))
    t_f = min(timed(dd.ppotrf) for _ in range(reps))
    flops = n ** 3 / 3
    return {"n": n, "uplo": "U", "reps": reps, "statistic": "min",
            "ms": round(t_br * 1e3, 3), "tflops": round(flops / t_br / 1e12, 3),
            "ppotrf_ms": round(t_f * 1e3, 3), "ppotrf_tflops": round(flops / t_f / 1e12, 3),
            "overhead_pct": round(100 * (t_br / t_f - 1), 2)}


def batched_leg(h, dd, args):
    import ctypes

    import numpy as np

    rng = np.random.RandomState(3)
    sizes = rng.randint(32, 513, size=args.batched)
    flops = float(np.sum(sizes.astype(np.float64) ** 3 / 3 + 2.0 * sizes.astype(np.float64) ** 2))
    out = {"members": int(args.batched), "orders": "uniform integers in [32, 512], seed 3", "nrhs": 1,
           "flops_convention": "sum n^3/3 + 2 n^2"}
    for dtype, name in ((np.float64, "fp64"), (np.float32, "fp32")):
        vb = dd.VBatch(h, sizes, nrhs=1, dtype=dtype)
        vb.generate(3, 0.5)
        es = np.dtype(dtype).itemsize
        a_bytes = int((vb.ld.astype(np.int64) * sizes).sum()) * es
        b_bytes = int(vb.ld.astype(np.int64).sum()) * es
        pristine_a = dd.batched._Device(h, a_bytes)
        pristine_b = dd.batched._Device(h, b_bytes)
        cp = lambda dst, src, n: dd.check_status(dd.lib.ddla_c_memcpy(h.ptr, ctypes.c_void_p(dst.ptr.value),
                                                                      ctypes.c_void_p(src.ptr.value), n))
        cp(pristine_a, vb.A, a_bytes)
        cp(pristine_b, vb.B, b_bytes)

        def restore():
            cp(vb.A, pristine_a, a_bytes)
            cp(vb.B, pristine_b, b_bytes)
            h.synchronize()

        t0 = time.perf_counter()
        occ, ordered = dd.pposv_vbatched_tune(vb, "L")
        tune_s = time.perf_counter() - t0
        sched = dd.vbatched_schedule(h, dtype)
        g = dd.BatchedGraph(vb, "L")
        times = []
        for rep in range(args.warmup + args.steps):
            restore()
            t0 = time.perf_counter()
            g.launch()
            h.synchronize()
            if rep >= args.warmup:
                times.append(time.perf_counter() - t0)
        t = sorted(times)[len(times) // 2]
        info = vb.get_info()
        logdet = vb.get_logdet().copy()
        x_batched = [vb.get_rhs(i) for i in range(0, args.batched, max(1, args.batched // 64))]
        g.destroy()
        # loop of per-matrix plain calls: pposv (factor + solve, synchronized) and diag_logdet
        restore()
        pfx = "d" if dtype == np.float64 else "s"
        pposv = getattr(dd.lib, f"ddla_{pfx}_pposv")
        logdet_fn = getattr(dd.lib, f"ddla_{pfx}_diag_logdet")
        res = np.zeros(2, dtype=dtype)
        loop_logdet = np.zeros(args.batched)
        info_loop = ctypes.c_int()
        t0 = time.perf_counter()
        for i in range(args.batched):
            n = int(sizes[i])
            a_ptr = ctypes.c_void_p(vb.A.ptr.value + int(vb.a_off[i]) * es)
            b_ptr = ctypes.c_void_p(vb.B.ptr.value + int(vb.b_off[i]) * es)
            da = dd.descinit(h, n, n, 64, 64, 0, 0, int(vb.ld[i]))
            db = dd.descinit(h, n, 1, 64, 64, 0, 0, int(vb.ld[i]))
            dd.check_status(pposv(h.ptr, b"L", b"L", b"N", n, 1, a_ptr, 1, 1,
                                  da.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), b_ptr, 1, 1,
                                  db.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(info_loop)))
            assert info_loop.value == 0
            dd.check_status(logdet_fn(h.ptr, n, a_ptr, int(vb.ld[i]), ctypes.c_void_p(res.ctypes.data)))
            loop_logdet[i] = 2.0 * float(res[0])
        h.synchronize()
        t_loop = time.perf_counter() - t0
        x_loop = [vb.get_rhs(i) for i in range(0, args.batched, max(1, args.batched // 64))]
        dx = max(float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max() /
                       max(1e-300, np.abs(b.astype(np.float64)).max())) for a, b in zip(x_batched, x_loop))
        out[name] = {
            "ms": round(t * 1e3, 3),
            "tflops": round(flops / t / 1e12, 3),
            "hbm_gb_s_a_read_write": round(2.0 * a_bytes / t / 1e9, 1),
            # fractions of both bounds: the matrices read and written once, against
            # HBM, and the flops against the matrix cores' dense peak of the type
            "roofline": {"mfma_peak_tflops": FP64_PEAK if dtype == np.float64 else FP32_PEAK,
                         "mfma_frac": round(flops / t / 1e12 / (FP64_PEAK if dtype == np.float64 else FP32_PEAK), 4),
                         "hbm_peak_gb_s": 8000.0, "hbm_frac": round(2.0 * a_bytes / t / 1e9 / 8000.0, 4),
                         "bound": "neither: each member's chain of dependent diagonal-block steps runs on one "
                                  "workgroup (latency)"},
            "loop_of_plain_calls_ms": round(t_loop * 1e3, 1),
            "speedup_vs_loop": round(t_loop / t, 1),
            "all_info_zero": bool((info == 0).all()),
            "max_logdet_diff_vs_loop": float(np.abs(logdet - loop_logdet).max()),
            "max_rel_solution_diff_vs_loop": dx,
            "tuned": {"workgroups_per_cu": occ, "largest_first": ordered,
                      "kernel": "left-looking" if sched["left_looking"] else "right-looking",
                      "blocked_path_above_order": None if sched["left_looking"] else 64 * sched["split_blocks"],
                      "blocked_step_columns": None if sched["left_looking"] else sched["step_columns"],
                      "tune_s": round(tune_s, 2)},
        }
        log(f"batched {name}: {t * 1e3:.3f} ms ({flops / t / 1e12:.2f} TF/s), loop {t_loop * 1e3:.0f} ms")
        del vb, pristine_a, pristine_b
    return out


def gp_leg(h, dd, args):
    import numpy as np
    import torch

    from libddla_amd import autograd as ag

    n, problems = args.gp_n, args.gp_problems
    rng = np.random.RandomState(4)
    data = [(rng.uniform(size=(n, 3)), rng.standard_normal(n)) for _ in range(problems)]
    dev = torch.device("cuda:0")

    def kernel(x, l):
        d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
        return torch.exp(-d2 / (2 * l * l)) + 1e-2 * torch.eye(n, dtype=torch.float64, device=dev), d2

    prepared = []
    for x, y in data:
        xt = torch.tensor(x, device=dev)
        prepared.append((xt, torch.tensor(y, device=dev)))

    def timed(loss_fn, handles=None):
        """Seconds for forward + backward of every problem (K formed outside), and problem 0's gradients;
        with handles, the problems run concurrently on them (one host thread and stream each)."""
        times, grads = [], []
        for rep in range(args.warmup + 1):
            ks = []
            for xt, yt in prepared:
                l = torch.tensor(0.2, dtype=torch.float64, device=dev, requires_grad=True)
                K, d2 = kernel(xt, l)
                Kd = K.detach().requires_grad_()
                ks.append((l, K, Kd, yt))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if handles:
                ag.backward_concurrently([(Kd, yt) for l, K, Kd, yt in ks], loss_fn, handles)
            else:
                for l, K, Kd, yt in ks:
                    loss = loss_fn(Kd, yt, h)
                    loss.backward()
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            if rep == args.warmup:
                times.append(t)
                for l, K, Kd, yt in ks[:1]:
                    K.backward(Kd.grad)  # chain dLoss/dK into dLoss/dl (outside the timed region)
                    grads.append((Kd.grad.detach().cpu().numpy(), float(l.grad), K.detach().cpu().numpy()))
        return times[0], grads[0]

    def timed_batched():
        """The same problems stacked (K: problems x n x n) through the batched Function, whose every launch
        serves all problems; seconds for forward + backward (K formed outside), problem 0's gradients."""
        result = None
        for rep in range(args.warmup + 1):
            l = torch.tensor(0.2, dtype=torch.float64, device=dev, requires_grad=True)
            parts = [kernel(xt, l) for xt, _ in prepared]
            K = torch.stack([p[0] for p in parts])
            Kd = K.detach().requires_grad_()
            Y = torch.stack([yt for _, yt in prepared])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            losses = ag.gp_negative_log_likelihood_batched(Kd, Y, h)
            losses.sum().backward()
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            if rep == args.warmup:
                K.backward(Kd.grad)  # dLoss/dl (summed over the problems) outside the timed region
                g0 = Kd.grad[0].detach().cpu().numpy()
                # problem 0's dl alone: its dK contracted with dK_0/dl
                dk0 = (parts[0][0].detach() - 1e-2 * torch.eye(n, dtype=torch.float64, device=dev)) * parts[0][1] / 0.2 ** 3
                dl0 = float((Kd.grad[0] * dk0).sum())
                result = (t, g0, dl0, parts[0][0].detach().cpu().numpy())
        return result

    t_b, kbar_b, dl_b, k = timed_batched()
    workers = [dd.Handle(dd.Backend.GPU) for _ in range(max(0, args.gp_workers))] if args.gp_workers > 1 else None
    t, (kbar, dl, _) = timed(ag.gp_negative_log_likelihood, workers)
    t_seq, _ = timed(ag.gp_negative_log_likelihood) if workers else (t, None)
    if args.no_gp_composed:
        t_comp, kbar_c, dl_c = None, None, None
    else:
        t_comp, (kbar_c, dl_c, _) = timed(ag.gp_negative_log_likelihood_composed, workers)
    for w in workers or []:
        w.destroy()

    def lower_product_flops(m):
        # tril(X^T Y) of lower-triangular X, Y in four k blocks [k0, k1) of growing order, masked
        total = 0.0
        for q in range(4):
            k0 = q * (m // 4) // 64 * 64
            k1 = m if q == 3 else (q + 1) * (m // 4) // 64 * 64
            if k1 > k0:
                total += k1 * (k1 + 1.0) * (k1 - k0)
        return total

    def lower_inverse_flops(m):
        # L^-1 by merging inverted 64 x 64 diagonal blocks: X21 = -X22 (L21 X11), two products per merge
        # (a merge of order-h blocks in 4 / 2 / 1 triangle-aware slices from h = 1024 / 256 / 64:
        # 2 h^3 (1 + 1 / slices))
        f, hh = 0.0, 64
        while hh < m:
            full = m // (2 * hh)
            o = full * 2 * hh
            slices = 4 if hh >= 1024 else 2 if hh >= 256 else 1
            f += full * 2.0 * hh ** 3 * (1.0 + 1.0 / slices)
            if o + hh < m:
                n2 = m - o - hh
                f += 2.0 * n2 * hh * hh + 2.0 * n2 * hh * n2
            hh *= 2
        return f

    # batched: forward potrf n^3/3 + L^-1 (lower_inverse_flops) + z = L^-1 y (2n^2); backward tril(X^T X)
    # (lower_product_flops) + a = X^T z (2n^2) + the gradient (n^2)
    flops_b = problems * (n ** 3 / 3 + lower_inverse_flops(n) + 2.0 * n * n + lower_product_flops(n) + 3.0 * n * n)
    # forward: potrf n^3/3, two one-column sweeps 2n^2; backward (fused): K^-1 from the factor --
    # tril(L^-1) by recursive inversion (n^3/2) and tril(L^-T L^-1) (lower_product_flops) -- and the
    # rank-1 update (2n^2)
    flops = problems * (n ** 3 / 3 + 4.0 * n * n + n ** 3 / 2 + lower_product_flops(n) + 2.0 * n * n)
    # composed: backward tril(L^T Lbar), one n-column solve (n^3), the lower-triangle solve (n^3/3) and
    # the solve adjoint (8 n^2)
    flops_comp = problems * (n ** 3 / 3 + 4.0 * n * n + lower_product_flops(n) + n ** 3 + n ** 3 / 3 + 8.0 * n * n)
    y = data[0][1]
    kinv = np.linalg.inv(k)
    alpha = kinv @ y
    kbar_ref = 0.5 * (kinv - np.outer(alpha, alpha))
    d2 = ((data[0][0][:, None, :] - data[0][0][None, :, :]) ** 2).sum(-1)
    dl_ref = float((kbar_ref * (k - 1e-2 * np.eye(n)) * d2 / 0.2 ** 3).sum())
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return {"problems": problems, "n": n, "ms": round(t_b * 1e3, 2), "tflops": round(flops_b / t_b / 1e12, 3),
            "fp64_peak_frac": round(flops_b / t_b / 1e12 / FP64_PEAK, 4),
            "flops_convention": "per problem: forward n^3/3 (potrf) + L^-1 by merging inverted 64 x 64 diagonal "
                                "blocks in triangle-aware slices (0.42 n^3 at n = 4096) + 2n^2; backward "
                                "tril(L^-T L^-1) in four k blocks (about n^2/2 * n) + 3n^2",
            "dK_rel_err_vs_fp64_closed_form": rel(kbar_b, kbar_ref),
            "dl_rel_err_vs_fp64_closed_form": abs(dl_b - dl_ref) / abs(dl_ref),
            "timed": "forward + backward of all problems through torch.autograd, K given: the batched Function "
                     "(libddla_amd.autograd.gp_negative_log_likelihood_batched, K stacked problems x n x n): "
                     "every launch -- each panel and trailing update of the factorization, each merge level of "
                     "L^-1, the products -- serves all problems",
            "per_problem": {
                "ms": round(t * 1e3, 2), "tflops": round(flops / t / 1e12, 3),
                "handles": max(1, args.gp_workers), "sequential_ms": round(t_seq * 1e3, 2),
                "flops_convention": "forward n^3/3 + 4n^2; backward K^-1 from the factor (recursive triangular "
                                    "inverse n^3/2 + tril(L^-T L^-1) in four k blocks, about n^2/2 * n) + 2n^2",
                "dK_rel_err_vs_fp64_closed_form": rel(kbar, kbar_ref),
                "dl_rel_err_vs_fp64_closed_form": abs(dl - dl_ref) / abs(dl_ref),
                "timed": "one fused Function per problem (ppotrf, ppotrs, diag_logdet; backward (K^-1 - alpha "
                         "alpha^T) / 2 with K^-1 from potri), the problems concurrently on `handles` handles "
                         "(libddla_amd.autograd.backward_concurrently: one host thread and stream each); "
                         "sequential_ms is one problem after another on one handle"},
            "composed": None if t_comp is None else {
                "ms": round(t_comp * 1e3, 2), "tflops": round(flops_comp / t_comp / 1e12, 3),
                "dK_rel_err_vs_fp64_closed_form": rel(kbar_c, kbar_ref),
                "dl_rel_err_vs_fp64_closed_form": abs(dl_c - dl_ref) / abs(dl_ref),
                "timed": "the same through the separate Cholesky, solve and log-determinant Functions and their "
                         "adjoints (tril(L^T Lbar), n-column and lower-triangle solves)"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--cpu-n", type=int, default=32768)
    ap.add_argument("--cpu-reps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--deltas", type=float, nargs="+", default=[4e-2, 4e-4, 4e-6])
    ap.add_argument("--tol", type=float, default=1e-14)
    ap.add_argument("--factors", type=int, nargs="+", default=[0, 1, 2, 3],
                    help="pposv_mixed factor modes: 0 FP32, 1 bf16x3, 2 bf16x6, 3 fp16x3")
    ap.add_argument("--batched", type=int, default=4096, help="members of the batched leg (0 skips it)")
    ap.add_argument("--gp-problems", type=int, default=16, help="problems of the GP-likelihood leg (0 skips it)")
    ap.add_argument("--gp-n", type=int, default=4096)
    ap.add_argument("--gp-workers", type=int, default=4,
                    help="handles (host threads and streams) the GP problems run on concurrently")
    ap.add_argument("--no-gp-composed", action="store_true", help="time only the fused GP likelihood Function")
    ap.add_argument("--no-bottom-right", action="store_true", help="skip the ppotrf_bottom_right leg")
    ap.add_argument("--uplo", choices=["L", "U"], default="L", help="triangle of the headline pposv")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs one process on one GPU; --gpus must be 1 "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import ctypes

    import numpy as np
    import torch

    import libddla_amd as dd

    n, nrhs = args.n, args.nrhs
    h = dd.Handle(dd.Backend.GPU)
    dev = torch.device("cuda:0")
    log(f"generating A (n={n}) and B (nrhs={nrhs}) on the GPU")
    a0 = torch.empty((n, n), dtype=torch.float64, device=dev)
    g = torch.empty((n, n), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    dd.check_status(dd.lib.ddla_d_spd_generate(h.ptr, n, 1, 0.1, ctypes.c_void_p(a0.data_ptr()), n,
                                               ctypes.c_void_p(g.data_ptr())))
    h.synchronize()
    b0 = torch.empty((nrhs, n), dtype=torch.float64, device=dev)  # column-major n x nrhs
    dd.check_status(dd.lib.ddla_c_fill_normal(h.ptr, 1, 1, 0, n * nrhs, ctypes.c_void_p(b0.data_ptr())))
    h.synchronize()
    a = torch.empty_like(a0)
    b = torch.empty_like(b0)
    A = dd.DistributedMatrix.from_torch(h, a, n, n, 512, 512)
    B = dd.DistributedMatrix.from_torch(h, b, n, nrhs, 512, 512)

    def restore():
        a.copy_(a0)
        b.copy_(b0)
        torch.cuda.synchronize()

    for w in range(args.warmup):
        restore()
        t0 = time.perf_counter()
        _, info = dd.pposv(A, B, args.uplo)
        log(f"warmup {w}: {time.perf_counter() - t0:.3f} s info={info}")
    h.profile(True)
    times = []
    for s in range(args.steps):
        restore()
        t0 = time.perf_counter()
        _, info = dd.pposv(A, B, args.uplo)
        times.append(time.perf_counter() - t0)
        assert info == 0, info
        log(f"step {s}: {times[-1]:.3f} s")
    prof = h.profile_read()
    h.profile(False)

    # backward error of the last solve against the pristine A, B (FP64, on the GPU)
    bwd = backward_error(a0, b0, b)

    t = sum(times) / len(times)
    flops = n ** 3 / 3 + 2.0 * n ** 2 * nrhs
    value = flops / t / 1e12
    cnt, ms, fl, _ = prof["trailing_update"]
    achieved = fl / (ms / 1e3) / 1e12 if ms > 0 else 0.0
    breakdown = {k: {"launches": v[0] // max(1, args.steps), "ms_per_step": round(v[1] / args.steps, 3)}
                 for k, v in prof.items() if v[0]}
    # the two triangular-solve sweeps read the lower factor once each: n^2 * 8 bytes in all
    solve_phase = None
    if prof.get("solve", (0, 0))[0]:
        solve_ms = prof["solve"][1] / args.steps
        solve_phase = {"ms_per_step": round(solve_ms, 3), "algorithmic_bytes": 8.0 * n * n,
                       "hbm_gb_s": round(8.0 * n * n / (solve_ms * 1e-3) / 1e9, 1), "peak_gb_s": 8000.0,
                       "frac": round(8.0 * n * n / (solve_ms * 1e-3) / 8e12, 4),
                       "bound": "the 64-row-block dependency chain (latency), not HBM"}
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "trailing_update_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Philox4x64-10 normals, seed 1; A = G G^T/n + 0.1 I)",
        "config": {"workload": f"FP64 pposv n={n} nrhs={nrhs} uplo={args.uplo}, A = G G^T/n + 0.1 I", "n": n,
                   "nrhs": nrhs, "nb": 512, "panel_width": 768 if n >= 16384 else 512, "uplo": args.uplo},
        "pct_fp64_peak": round(100 * value / FP64_PEAK, 2),
        "backward_error": bwd,
        "roofline": {"bound": "mfma", "kernel": "trailing_update (FP64 MFMA syrk)", "achieved": round(achieved, 3),
                     "peak": FP64_PEAK, "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK, 4),
                     "traffic": traffic,
                     "traffic_source": "profiles/trailing_update_traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE "
                                       "passes over this workload, not measured in this run)",
                     "launches_per_step": cnt // max(1, args.steps)},
        "breakdown": breakdown,
        "solve_phase": solve_phase,
    }
    if not args.no_mixed:
        log(f"mixed-precision pposv, deltas {args.deltas}")
        rows = mixed_leg(h, dd, args, a0, a, b0, b, g, A, B, t)
        out["mixed"] = {
            "solver": "pposv_mixed: FP32 factorization on the matrix cores + FP64 iterative refinement, LAPACK "
                      "dsposv stopping test",
            "tol": args.tol, "seed": 2, "max_iter": 30,
            "forward_error_reference": "FP64 pposv solution of the same system on the GPU",
            "factors": {"fp32": "FP32 MFMA trailing updates",
                        "bf16x3": "bfloat16 split products x0y0 + x0y1 + x1y0, FP32 accumulation",
                        "bf16x6": "bfloat16 split products with i + j <= 2 (three planes), FP32 accumulation",
                        "fp16x3": "half-precision split products x0y0 + x0y1 + x1y0 of panel rows scaled by "
                                  "powers of two, FP32 accumulation"},
            "per_delta": rows,
            "worst_time_ratio_vs_fp64": {FACTOR_NAMES[f]: max(r["time_ratio_vs_fp64"] for r in rows
                                                              if r["factor"] == FACTOR_NAMES[f])
                                         for f in args.factors},
            "worst_backward_error": max(r["backward_error"] for r in rows),
        }
    del g
    if not args.no_bottom_right:
        log(f"bottom-right leg: ppotrf_bottom_right n={n}")
        out["bottom_right"] = bottom_right_leg(h, dd, args, a0, a, A)
    if args.batched > 0:
        log(f"batched leg: {args.batched} members")
        out["batched"] = batched_leg(h, dd, args)
    if args.gp_problems > 0:
        log(f"GP-likelihood leg: {args.gp_problems} problems, n={args.gp_n}")
        out["gp"] = gp_leg(h, dd, args)
    if not args.no_cpu:
        log(f"CPU baseline: OpenBLAS dpotrf+dpotrs at n={args.cpu_n}")
        out["cpu_baseline"] = cpu_baseline(h, dd, args.cpu_n, nrhs, args.cpu_reps)
    print(json.dumps(out), flush=True)
    h.destroy()


if __name__ == "__main__":
    main()
