"""Benchmark of the gfx950 L3 engine on one MI355X.

Workload (one step): DGEMM M=N=K=8192 with op(A)=A^T, alpha=1, beta=0,
entries N(0,1), operands resident in HBM, called through the unchanged
cblas_dgemm C ABI with device pointers.  The line also carries DTRSM
(Left/Lower/NoTrans/NonUnit, m=8192, nrhs=4096) and DSYRK (lower, n=8192,
k=2048) from the same workload family, and native FP32 SGEMM at 8192.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints ONE JSON line.  `roofline` is for the dominant kernel (the FP64 MFMA
GEMM kernel): achieved = 2*M*N*K per launch / average launch time measured
with HIP events on the library's own stream; peak = MI355X dense FP64 MFMA
(78.6 TFLOP/s).  `cpu_baseline` times the reference's own linux_avx2 build
(oracle/_ref, compiled from the reference sources) on one host core over a
bounded slice of the same DGEMM (`value`), and beside it on slices of the
DTRSM and SGEMM workloads and, with nproc concurrent single-threaded workers,
on the batched problems and a sample of the serving requests; it runs before
the process touches the GPU.
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

PEAK = {"f64": 78.6, "f32": 157.3}  # dense MFMA TFLOP/s, MI355X
METRIC = ("TFLOP/s & %MFMA peak, D/SGEMM+DTRSM n=8192; p50/p99 mixed-L3 request "
          "latency")
COL, ROW, NOTR, TR, UP, LO, NONUNIT, LEFT, RIGHT = 102, 101, 111, 112, 121, 122, 131, 141, 142
SERVE_RATE = 16000.0  # requests/s of the serving workload's Poisson arrivals


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=8192)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-extra", action="store_true")
    p.add_argument("--no-batched", action="store_true")
    p.add_argument("--no-serving", action="store_true")
    p.add_argument("--serve-requests", type=int, default=2000)
    p.add_argument("--serve-rate", type=float, default=SERVE_RATE,
                   help="Poisson arrival rate of the serving workload, requests/s")
    p.add_argument("--serve-streams", type=int, default=0,
                   help="device streams of the serving loop (0: the library default)")
    p.add_argument("--serve-slots", type=int, default=0,
                   help="launches in flight per stream (0: the library default)")
    p.add_argument("--precision-n", type=int, default=16384,
                   help="SGEMM size of the native-vs-BF16x3 comparison (0: skip)")
    p.add_argument("--cpu-slice", type=int, default=2048,
                   help="M = N of the CPU-baseline DGEMM slice (K stays 8192)")
    p.add_argument("--only", choices=["dgemm", "dtrsm", "dsyrk", "sgemm", "sgemm_bf16x3",
                                      "batched", "batched_trsm", "serving"],
                   help="profiling aid: run just this operation (3 calls) and exit")
    p.add_argument("--traffic-probe", action="store_true",
                   help="profiling aid: one 512 MiB device copy (counter "
                        "calibration) then one DGEMM, for rocprofv3 --pmc runs")
    return p.parse_args()


def timed_calls(torch, lib, fn, steps, warmup):
    """Run fn() warmup+steps times; return (wall ms per call, event ms per
    call on the library stream)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(lib.mublis_rt_stream())
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return wall, e0.elapsed_time(e1) / steps


def _ref_lib():
    """The reference's own linux_avx2 build (oracle/_ref, compiled from the
    reference sources by oracle/Makefile), or None when it is not built."""
    ref = os.path.join(ROOT, "oracle", "_ref", "libref_avx2.so")
    if not os.path.exists(ref):
        return None
    lib = ctypes.CDLL(ref)
    I, P = ctypes.c_int, ctypes.c_void_p
    for p, ct in (("s", ctypes.c_float), ("d", ctypes.c_double)):
        getattr(lib, f"cblas_{p}gemm").argtypes = [I] * 6 + [ct, P, I, P, I, ct, P, I]
        getattr(lib, f"cblas_{p}syrk").argtypes = [I] * 5 + [ct, P, I, ct, P, I]
        getattr(lib, f"cblas_{p}syr2k").argtypes = [I] * 5 + [ct, P, I, P, I, ct, P, I]
        getattr(lib, f"cblas_{p}trsm").argtypes = [I] * 7 + [ct, P, I, P, I]
    return lib


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_workers():
    """Concurrent CPU workers: the cores this process may run on, at most 16
    (a one-GPU box's share of its host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def _ref_call(lib, r, rng):
    """Operands (U[-1, 1], column-major) and a closure running request or
    problem `r` through the reference build; returns (closure, flop)."""
    op = r["op"]
    f32 = np.float32
    if op in ("sgemm", "gemm"):
        m, n, k = r["m"], r["n"], r["k"]
        a = rng.uniform(-1, 1, (k, m)).astype(f32)  # column-major m x k, ld m
        b = rng.uniform(-1, 1, (n, k)).astype(f32)
        c = rng.uniform(-1, 1, (n, m)).astype(f32)
        beta = r.get("beta", 0.5)
        return (lambda: lib.cblas_sgemm(COL, NOTR, NOTR, m, n, k, 1.0, a.ctypes.data, m,
                                        b.ctypes.data, k, beta, c.ctypes.data, m),
                2.0 * m * n * k), (a, b, c)
    if op in ("ssyrk", "ssyr2k", "syrk"):
        n, k = r["n"], r["k"]
        uplo = UP if r.get("upper") else LO
        a = rng.uniform(-1, 1, (k, n)).astype(f32)
        c = rng.uniform(-1, 1, (n, n)).astype(f32)
        beta = r.get("beta", 0.5)
        if op == "ssyr2k":
            b = rng.uniform(-1, 1, (k, n)).astype(f32)
            return (lambda: lib.cblas_ssyr2k(COL, uplo, NOTR, n, k, 1.0, a.ctypes.data, n,
                                             b.ctypes.data, n, beta, c.ctypes.data, n),
                    2.0 * n * (n + 1) * k), (a, b, c)
        return (lambda: lib.cblas_ssyrk(COL, uplo, NOTR, n, k, 1.0, a.ctypes.data, n, beta,
                                        c.ctypes.data, n),
                1.0 * n * (n + 1) * k), (a, c)
    m, n = r["m"], r["n"]  # dtrsm: Left, Lower, NoTrans, NonUnit; diagonal m
    a = rng.uniform(-1, 1, (m, m))
    np.fill_diagonal(a, float(m))
    b = rng.uniform(-1, 1, (n, m))
    return (lambda: lib.cblas_dtrsm(COL, LEFT, LO, NOTR, NONUNIT, m, n, 1.0, a.ctypes.data, m,
                                    b.ctypes.data, m),
            float(m) * m * n), (a, b)


_BARRIER = None


def _cpu_worker(items, reps, seed, out):
    """One concurrent CPU worker: builds its problems, waits for the others,
    then runs them `reps` times; puts (seconds spent in the calls, flop)."""
    lib = _ref_lib()
    rng = np.random.default_rng(seed)
    busy, flop = 0.0, 0.0
    try:
        _BARRIER.wait(timeout=300)
        for r in items:
            (fn, fl), keep = _ref_call(lib, r, rng)
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            busy += time.perf_counter() - t0
            flop += fl * reps
            del keep
    finally:
        out.put((busy, flop))


def _concurrent(items, reps, workers):
    """Aggregate throughput of `workers` processes, each running its share of
    `items` (dealt largest first onto the least loaded worker) `reps` times:
    total flop / the longest worker's time in the calls.  The workers are
    forked before this process touches the GPU."""
    import multiprocessing as mp
    global _BARRIER
    ctx = mp.get_context("fork")
    shares, load = [[] for _ in range(workers)], [0.0] * workers
    for r in sorted(items, key=lambda r: -r["flop"]):
        w = load.index(min(load))
        shares[w].append(r)
        load[w] += r["flop"]
    _BARRIER = ctx.Barrier(workers)
    out = ctx.Queue()
    procs = [ctx.Process(target=_cpu_worker, args=(shares[w], reps, 1000 + w, out))
             for w in range(workers)]
    for p in procs:
        p.start()
    res = [out.get(timeout=600) for _ in procs]
    for p in procs:
        p.join(timeout=60)
    busy = max(b for b, _ in res)
    flop = sum(f for _, f in res)
    return flop / busy / 1e12, busy


def cpu_baseline(n_k, m_slice):
    """The reference's own linux_avx2 build (-O3 -mavx2 -mfma -ffast-math) on
    the GPU box's host.  `value`: one thread, DGEMM op(A)=A^T on an m_slice x
    m_slice x n_k slice of the workload.  Beside it, one thread on slices of
    the DTRSM and SGEMM workloads, and nproc concurrent single-threaded
    workers (the reference has no threading inside a call) on the batched
    configuration's problems and on a sample of the serving requests."""
    lib = _ref_lib()
    if lib is None:
        return None
    rng = np.random.default_rng(314159)
    m = m_slice
    a = np.asfortranarray(rng.standard_normal((n_k, m)))
    b = np.asfortranarray(rng.standard_normal((n_k, m)))
    c = np.zeros((m, m), order="F")
    t0 = time.perf_counter()
    lib.cblas_dgemm(COL, TR, NOTR, m, m, n_k, 1.0, a.ctypes.data, n_k, b.ctypes.data, n_k,
                    0.0, c.ctypes.data, m)
    dt = time.perf_counter() - t0
    flops = 2.0 * m * m * n_k
    del a, b, c
    single = {}
    # DTRSM L/L/N/NonUnit m = n_k on 256 of the 4096 right-hand sides
    # (off-diagonal N(0,1)/sqrt(m), diagonal 1 + |N(0,1)|, as on the GPU)
    nrhs = 256
    at = np.tril(rng.standard_normal((n_k, n_k)) / n_k ** 0.5, -1)
    at[np.diag_indices(n_k)] = 1.0 + np.abs(rng.standard_normal(n_k))
    at = np.asfortranarray(at)
    bt = np.asfortranarray(rng.standard_normal((n_k, nrhs)))
    t0 = time.perf_counter()
    lib.cblas_dtrsm(COL, LEFT, LO, NOTR, NONUNIT, n_k, nrhs, 1.0, at.ctypes.data, n_k,
                    bt.ctypes.data, n_k)
    t = time.perf_counter() - t0
    single[f"dtrsm_{n_k}x{nrhs}"] = {"tflops": round(n_k * n_k * nrhs / t / 1e12, 5),
                                     "seconds": round(t, 2)}
    del at, bt
    # SGEMM op(A)=A^T on an m_slice x m_slice x n_k slice
    a = np.asfortranarray(rng.standard_normal((n_k, m)).astype(np.float32))
    b = np.asfortranarray(rng.standard_normal((n_k, m)).astype(np.float32))
    c = np.zeros((m, m), dtype=np.float32, order="F")
    t0 = time.perf_counter()
    lib.cblas_sgemm(COL, TR, NOTR, m, m, n_k, 1.0, a.ctypes.data, n_k, b.ctypes.data, n_k,
                    0.0, c.ctypes.data, m)
    t = time.perf_counter() - t0
    single[f"sgemm_{m}x{m}x{n_k}"] = {"tflops": round(2.0 * m * m * n_k / t / 1e12, 5),
                                      "seconds": round(t, 2)}
    del a, b, c
    concurrent = {}
    workers = _cpu_workers()
    from mublis_amd.workloads import grouped_shapes, serving_plan
    from mublis_amd import serving as S
    ms, ns, ks, _, sn, sk, _ = grouped_shapes()
    items = [dict(op="gemm", m=int(x), n=int(y), k=int(z), beta=1.0, flop=2.0 * x * y * z)
             for x, y, z in zip(ms, ns, ks)]
    items += [dict(op="syrk", n=int(x), k=int(z), beta=1.0, flop=1.0 * x * (x + 1) * z)
              for x, z in zip(sn, sk)]
    reps = 8
    tf, busy = _concurrent(items, reps, workers)
    concurrent["batched_4096sgemm_1024ssyrk"] = {
        "tflops": round(tf, 4), "workers": workers, "reps": reps,
        "seconds": round(busy, 2)}
    names = {S.SGEMM: "sgemm", S.SSYRK: "ssyrk", S.SSYR2K: "ssyr2k", S.DTRSM: "dtrsm"}
    sample, total = [], 0.0
    for r in serving_plan():
        if total >= 1.2e12:
            break
        q = dict(op=names[r["op"]], flop=r["flop"], upper=r.get("uplo") == S.UPPER,
                 m=r.get("m"), n=r["n"], k=r.get("k"))
        sample.append(q)
        total += r["flop"]
    tf, busy = _concurrent(sample, 1, workers)
    concurrent["serving_sample"] = {
        "tflops": round(tf, 4), "requests_per_s": round(len(sample) / busy, 1),
        "requests": len(sample), "of": 2000, "workers": workers, "seconds": round(busy, 2)}
    return {"value": round(flops / dt / 1e12, 5), "unit": "TFLOP/s", "cores": 1,
            "kind": "reference",
            "sample": f"DGEMM op(A)=A^T M=N={m} K={n_k} N(0,1), reference linux_avx2 "
                      f"build (-O3 -mavx2 -mfma -ffast-math), one thread, {dt:.2f} s",
            "cpu_model": _cpu_model(), "cores_available": workers,
            "single_thread": single,
            "concurrent": concurrent}


def profile_mode(args, torch, M, lib, dev, g):
    """Run one operation of the workload a few times with nothing else on the
    GPU, for kernel-trace or counter collection."""
    n = args.n
    if args.traffic_probe:
        src = torch.empty(1 << 26, dtype=torch.float64, device=dev)   # 512 MiB
        src.fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)                    # calibration: 512 MiB read + written
        torch.cuda.synchronize()
        what, calls = "dgemm", 1
    else:
        # The GEMMs run bench.py's own warm-up + timed call count, so that a
        # traced average compares with the timed figure; the rest a few calls.
        what = args.only
        calls = args.warmup + args.steps if what in ("dgemm", "sgemm") else 3
    if what == "sgemm_bf16x3":
        nn = args.precision_n or n
        A = torch.randn((nn, nn), device=dev, generator=g)
        B = torch.randn((nn, nn), device=dev, generator=g)
        C = torch.zeros((nn, nn), device=dev)
        f = lib.mublis_sgemm_bf16x3
        f.argtypes = [ctypes.c_int] * 5 + [ctypes.c_float, ctypes.c_void_p, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                                           ctypes.c_int, ctypes.c_int]
        for _ in range(calls):
            f(1, 0, nn, nn, nn, 1.0, M.ptr(A), 1, nn, M.ptr(B), 1, nn, 0.0, M.ptr(C), 1, nn)
    elif what in ("dgemm", "sgemm"):
        dt = torch.float64 if what == "dgemm" else torch.float32
        A = torch.randn((n, n), dtype=dt, device=dev, generator=g)
        B = torch.randn((n, n), dtype=dt, device=dev, generator=g)
        C = torch.zeros((n, n), dtype=dt, device=dev)
        f = lib.cblas_dgemm if what == "dgemm" else lib.cblas_sgemm
        for _ in range(calls):
            f(COL, TR, NOTR, n, n, n, 1.0, M.ptr(A), n, M.ptr(B), n, 0.0, M.ptr(C), n)
    elif what == "dtrsm":
        m, nrhs = n, n // 2
        At = torch.randn((m, m), dtype=torch.float64, device=dev, generator=g) / m ** 0.5
        d = 1.0 + torch.randn(m, dtype=torch.float64, device=dev, generator=g).abs()
        At = torch.triu(At, 1) + torch.diag(d)
        B0 = torch.randn((nrhs, m), dtype=torch.float64, device=dev, generator=g)
        for _ in range(calls):
            Bt = B0.clone()
            lib.cblas_dtrsm(COL, LEFT, LO, NOTR, NONUNIT, m, nrhs, 1.0, M.ptr(At), m,
                            M.ptr(Bt), m)
    elif what in ("batched", "batched_trsm"):
        from mublis_amd.workloads import grouped_problems, grouped_trsm_problems
        ops, keep, specs = (grouped_problems if what == "batched" else grouped_trsm_problems)(torch)
        plan = M.GroupedPlan(ops)
        for _ in range(calls):
            plan.execute(lib.mublis_rt_stream())
        lib.mublis_rt_sync(None)
        plan.close()
    elif what == "dsyrk":
        k = n // 4
        As = torch.randn((k, n), dtype=torch.float64, device=dev, generator=g)
        C = torch.zeros((n, n), dtype=torch.float64, device=dev)
        for _ in range(calls):
            lib.cblas_dsyrk(COL, LO, NOTR, n, k, 1.0, M.ptr(As), n, 0.0, M.ptr(C), n)
    torch.cuda.synchronize()
    print(f"profile mode done: {what} x{calls}", flush=True)


def sampled_gemm_error(torch, At, Bt, Ct, count=64, seed=0):
    """Normwise relative error of `count` sampled entries of C = op(A) B
    (tensors in the benchmark's convention: C(i,j) = Ct[j, i] = sum_p
    At[i, p] Bt[j, p]) against long-double sums of the same inputs."""
    n_i, n_j = At.shape[0], Bt.shape[0]
    rng = np.random.default_rng(seed)
    I = torch.from_numpy(rng.integers(0, n_i, count)).to(At.device)
    J = torch.from_numpy(rng.integers(0, n_j, count)).to(At.device)
    rows = At[I].double().cpu().numpy().astype(np.longdouble)
    cols = Bt[J].double().cpu().numpy().astype(np.longdouble)
    exact = np.sum(rows * cols, axis=1)
    got = Ct[J, I].double().cpu().numpy()
    return float(np.linalg.norm(got - exact) / np.linalg.norm(exact))


def trsm_residual(torch, At, Xt, Bt, seed=1):
    """||A (X v) - B v|| / (||A||_F ||X||_F ||v||) for a random v, in long
    double via the oracle's matrix-vector helpers (lower, no transpose).
    At, Xt, Bt are column-major matrices stored as transposed tensors."""
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    double = At.dtype == torch.float64
    gemv = lib.orc_dgemv_ld if double else lib.orc_sgemv_ld
    trmv = lib.orc_dtrmv_ld if double else lib.orc_strmv_ld
    vp, ci = ctypes.c_void_p, ctypes.c_int
    gemv.argtypes = [ci, ci, vp, ci, vp, vp]
    trmv.argtypes = [ci, ci, ci, ci, vp, ci, vp, vp]
    m, n = At.shape[0], Xt.shape[0]
    A, X, B = At.cpu().numpy(), Xt.cpu().numpy(), Bt.cpu().numpy()
    v = np.random.default_rng(seed).standard_normal(n)
    xv, bv, ay = np.zeros(m), np.zeros(m), np.zeros(m)
    gemv(m, n, X.ctypes.data, m, v.ctypes.data, xv.ctypes.data)
    gemv(m, n, B.ctypes.data, m, v.ctypes.data, bv.ctypes.data)
    trmv(LO, NOTR, NONUNIT, m, A.ctypes.data, m, xv.ctypes.data, ay.ctypes.data)
    a64 = np.linalg.norm(A.astype(np.float64))
    x64 = np.linalg.norm(X.astype(np.float64))
    return float(np.linalg.norm(ay - bv) / (a64 * x64 * np.linalg.norm(v)))


def run_extras(args, torch, M, lib, dev, g, A, B, C, extra):
    """The DGEMM with host pointers; DTRSM, DSYRK and SGEMM at 8192; the
    precision workloads (SGEMM 16384
    native vs BF16x3 split, STRSM 8192 x 1024 plain vs FP64-refined)."""
    n = args.n
    errs = extra["errors_vs_fp64"]
    steps = max(3, args.steps // 2)
    flops = 2.0 * n ** 3
    # The same DGEMM with host-pointer operands (numpy arrays): end to end,
    # PCIe transfers included (pinned staging; C streamed by column panels)
    Ah, Bh = A.cpu().numpy(), B.cpu().numpy()
    Ch = np.zeros((n, n))

    def dgemm_host():
        lib.cblas_dgemm(COL, TR, NOTR, n, n, n, 1.0, Ah.ctypes.data, n, Bh.ctypes.data, n,
                        0.0, Ch.ctypes.data, n)
    w_h, _ = timed_calls(torch, lib, dgemm_host, 3, 1)
    extra["dgemm_8192_host_pointers_ms"] = round(w_h, 2)
    extra["dgemm_8192_host_pointers_tflops"] = round(flops / (w_h * 1e-3) / 1e12, 3)
    cols = np.random.default_rng(7).integers(0, n, 16)
    extra["dgemm_8192_host_pointers_bits_match"] = bool(
        np.array_equal(Ch[cols], C[torch.from_numpy(cols).to(dev)].cpu().numpy()))
    del Ah, Bh, Ch
    # DTRSM L/L/N/NonUnit m=8192 nrhs=4096 (off-diag N(0,1)/sqrt(m), diag 1+|N|)
    m, nrhs = n, n // 2
    At = torch.randn((m, m), dtype=torch.float64, device=dev, generator=g) / m ** 0.5
    d = 1.0 + torch.randn(m, dtype=torch.float64, device=dev, generator=g).abs()
    At = torch.triu(At, 1) + torch.diag(d)
    B0 = torch.randn((nrhs, m), dtype=torch.float64, device=dev, generator=g)
    Bt = B0.clone()

    def dtrsm():
        Bt.copy_(B0)
        lib.cblas_dtrsm(COL, LEFT, LO, NOTR, NONUNIT, m, nrhs, 1.0, M.ptr(At), m,
                        M.ptr(Bt), m)

    def copy_only():
        Bt.copy_(B0)
    # three untimed solves first: the host-pointer DGEMM before leaves the
    # GPU idle between its transfers, and the first solves after it run slow
    w_t, _ = timed_calls(torch, lib, dtrsm, steps, 3)
    w_c, _ = timed_calls(torch, lib, copy_only, steps, 1)
    t_trsm = max(w_t - w_c, 1e-6)
    extra["dtrsm_8192x4096_tflops"] = round(m * m * nrhs / (t_trsm * 1e-3) / 1e12, 3)
    extra["dtrsm_pct_fp64_peak"] = round(100 * extra["dtrsm_8192x4096_tflops"] /
                                         PEAK["f64"], 2)
    dtrsm()  # the copy-only timing above restored B; solve once more to check it
    torch.cuda.synchronize()
    errs["dtrsm_8192x4096_residual"] = trsm_residual(torch, At, Bt, B0)
    del At, B0, Bt
    # DSYRK lower n=8192 k=2048
    k = n // 4
    As = torch.randn((k, n), dtype=torch.float64, device=dev, generator=g)  # A: n x k

    def dsyrk():
        lib.cblas_dsyrk(COL, LO, NOTR, n, k, 1.0, M.ptr(As), n, 0.0, M.ptr(C), n)
    w_s, _ = timed_calls(torch, lib, dsyrk, steps, 1)
    extra["dsyrk_8192_2048_tflops"] = round(n * (n + 1) * k / (w_s * 1e-3) / 1e12, 3)
    extra["dsyrk_pct_fp64_peak"] = round(100 * extra["dsyrk_8192_2048_tflops"] /
                                         PEAK["f64"], 2)
    del As
    # native FP32 SGEMM 8192 (op(A)=A^T, same shape)
    Af, Bf = A.float(), B.float()
    Cf = torch.zeros((n, n), dtype=torch.float32, device=dev)

    def sgemm():
        lib.cblas_sgemm(COL, TR, NOTR, n, n, n, 1.0, M.ptr(Af), n, M.ptr(Bf), n, 0.0,
                        M.ptr(Cf), n)
    w_f, _ = timed_calls(torch, lib, sgemm, steps, 1)
    extra["sgemm_8192_tflops"] = round(flops / (w_f * 1e-3) / 1e12, 3)
    extra["sgemm_pct_fp32_peak"] = round(100 * extra["sgemm_8192_tflops"] / PEAK["f32"], 2)
    errs["sgemm_8192"] = sampled_gemm_error(torch, Af, Bf, Cf)
    del Af, Bf, Cf
    if args.precision_n:
        run_precision(args, torch, M, lib, dev, g, extra)
    if not args.no_batched:
        run_batched(args, torch, M, lib, extra)
    if not args.no_serving:
        run_serving(args, torch, M, lib, extra)


# Grouped plans are timed as the mean of back-to-back replays after untimed
# ones: the first replays after the preceding workloads run while the clock
# settles (measured: the first 10 back-to-back replays of the batched plan
# 0.93 ms each, the next 20 0.89).
REPLAY_STEPS, REPLAY_WARMUP = 20, 10


def run_batched(args, torch, M, lib, extra):
    """Batched workload: 4096 SGEMM (M, N, K log-uniform [16, 512]) + 1024
    SSYRK lower (n log-uniform [32, 1024], k [16, 512]), alpha = beta = 1,
    U[-1, 1], one grouped launch replayed from a hipGraph; against the
    same problems issued as serial cblas_* calls."""
    from mublis_amd.workloads import grouped_problems
    ops, keep, specs = grouped_problems(torch)
    plan = M.GroupedPlan(ops)
    info = plan.info()
    stream = lib.mublis_rt_stream()

    def replay():
        plan.execute(stream)
    w_g, ev_g = timed_calls(torch, lib, replay, REPLAY_STEPS, REPLAY_WARMUP)
    plan.close()

    def serial():
        for s in specs:
            if s["op"] == "gemm":
                lib.cblas_sgemm(COL, NOTR, NOTR, s["m"], s["n"], s["k"], 1.0, M.ptr(s["A"]),
                                s["lda"], M.ptr(s["B"]), s["ldb"], 1.0, M.ptr(s["C"]), s["m"])
            else:
                lib.cblas_ssyrk(COL, LO, NOTR, s["n"], s["k"], 1.0, M.ptr(s["A"]), s["lda"],
                                1.0, M.ptr(s["C"]), s["n"])
    w_s, _ = timed_calls(torch, lib, serial, 1, 0)
    ideal = max(info["flops"] / (PEAK["f32"] * 1e12), info["bytes"] / 8e12)
    extra["batched_4096sgemm_1024ssyrk"] = {
        "tflops": round(info["flops"] / (ev_g * 1e-3) / 1e12, 3),
        "ms_per_launch": round(ev_g, 4),
        "flops": info["flops"], "min_hbm_bytes": info["bytes"],
        "roofline_frac": round(ideal / (ev_g * 1e-3), 4),
        "bound": "mfma" if info["flops"] / (PEAK["f32"] * 1e12) >= info["bytes"] / 8e12
                 else "hbm",
        "work_items": info["items"], "problems_per_tile_shape_32_64_128": info["tiles"],
        "serial_cblas_calls_ms": round(w_s, 3),
        "speedup_vs_serial_calls": round(w_s / ev_g, 2),
    }
    run_batched_trsm(torch, M, lib, extra)


def run_batched_trsm(torch, M, lib, extra):
    """Batched STRSM: 1024 problems with m, n log-uniform [16, 512], side,
    uplo and transpose random, diagonally dominant A, in one grouped plan
    (FP64 inverses of every diagonal block in one launch, then 16- or 32-column
    panels) replayed from its hipGraph; against the same problems as serial
    cblas_strsm calls.  Each replay solves in place again."""
    from mublis_amd.workloads import grouped_trsm_problems
    ops, keep, specs = grouped_trsm_problems(torch)
    plan = M.GroupedPlan(ops)
    info = plan.info()
    stream = lib.mublis_rt_stream()

    def replay():
        plan.execute(stream)
    _, ev_g = timed_calls(torch, lib, replay, REPLAY_STEPS, REPLAY_WARMUP)
    plan.close()

    def serial():
        for s in specs:
            lib.cblas_strsm(COL, RIGHT if s["side"] else LEFT, UP if s["upper"] else LO,
                            TR if s["trans"] else NOTR, NONUNIT, s["m"], s["n"], 1.0,
                            M.ptr(s["A"]), s["order"], M.ptr(s["B"]), s["m"])
    w_s, _ = timed_calls(torch, lib, serial, 1, 0)
    extra["batched_1024strsm"] = {
        "tflops": round(info["flops"] / (ev_g * 1e-3) / 1e12, 3),
        "ms_per_launch": round(ev_g, 4),
        "flops": info["flops"], "work_items": info["items"],
        "serial_cblas_calls_ms": round(w_s, 3),
        "speedup_vs_serial_calls": round(w_s / ev_g, 2),
    }


def serial_call(lib, S, r):
    """One serving request as the plain synchronous cblas_* call on the same
    device memory."""
    S.plain_call(lib, r)


def fifo_latencies(arrivals, service):
    """Latencies of a single FIFO server with these arrival times and service
    times (the serial per-call path at the same arrival process)."""
    free, out = 0.0, np.empty(len(arrivals))
    for i, (t, d) in enumerate(zip(arrivals, service)):
        start = max(t, free)
        free = start + d
        out[i] = free - t
    return out


def serve_run(S, srv, reqs, arrivals):
    """Submit reqs at the given arrival offsets (seconds) and wait for all;
    returns (stats, largest submission delay in ms).  The arrivals are
    issued by the library's native replay loop (mublis_server_replay):
    Python client threads delayed arrivals by 5-30 ms on some runs (the
    interpreter's threads not scheduled on time), which queued requests
    and moved p99 from under 1 ms to 7-20 ms with the same library."""
    srv.reset_stats()
    cr = [S.Server.request(r["op"], **{k: v for k, v in r.items() if k not in ("op", "flop")})
          for r in reqs]
    tickets, lag = srv.replay(cr, arrivals)
    for t in tickets:
        st, _ = srv.wait(t)
        assert st == 0, st
    return srv.stats(), lag


def run_serving(args, torch, M, lib, extra):
    """Serving workload: Poisson arrivals at a fixed rate (replayed by the
    library's native loop), 60% SGEMM / 20% SSYRK+SSYR2K / 20% DTRSM (sizes in
    mublis_amd.workloads.serving_requests), operands device-resident by
    handle in a 64 GB budget; against the same requests issued as serial
    synchronous cblas_* calls."""
    import mublis_amd.serving as S
    from mublis_amd.workloads import serving_requests
    srv = S.Server(hbm_budget=64 << 30, streams=args.serve_streams,
                   slots_per_stream=args.serve_slots)
    try:
        reqs, _ = serving_requests(torch, lib, srv, count=args.serve_requests)
        flop = sum(r["flop"] for r in reqs)
        for r in reqs[:40]:  # warm both paths
            serial_call(lib, S, r)
        serve_run(S, srv, reqs[:40], np.zeros(40))
        service = np.empty(len(reqs))
        t_all = time.perf_counter()
        for i, r in enumerate(reqs):
            t = time.perf_counter()
            serial_call(lib, S, r)
            service[i] = time.perf_counter() - t
        serial_s = time.perf_counter() - t_all
        rng = np.random.default_rng(7)
        arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, len(reqs)))
        st, lag = serve_run(S, srv, reqs, arrivals)
        fifo = 1e3 * fifo_latencies(arrivals, service)
        burst, _ = serve_run(S, srv, reqs, np.zeros(len(reqs)))
        mix, by_op = {}, {}
        for r, d in zip(reqs, service):
            name = S.OP_NAMES[r["op"]]
            mix[name] = mix.get(name, 0) + 1
            f, t = by_op.get(name, (0.0, 0.0))
            by_op[name] = (f + r["flop"], t + d)
        extra["serving_mixed_l3"] = {
            "requests": len(reqs), "mix": mix, "arrival_rate_per_s": args.serve_rate,
            "tflops": round(st["tflops"], 3),
            "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
            "mean_ms": round(st["mean_ms"], 3), "max_ms": round(st["max_ms"], 3),
            "elapsed_s": round(st["elapsed_s"], 4),
            "serial_cblas_tflops": round(flop / serial_s / 1e12, 3),
            "serial_cblas_s": round(serial_s, 4),
            "serial_s_and_tflops_by_op": {k: [round(t, 4), round(f / t / 1e12, 2)]
                                          for k, (f, t) in by_op.items()},
            "serial_fifo_same_arrivals_p50_ms": round(float(np.percentile(fifo, 50)), 3),
            "serial_fifo_same_arrivals_p99_ms": round(float(np.percentile(fifo, 99)), 3),
            "throughput_vs_serial": round(st["tflops"] / (flop / serial_s / 1e12), 2),
            "burst_tflops": round(burst["tflops"], 3),
            "burst_p50_ms": round(burst["p50_ms"], 3), "burst_p99_ms": round(burst["p99_ms"], 3),
            "launches": st["launches"], "grouped_launches": st["grouped_launches"],
            "grouped_requests": st["grouped_requests"], "max_pending": st["max_pending"],
            "max_inflight_launches": st["max_inflight"],
            "hbm_budget_gb": round(st["hbm_budget"] / 2 ** 30, 1),
            "hbm_peak_gb": round(st["hbm_peak"] / 2 ** 30, 2),
            "max_arrival_lag_ms": round(lag, 3),
            "hip_hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
        }
    finally:
        srv.close()


def run_precision(args, torch, M, lib, dev, g, extra):
    errs = extra["errors_vs_fp64"]
    n = args.precision_n
    steps = 3
    # rows of op(A) and of B scaled by 10^U(-3,3); op(A) = A^T so the tensor
    # rows are the rows of op(A)
    At = torch.randn((n, n), device=dev, generator=g) * \
        10.0 ** (torch.rand((n, 1), device=dev, generator=g) * 6 - 3)
    Bt = torch.randn((n, n), device=dev, generator=g) * \
        10.0 ** (torch.rand((1, n), device=dev, generator=g) * 6 - 3)
    Ct = torch.empty((n, n), device=dev)
    f = lib.mublis_sgemm_bf16x3
    f.argtypes = [ctypes.c_int] * 5 + [ctypes.c_float, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.c_int]
    f.restype = ctypes.c_int

    def native():
        lib.cblas_sgemm(COL, TR, NOTR, n, n, n, 1.0, M.ptr(At), n, M.ptr(Bt), n, 0.0,
                        M.ptr(Ct), n)

    def split():
        f(1, 0, n, n, n, 1.0, M.ptr(At), 1, n, M.ptr(Bt), 1, n, 0.0, M.ptr(Ct), 1, n)
    fl = 2.0 * n ** 3
    w_n, _ = timed_calls(torch, lib, native, steps, 1)
    errs[f"sgemm_{n}_native_rowscaled"] = sampled_gemm_error(torch, At, Bt, Ct, 128)
    w_s, _ = timed_calls(torch, lib, split, steps, 1)
    errs[f"sgemm_{n}_bf16x3_rowscaled"] = sampled_gemm_error(torch, At, Bt, Ct, 128)
    extra[f"sgemm_{n}_native_tflops"] = round(fl / (w_n * 1e-3) / 1e12, 3)
    extra[f"sgemm_{n}_bf16x3_tflops"] = round(fl / (w_s * 1e-3) / 1e12, 3)
    extra[f"sgemm_{n}_bf16x3_bf16_mfma_util_pct"] = round(
        100 * 6 * fl / (w_s * 1e-3) / 2516.6e12, 2)
    del At, Bt, Ct
    torch.cuda.empty_cache()
    # STRSM L/L m=8192 nrhs=1024: diag U[1,2], off-diag N(0,1)/sqrt(m), 2 sweeps
    m, nrhs = 8192, 1024
    At = torch.randn((m, m), device=dev, generator=g) / m ** 0.5
    At = torch.triu(At, 1) + torch.diag(1 + torch.rand(m, device=dev, generator=g))
    B0 = torch.randn((nrhs, m), device=dev, generator=g)
    Xp, Xr = B0.clone(), B0.clone()
    r = lib.mublis_strsm_refined
    r.argtypes = [ctypes.c_int] * 6 + [ctypes.c_float, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int]
    r.restype = ctypes.c_int

    def plain():
        Xp.copy_(B0)
        lib.cblas_strsm(COL, LEFT, LO, NOTR, NONUNIT, m, nrhs, 1.0, M.ptr(At), m,
                        M.ptr(Xp), m)

    def refined():
        Xr.copy_(B0)
        r(0, 1, 0, 0, m, nrhs, 1.0, M.ptr(At), 1, m, M.ptr(Xr), 1, m, 2)
    w_p, _ = timed_calls(torch, lib, plain, steps, 1)
    w_r, _ = timed_calls(torch, lib, refined, steps, 1)
    fl = float(m) * m * nrhs
    extra["strsm_8192x1024_tflops"] = round(fl / (w_p * 1e-3) / 1e12, 3)
    extra["strsm_8192x1024_refined2_tflops"] = round(fl / (w_r * 1e-3) / 1e12, 3)
    Ad, Xd = At.double(), B0.double()
    lib.cblas_dtrsm(COL, LEFT, LO, NOTR, NONUNIT, m, nrhs, 1.0, M.ptr(Ad), m, M.ptr(Xd), m)
    torch.cuda.synchronize()
    errs["strsm_8192x1024_plain_fwd_vs_dtrsm"] = ((Xp.double() - Xd).norm() / Xd.norm()).item()
    errs["strsm_8192x1024_refined_fwd_vs_dtrsm"] = ((Xr.double() - Xd).norm() /
                                                    Xd.norm()).item()
    errs["strsm_8192x1024_refined_residual"] = trsm_residual(torch, At, Xr, B0)
    # the floor for an FP32 answer: the FP64 solution rounded to FP32, same
    # FP32 triangle and right-hand side
    errs["dtrsm_8192x1024_fp32_rounded_residual"] = trsm_residual(torch, At, Xd.float(), B0)
    errs["dtrsm_8192x1024_residual"] = trsm_residual(torch, Ad, Xd, B0.double())


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)
    # The serving loop keeps 4 streams busy at once; HIP spreads a process's
    # streams over GPU_MAX_HW_QUEUES hardware queues (4 by default), so when
    # streams created earlier hold queues (the library's per-thread stream,
    # torch's, the host-staging upload and download streams of the
    # host-pointer DGEMM) two server streams share a queue and run in order
    # (measured: 4 queues p50 0.79 ms; 8 queues with the staging streams
    # present p50 0.24 / p99 8.7 ms; 16 queues 0.13 / 0.96 ms, as the server
    # alone).  Raised to 16 unless set higher; must happen before HIP
    # initialises.
    try:
        hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
    except ValueError:
        hwq = 4
    if hwq < 16:
        os.environ["GPU_MAX_HW_QUEUES"] = "16"
    # The CPU baseline runs first: its concurrent workers are forked from a
    # process that has not touched the GPU yet.
    cpu = None
    if not (args.no_cpu or args.only or args.traffic_probe):
        cpu = cpu_baseline(args.n, args.cpu_slice)
    import torch
    import mublis_amd as M

    lib = M.load()
    M.check(lib.mublis_rt_init(), "mublis_rt_init")
    dev = torch.device("cuda:0")
    n = args.n
    g = torch.Generator(device=dev).manual_seed(314159)
    if args.only == "serving":
        extra = {}
        run_serving(args, torch, M, lib, extra)
        print(json.dumps(extra), flush=True)
        return
    if args.only or args.traffic_probe:
        return profile_mode(args, torch, M, lib, dev, g)
    # Column-major K x M A is the row-major (M, K) tensor; op(A) = A^T.
    A = torch.randn((n, n), dtype=torch.float64, device=dev, generator=g)
    B = torch.randn((n, n), dtype=torch.float64, device=dev, generator=g)
    C = torch.zeros((n, n), dtype=torch.float64, device=dev)
    pa, pb, pc = M.ptr(A), M.ptr(B), M.ptr(C)

    def dgemm():
        lib.cblas_dgemm(COL, TR, NOTR, n, n, n, 1.0, pa, n, pb, n, 0.0, pc, n)

    wall, ev = timed_calls(torch, lib, dgemm, args.steps, args.warmup)
    flops = 2.0 * n * n * n
    tflops = flops / (wall * 1e-3) / 1e12
    achieved = flops / (ev * 1e-3) / 1e12
    assert torch.isfinite(C[:64, :64]).all().item()

    extra = {"errors_vs_fp64": {}}
    errs = extra["errors_vs_fp64"]
    errs["dgemm_8192"] = sampled_gemm_error(torch, A, B, C)
    if not args.no_extra:
        run_extras(args, torch, M, lib, dev, g, A, B, C, extra)

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "dgemm_8192_transA_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": round(tflops, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic N(0,1) generated on device (seed 314159)",
        "config": {"workload": f"DGEMM M=N=K={n} op(A)=A^T alpha=1 beta=0, cblas_dgemm "
                               "with device-resident operands",
                   "pct_fp64_mfma_peak": round(100 * tflops / PEAK["f64"], 2),
                   **extra},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK["f64"],
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK["f64"], 4),
                     "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
