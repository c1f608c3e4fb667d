"""Synthetic workloads of the benchmark configurations, built on the device.

The sizes are drawn on the host first (grouped_shapes, serving_plan), so a
CPU run can draw the same problems without a device.

`grouped_problems` draws the batched configuration: GEMM problems with M,
N, K log-uniform in a range and SYRK problems with n and k log-uniform,
entries uniform[-1, 1], column-major device tensors, plus the
`SbatchOp` descriptors that point at them.  `grouped_trsm_problems` draws
batched STRSM problems the same way.  Tensors are stored transposed
(shape (cols, ld)) so that a row-major torch tensor is a column-major
matrix.
"""
import math

import numpy as np

from . import BATCH_GEMM, BATCH_SYR2K, BATCH_SYRK, BATCH_TRSM, SbatchOp


def _logu(rng, lo, hi, size):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size)).astype(np.int64).clip(lo, hi)


_ALIGN = 64  # elements: every carved operand starts 256-byte aligned


def _carve(torch, shapes, gen, device, dtype=None):
    """Views of the given (cols, ld) shapes into ONE device allocation filled
    uniform[-1, 1] by ONE kernel.  Building thousands of operands tensor by
    tensor issued one allocation and two or three fill kernels each (about
    43,000 launches for the batched workload); this issues one."""
    offs, total = [], 0
    for c, r in shapes:
        offs.append(total)
        total += (c * r + _ALIGN - 1) // _ALIGN * _ALIGN
    pool = torch.empty(max(total, 1), device=device, dtype=dtype or torch.float32)
    pool.uniform_(-1.0, 1.0, generator=gen)
    return pool, [pool[o:o + c * r].view(c, r) for o, (c, r) in zip(offs, shapes)]


def grouped_shapes(seed=314159, n_gemm=4096, n_syrk=1024, gemm_range=(16, 512),
                   syrk_n=(32, 1024), syrk_k=(16, 512), vary_flags=False, n_syr2k=0):
    """The batched configuration's problem sizes and flags, drawn on the host:
    (ms, ns, ks, gflags, sn, sk, sflags) -- GEMM M, N, K and (trans_a,
    trans_b), then SYRK (and SYR2K, last) n, k and (trans, upper)."""
    rng = np.random.default_rng(seed)
    ms = _logu(rng, *gemm_range, n_gemm)
    ns = _logu(rng, *gemm_range, n_gemm)
    ks = _logu(rng, *gemm_range, n_gemm)
    gflags = [(int(rng.integers(0, 2)), int(rng.integers(0, 2))) if vary_flags else (0, 0)
              for _ in range(n_gemm)]
    sn = _logu(rng, *syrk_n, n_syrk + n_syr2k)
    sk = _logu(rng, *syrk_k, n_syrk + n_syr2k)
    sflags = [(int(rng.integers(0, 2)), int(rng.integers(0, 2))) if vary_flags else (0, 0)
              for _ in range(n_syrk + n_syr2k)]
    return ms, ns, ks, gflags, sn, sk, sflags


def grouped_problems(torch, seed=314159, n_gemm=4096, n_syrk=1024, gemm_range=(16, 512),
                     syrk_n=(32, 1024), syrk_k=(16, 512), alpha=1.0, beta=1.0,
                     vary_flags=False, device="cuda", n_syr2k=0):
    """Returns (ops, keep, specs): SbatchOp list, the tensors they point at
    (kept alive by the caller) and per-problem dicts for checking.  SYR2K
    problems (n_syr2k, drawn like the SYRK ones) come last."""
    g = torch.Generator(device=device).manual_seed(seed)
    ms, ns, ks, gflags, sn, sk, sflags = grouped_shapes(seed, n_gemm, n_syrk, gemm_range,
                                                        syrk_n, syrk_k, vary_flags, n_syr2k)
    # column-major rows x cols = tensor (cols, rows)
    shapes = []
    for i in range(n_gemm):
        m, n, k = int(ms[i]), int(ns[i]), int(ks[i])
        ta, tb = gflags[i]
        shapes += [(m, k) if ta else (k, m), (k, n) if tb else (n, k), (n, m)]
    for i in range(n_syrk + n_syr2k):
        n, k = int(sn[i]), int(sk[i])
        tr = sflags[i][0]
        a_shape = (n, k) if tr else (k, n)
        shapes += [a_shape] + ([a_shape] if i >= n_syrk else []) + [(n, n)]
    pool, views = _carve(torch, shapes, g, device)
    it = iter(views)
    ops, keep, specs = [], [pool], []
    for i in range(n_gemm):
        m, n, k = int(ms[i]), int(ns[i]), int(ks[i])
        ta, tb = gflags[i]
        A, B, C = next(it), next(it), next(it)
        lda, ldb = A.shape[1], B.shape[1]
        ops.append(SbatchOp(BATCH_GEMM, ta, tb, 0, m, n, k, alpha, beta, A.data_ptr(), lda,
                            B.data_ptr(), ldb, C.data_ptr(), m))
        specs.append(dict(op="gemm", m=m, n=n, k=k, ta=ta, tb=tb, A=A, B=B, C=C, lda=lda,
                          ldb=ldb))
    for i in range(n_syrk + n_syr2k):
        n, k = int(sn[i]), int(sk[i])
        tr, up = sflags[i]
        two = i >= n_syrk
        A = next(it)
        B = next(it) if two else None
        C = next(it)
        ops.append(SbatchOp(BATCH_SYR2K if two else BATCH_SYRK, tr, 0, 2 if up else 1, n, n,
                            k, alpha, beta, A.data_ptr(), A.shape[1],
                            B.data_ptr() if two else None, B.shape[1] if two else 1,
                            C.data_ptr(), n))
        specs.append(dict(op="syr2k" if two else "syrk", n=n, k=k, trans=tr, upper=up, A=A,
                          B=B, C=C, lda=A.shape[1], ldb=B.shape[1] if two else 1))
    return ops, keep, specs


def grouped_trsm_problems(torch, seed=271828, count=1024, size=(16, 512), alpha=1.0,
                          device="cuda"):
    """`count` STRSM problems for a grouped launch: B m x n with m, n
    log-uniform in `size`, side, uplo and trans_a drawn uniformly, non-unit
    diagonal; A's stored triangle has off-diagonal entries uniform[-1, 1] /
    order and diagonal uniform[1, 2] (diagonally dominant), B uniform[-1, 1].
    The triangles are drawn on the host and uploaded in one copy, the
    right-hand sides carved from one device pool.  Returns (ops, keep,
    specs) like grouped_problems."""
    rng = np.random.default_rng(seed)
    g = torch.Generator(device=device).manual_seed(seed)
    ms = _logu(rng, *size, count)
    ns = _logu(rng, *size, count)
    flags = [tuple(int(x) for x in rng.integers(0, 2, size=3)) for _ in range(count)]
    orders = [int(ns[i]) if flags[i][0] else int(ms[i]) for i in range(count)]
    offs, total = [], 0
    for o in orders:
        offs.append(total)
        total += (o * o + _ALIGN - 1) // _ALIGN * _ALIGN
    host = np.empty(max(total, 1), np.float32)
    for o, off in zip(orders, offs):
        a = rng.uniform(-1.0, 1.0, (o, o)) / o
        np.fill_diagonal(a, 1.0 + rng.uniform(0.0, 1.0, o))
        host[off:off + o * o] = a.ravel()
    apool = torch.from_numpy(host).to(device)
    bpool, bviews = _carve(torch, [(int(ns[i]), int(ms[i])) for i in range(count)], g, device)
    ops, keep, specs = [], [apool, bpool], []
    for i in range(count):
        m, n = int(ms[i]), int(ns[i])
        side, up, tr = flags[i]
        order = orders[i]
        A = apool[offs[i]:offs[i] + order * order].view(order, order)
        B = bviews[i]
        ops.append(SbatchOp(BATCH_TRSM, tr, 0, 2 if up else 1, m, n, 0, alpha, 0.0,
                            A.data_ptr(), order, None, 1, B.data_ptr(), m, side, 0))
        specs.append(dict(op="trsm", m=m, n=n, side=side, upper=up, trans=tr, A=A, B=B,
                          order=order))
    return ops, keep, specs


def _log_uniform(rng, lo, hi):
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


_SERVE_POOL = 1 << 25  # elements of each random pool; covers the largest operand (4096^2)


def serving_plan(seed=2024, count=2000):
    """The mixed-L3 serving workload's request shapes, drawn on the host.

    60% SGEMM (M, N, K log-uniform [64, 4096]), 20% SSYRK/SSYR2K (half each;
    n log-uniform [128, 4096], k log-uniform [64, 2048], triangle chosen at
    random), 20% DTRSM (Left, Lower, NoTrans, NonUnit; m log-uniform
    [256, 4096], nrhs log-uniform [1, 512]).  Each entry is a dict with "op"
    (a mublis_amd.serving op code), its sizes, "uplo" where it applies and
    "mats": the operands in draw order, (role, dtype, rows, cols, offset)
    with `offset` the operand's start in its random pool.  Pure Python: the
    CPU baseline draws the same requests without a device."""
    from . import serving as S
    import random
    rng = random.Random(seed)
    plan = []

    def mat(r, role, dt, rows, cols):
        r["mats"].append((role, dt, rows, cols, rng.randrange(0, _SERVE_POOL - rows * cols + 1)))

    for _ in range(count):
        u = rng.random()
        if u < 0.6:
            m, n, k = (_log_uniform(rng, 64, 4096) for _ in range(3))
            r = dict(op=S.SGEMM, m=m, n=n, k=k, mats=[])
            mat(r, "a", S.F32, m, k)
            mat(r, "b", S.F32, k, n)
            mat(r, "c", S.F32, m, n)
            r["flop"] = 2.0 * m * n * k
        elif u < 0.8:
            n, k = _log_uniform(rng, 128, 4096), _log_uniform(rng, 64, 2048)
            two = rng.random() < 0.5
            uplo = S.LOWER if rng.random() < 0.5 else S.UPPER
            r = dict(op=S.SSYR2K if two else S.SSYRK, n=n, k=k, uplo=uplo, mats=[])
            mat(r, "a", S.F32, n, k)
            mat(r, "c", S.F32, n, n)
            if two:
                mat(r, "b", S.F32, n, k)
            r["flop"] = (2.0 if two else 1.0) * n * (n + 1) * k
        else:
            m, nrhs = _log_uniform(rng, 256, 4096), _log_uniform(rng, 1, 512)
            r = dict(op=S.DTRSM, m=m, n=nrhs, uplo=S.LOWER, mats=[])
            mat(r, "a", S.F64, m, m)
            mat(r, "b", S.F64, m, nrhs)
            r["flop"] = float(m) * m * nrhs
        plan.append(r)
    return plan


def serving_requests(torch, lib, server, seed=2024, count=2000, alpha=1.0, beta=0.5):
    """The mixed-L3 serving workload (serving_plan) on device-resident
    matrices: entries U[-1, 1], copied from a shared random pool on the
    device at the plan's offsets; DTRSM triangles diagonally dominant
    (diagonal m, off-diagonal U[-1, 1]).  Returns (requests, matrices): each
    request is a dict of mublis_submit keyword arguments plus "op" and
    "flop"."""
    from . import serving as S
    gen = torch.Generator(device="cuda").manual_seed(seed)
    pools = {S.F32: torch.rand(_SERVE_POOL, device="cuda", generator=gen) * 2 - 1,
             S.F64: torch.rand(_SERVE_POOL, device="cuda", generator=gen,
                               dtype=torch.float64) * 2 - 1}
    copy = {S.F32: lib.mublis_gfx950_scopy_dev, S.F64: lib.mublis_gfx950_dcopy_dev}
    mats, reqs, keep = [], [], []
    for p in serving_plan(seed, count):
        r = {k: v for k, v in p.items() if k != "mats"}
        if p["op"] != S.DTRSM:
            r["alpha"], r["beta"] = alpha, beta
        else:
            r["alpha"] = alpha
        for role, dt, rows, cols, off in p["mats"]:
            h = server.matrix(dt, rows, cols)
            pool = pools[dt]
            addr = pool.data_ptr() + off * pool.element_size()
            st = copy[dt](rows, cols, addr, 1, rows, h.ptr, 1, h.ld, None)
            assert st == 0, st
            mats.append(h)
            r[role] = h
            if p["op"] == S.DTRSM and role == "a":
                m = rows
                diag = torch.full((m,), float(m), dtype=torch.float64, device="cuda")
                st = lib.mublis_gfx950_dcopy_dev(1, m, diag.data_ptr(), 1, 1, h.ptr, 1,
                                                 h.ld + 1, None)
                assert st == 0, st
                keep.append(diag)  # alive until the copies have run
        reqs.append(r)
    lib.mublis_rt_sync(None)
    torch.cuda.synchronize()
    return reqs, mats
