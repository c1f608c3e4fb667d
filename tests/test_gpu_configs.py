"""GPU checks of three benchmark configurations at their full sizes, through
the C ABI.

* Plain cblas_sgemm, row-major, no transpose, M = N = K = 1024, alpha = 1,
  beta = 0.5, A/B/C uniform[-1, 1] from seed 314159 (numpy default_rng):
  the row-major call is the column-major problem C' = B' A'
  (frame/compat/cblas/cblas.c:188-200).  Bar: ||C - C*||_F <= 2 k eps
  (|alpha| ||A|| ||B|| + |beta| ||C0||) against the oracle's orc_sgemm
  (double accumulation, rounded once) and within twice that of the
  reference's own linux_avx2 build when it is present; host pointers and
  device pointers.
* The batched workload's full plan (4096 SGEMM, M, N, K log-uniform
  [16, 512], plus 1024 SSYRK lower, n log-uniform [32, 1024], k log-uniform
  [16, 512], alpha = beta = 1, uniform[-1, 1]; the plan the benchmark
  replays): every problem bit-identical to its single cblas_sgemm /
  cblas_ssyrk call, and 64 sampled problems within 2 k eps (|alpha| ||A||
  ||B|| + |beta| ||C0||) of the oracle.
* The serving workload's full request set (2,000 requests: 60 % SGEMM,
  20 % SSYRK / SSYR2K, 20 % DTRSM, sizes up to 4096, operands resident in a
  64 GB budget), submitted from 8 threads at once: every result
  bit-identical to the plain cblas_* call on the same operands, and a
  sample within the oracle bound (GEMM / SYRK / SYR2K: 4 (k + 2) eps
  (|alpha| ||A|| ||B|| + |beta| ||C0||), twice the product term for SYR2K;
  DTRSM: residual ||A X - alpha B|| <= 4 m eps ||A|| ||X||).
"""
import ctypes
import os
import threading

import numpy as np
import pytest

import l3cases as L
import mublis_amd as M

pytestmark = pytest.mark.gpu

EPS32 = np.finfo(np.float32).eps
EPS64 = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def lib():
    lib = M.load()
    assert lib.mublis_rt_init() == 0
    return lib


@pytest.fixture(scope="module")
def orc():
    return L.Cblas(L.ORACLE_LIB, prefix="orc_", returns_status=True)


# -------------------------------------------- row-major SGEMM 1024^3 ----
@pytest.mark.parametrize("dev", [False, True], ids=["host", "device"])
def test_rowmajor_sgemm_1024(lib, orc, dev):
    n = 1024
    rng = np.random.default_rng(314159)
    A = rng.uniform(-1.0, 1.0, (n, n)).astype(np.float32)   # C order = row-major
    B = rng.uniform(-1.0, 1.0, (n, n)).astype(np.float32)
    C0 = rng.uniform(-1.0, 1.0, (n, n)).astype(np.float32)
    if dev:
        import torch
        dA, dB, dC = (torch.from_numpy(x).cuda() for x in (A, B, C0.copy()))
        lib.cblas_sgemm(L.ROW, L.NOTR, L.NOTR, n, n, n, 1.0, M.ptr(dA), n, M.ptr(dB), n, 0.5,
                        M.ptr(dC), n)
        torch.cuda.synchronize()
        got = dC.cpu().numpy()
    else:
        got = C0.copy()
        lib.cblas_sgemm(L.ROW, L.NOTR, L.NOTR, n, n, n, 1.0, A.ctypes.data, n, B.ctypes.data,
                        n, 0.5, got.ctypes.data, n)
    want = C0.copy()
    assert orc.call(np.float32, "gemm", L.ROW, L.NOTR, L.NOTR, n, n, n, 1.0, A, n, B, n, 0.5,
                    want, n) == 0
    bound = 2 * n * EPS32 * (np.linalg.norm(A.astype(np.float64)) *
                             np.linalg.norm(B.astype(np.float64)) +
                             0.5 * np.linalg.norm(C0.astype(np.float64)))
    err = np.linalg.norm(got.astype(np.float64) - want)
    assert err <= bound, (err, bound)
    # exact float64 product as a second reference for the oracle itself
    exact = A.astype(np.float64) @ B.astype(np.float64) + 0.5 * C0.astype(np.float64)
    assert np.linalg.norm(want - exact) <= 1e-3 * bound
    if os.path.exists(L.REF_AVX2_LIB):
        ref = L.Cblas(L.REF_AVX2_LIB)
        r = C0.copy()
        ref.call(np.float32, "gemm", L.ROW, L.NOTR, L.NOTR, n, n, n, 1.0, A, n, B, n, 0.5, r, n)
        assert np.linalg.norm(got.astype(np.float64) - r) <= 2 * bound
    print(f"row-major sgemm 1024: error vs oracle {err:.3e} (bound {bound:.3e})")


# ---------------------------------------- the full grouped batched plan --
def test_full_grouped_plan(lib, orc):
    torch = pytest.importorskip("torch")
    from mublis_amd.workloads import grouped_problems
    ops, keep, specs = grouped_problems(torch)
    assert len(ops) == 5120
    initial = [s["C"].clone() for s in specs]
    plan = M.GroupedPlan(ops)
    try:
        plan.execute()
        torch.cuda.synchronize()
    finally:
        plan.close()
    grouped = [s["C"].clone() for s in specs]
    bad = []
    for i, (s, c0) in enumerate(zip(specs, initial)):
        C = c0.clone()
        if s["op"] == "gemm":
            lib.cblas_sgemm(102, 111, 111, s["m"], s["n"], s["k"], 1.0, M.ptr(s["A"]),
                            s["lda"], M.ptr(s["B"]), s["ldb"], 1.0, M.ptr(C), s["m"])
        else:
            lib.cblas_ssyrk(102, 122, 111, s["n"], s["k"], 1.0, M.ptr(s["A"]), s["lda"], 1.0,
                            M.ptr(C), s["n"])
        if not torch.equal(C, grouped[i]):
            bad.append(i)
    torch.cuda.synchronize()
    assert not bad, f"{len(bad)} of {len(specs)} problems differ from their single call"
    # 64 problems against the oracle: 48 SGEMM and 16 SSYRK, drawn by seed
    rng = np.random.default_rng(64)
    n_gemm = sum(1 for s in specs if s["op"] == "gemm")
    pick = list(rng.choice(n_gemm, 48, replace=False)) + \
        list(n_gemm + rng.choice(len(specs) - n_gemm, 16, replace=False))
    for i in pick:
        s = specs[i]
        c0 = initial[i].cpu().numpy().T.copy(order="F")
        got = grouped[i].cpu().numpy().T
        a = np.asfortranarray(s["A"].cpu().numpy().T)
        want = c0.copy(order="F")
        if s["op"] == "gemm":
            b = np.asfortranarray(s["B"].cpu().numpy().T)
            assert orc.call(np.float32, "gemm", 102, 111, 111, s["m"], s["n"], s["k"], 1.0, a,
                            s["lda"], b, s["ldb"], 1.0, want, s["m"]) == 0
            scale = np.linalg.norm(a) * np.linalg.norm(b) + np.linalg.norm(c0)
            err = np.linalg.norm(got - want)
        else:
            assert orc.call(np.float32, "syrk", 102, 122, 111, s["n"], s["k"], 1.0, a,
                            s["lda"], 1.0, want, s["n"]) == 0
            low = np.tril(np.ones((s["n"], s["n"]), bool))
            scale = np.linalg.norm(a) ** 2 + np.linalg.norm(c0[low])
            err = np.linalg.norm(np.where(low, got - want, 0))
            np.testing.assert_array_equal(got[~low], c0[~low])
        assert err <= 2 * s["k"] * EPS32 * scale, (i, s["op"], err)


# ------------------------------------------- the full serving request set --
def _output(r):
    """The matrix a request writes: C, or B for a TRSM."""
    return r["b"] if r["op"] in (4, 5) else r["c"]


def _snapshot(torch, lib, mats):
    """Device copies of the given matrices (as (cols, ld) tensors)."""
    out = []
    for h in mats:
        dt = torch.float64 if h.dtype == 1 else torch.float32
        t = torch.empty((h.cols, h.ld), dtype=dt, device="cuda")
        cp = lib.mublis_gfx950_dcopy_dev if h.dtype == 1 else lib.mublis_gfx950_scopy_dev
        assert cp(h.rows, h.cols, h.ptr, 1, h.ld, t.data_ptr(), 1, h.ld, None) == 0
        out.append(t)
    lib.mublis_rt_sync(None)
    return out


def _restore(torch, lib, mats, snaps):
    for h, t in zip(mats, snaps):
        cp = lib.mublis_gfx950_dcopy_dev if h.dtype == 1 else lib.mublis_gfx950_scopy_dev
        assert cp(h.rows, h.cols, t.data_ptr(), 1, h.ld, h.ptr, 1, h.ld, None) == 0
    lib.mublis_rt_sync(None)


def _dense(t, h):
    return t[:, :h.rows].cpu().numpy().T.astype(np.float64)


def test_full_serving_request_set(lib):
    torch = pytest.importorskip("torch")
    import mublis_amd.serving as S
    from mublis_amd.workloads import serving_requests
    srv = S.Server(hbm_budget=64 << 30)
    try:
        reqs, _ = serving_requests(torch, lib, srv, count=2000)
        outs = [_output(r) for r in reqs]
        initial = _snapshot(torch, lib, outs)
        kw = [{k: v for k, v in r.items() if k not in ("op", "flop")} for r in reqs]
        tickets = [None] * len(reqs)

        def client(c):
            for i in range(c, len(reqs), 8):
                tickets[i] = srv.submit(reqs[i]["op"], **kw[i])
        ts = [threading.Thread(target=client, args=(c,)) for c in range(8)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        for t in tickets:
            assert srv.wait(t)[0] == 0
        served = _snapshot(torch, lib, outs)
        _restore(torch, lib, outs, initial)
        for r in reqs:
            S.plain_call(lib, r)
        lib.mublis_rt_sync(None)
        plain = _snapshot(torch, lib, outs)
        bad = [i for i, (h, x, y) in enumerate(zip(outs, served, plain))
               if not torch.equal(x[:, :h.rows], y[:, :h.rows])]
        assert not bad, (f"{len(bad)} of {len(reqs)} served results differ from the plain "
                         f"call, first {S.OP_NAMES[reqs[bad[0]]['op']]}")
        # a sample against the oracle: requests small enough for the CPU
        orc = L.Cblas(L.ORACLE_LIB, prefix="orc_", returns_status=True)
        rng = np.random.default_rng(5)
        small = [i for i, r in enumerate(reqs) if r["flop"] <= 2e8]
        by_op = {}
        for i in small:
            by_op.setdefault(reqs[i]["op"], []).append(i)
        sample = []
        for op, idx in sorted(by_op.items()):
            sample += list(rng.choice(idx, min(8, len(idx)), replace=False))
        assert len(sample) >= 24
        for i in sample:
            r = reqs[i]
            h = _output(r)
            got = _dense(served[i], h)
            c0 = _dense(initial[i], h)
            if r["op"] == S.DTRSM:
                a = r["a"].read()
                x = got
                res = np.linalg.norm(np.tril(a) @ x - r["alpha"] * c0)
                assert res <= 4 * r["m"] * EPS64 * np.linalg.norm(np.tril(a)) * \
                    np.linalg.norm(x), (i, res)
                continue
            a = r["a"].read()
            want = np.asfortranarray(c0.astype(np.float32))
            c00 = want.astype(np.float64)
            if r["op"] == S.SGEMM:
                b = r["b"].read()
                assert orc.call(np.float32, "gemm", 102, 111, 111, r["m"], r["n"], r["k"],
                                r["alpha"], a, r["m"], b, r["k"], r["beta"], want,
                                r["m"]) == 0
                prod = np.linalg.norm(a) * np.linalg.norm(b)
                write = np.ones_like(got, bool)
            else:
                lo = r["uplo"] == S.LOWER
                write = np.tril(np.ones_like(got, bool)) if lo else \
                    np.triu(np.ones_like(got, bool))
                if r["op"] == S.SSYRK:
                    assert orc.call(np.float32, "syrk", 102, 122 if lo else 121, 111, r["n"],
                                    r["k"], r["alpha"], a, r["n"], r["beta"], want,
                                    r["n"]) == 0
                    prod = np.linalg.norm(a) ** 2
                else:
                    b = r["b"].read()
                    assert orc.call(np.float32, "syr2k", 102, 122 if lo else 121, 111, r["n"],
                                    r["k"], r["alpha"], a, r["n"], b, r["n"], r["beta"], want,
                                    r["n"]) == 0
                    prod = 2 * np.linalg.norm(a) * np.linalg.norm(b)
                np.testing.assert_array_equal(got[~write], c00[~write])
            err = np.linalg.norm(np.where(write, got - want, 0))
            bound = 4 * (r["k"] + 2) * EPS32 * (abs(r["alpha"]) * prod +
                                                abs(r["beta"]) * np.linalg.norm(c00[write]))
            assert err <= bound, (i, S.OP_NAMES[r["op"]], err, bound)
    finally:
        srv.close()
