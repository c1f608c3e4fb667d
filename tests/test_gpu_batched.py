"""GPU checks of the grouped persistent batched kernel (uneven FP32 GEMM,
SYRK and TRSM problems in one hipGraph-captured plan), through the C ABI.

Bars: every GEMM/SYRK problem's result is bit-identical to the
corresponding single cblas_sgemm / cblas_ssyrk call on the same device
operands (the grouped kernel runs the same per-tile MFMA routine and the
same K order); forced tile shapes give identical bits; graph replays
accumulate exactly like repeated calls; a sample is within the normwise
FP32 bound of the oracle.  TRSM problems: the bars of check_trsm below.
"""
import ctypes
import os

import numpy as np
import pytest

import l3cases as L
import mublis_amd as M
from mublis_amd.workloads import grouped_problems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    lib = M.load()
    assert lib.mublis_rt_init() == 0
    return lib


def single_calls(lib, torch, specs, initial, alpha, beta, times=1):
    outs = []
    for s, c0 in zip(specs, initial):
        C = c0.clone()
        for _ in range(times):
            if s["op"] == "gemm":
                lib.cblas_sgemm(102, 112 if s["ta"] else 111, 112 if s["tb"] else 111, s["m"],
                                s["n"], s["k"], alpha, M.ptr(s["A"]), s["lda"], M.ptr(s["B"]),
                                s["ldb"], beta, M.ptr(C), s["m"])
            elif s["op"] == "syrk":
                lib.cblas_ssyrk(102, 121 if s["upper"] else 122, 112 if s["trans"] else 111,
                                s["n"], s["k"], alpha, M.ptr(s["A"]), s["lda"], beta, M.ptr(C),
                                s["n"])
            else:
                lib.cblas_ssyr2k(102, 121 if s["upper"] else 122, 112 if s["trans"] else 111,
                                 s["n"], s["k"], alpha, M.ptr(s["A"]), s["lda"], M.ptr(s["B"]),
                                 s["ldb"], beta, M.ptr(C), s.get("ldc", s["n"]))
        outs.append(C)
    torch.cuda.synchronize()
    return outs


@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (0.75, -0.5), (1.0, 0.0), (0.0, 2.0)])
def test_grouped_equals_single_calls(lib, alpha, beta):
    torch = pytest.importorskip("torch")
    ops, keep, specs = grouped_problems(torch, seed=7, n_gemm=300, n_syrk=80,
                                        syrk_n=(32, 512), alpha=alpha, beta=beta,
                                        vary_flags=True)
    initial = [s["C"].clone() for s in specs]
    want = single_calls(lib, torch, specs, initial, alpha, beta)
    for s, c0 in zip(specs, initial):
        s["C"].copy_(c0)
    plan = M.GroupedPlan(ops)
    try:
        plan.execute()
        torch.cuda.synchronize()
        info = plan.info()
    finally:
        plan.close()
    assert info["items"] > 0 and sum(info["tiles"]) == len(ops)
    bad = [i for i, (s, w) in enumerate(zip(specs, want)) if not torch.equal(s["C"], w)]
    assert not bad, f"{len(bad)} problems differ, first {specs[bad[0]]['op']}"


def _syr2k_edge(torch, n, k, lda, ldb, trans, upper, seed):
    """One SYR2K problem with explicit leading dimensions (column-major
    storage as (cols, ld) tensors) for the extent-1 layout cases."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    cols = n if trans else k
    A = torch.rand((cols, lda), device="cuda", generator=g) * 2 - 1
    B = torch.rand((cols, ldb), device="cuda", generator=g) * 2 - 1
    C = torch.rand((n, n), device="cuda", generator=g) * 2 - 1
    op = M.SbatchOp(M.BATCH_SYR2K, trans, 0, 2 if upper else 1, n, n, k, 1.0, 1.0,
                    A.data_ptr(), lda, B.data_ptr(), ldb, C.data_ptr(), n)
    return op, dict(op="syr2k", n=n, k=k, trans=trans, upper=upper, A=A, B=B, C=C, lda=lda,
                    ldb=ldb)


def test_grouped_syr2k_equals_single_calls(lib):
    """Grouped SYR2K problems (both products as K segments of one tile loop)
    give the bits of single cblas_ssyr2k calls, including n = 1 with lda = 1
    and ldb = 5 (and the reverse, and transposed): the two products' operand
    strides differ there, and both paths read them with one layout after
    normalising the stride of the extent-1 dimension."""
    torch = pytest.importorskip("torch")
    ops, keep, specs = grouped_problems(torch, seed=12, n_gemm=30, n_syrk=10, n_syr2k=90,
                                        syrk_n=(1, 400), syrk_k=(1, 300), vary_flags=True)
    edges = [(1, 7, 1, 5, 0, 0), (1, 7, 5, 1, 0, 1), (1, 9, 9, 12, 1, 0), (1, 9, 12, 9, 1, 1),
             (1, 1, 1, 3, 0, 0), (2, 33, 2, 6, 0, 1), (40, 1, 40, 41, 1, 0)]
    for i, (n, k, lda, ldb, tr, up) in enumerate(edges):
        o, sp = _syr2k_edge(torch, n, k, lda, ldb, tr, up, seed=100 + i)
        ops.append(o)
        specs.append(sp)
    initial = [s["C"].clone() for s in specs]
    want = single_calls(lib, torch, specs, initial, 1.0, 1.0)
    for s, c0 in zip(specs, initial):
        s["C"].copy_(c0)
    plan = M.GroupedPlan(ops)
    try:
        plan.execute()
        torch.cuda.synchronize()
    finally:
        plan.close()
    bad = [i for i, (s, w) in enumerate(zip(specs, want)) if not torch.equal(s["C"], w)]
    assert not bad, [(specs[i]["op"], specs[i]["n"], specs[i]["k"]) for i in bad[:5]]
    # the edge problems against the oracle, so equal bits are also right bits
    orc = L.Cblas(L.ORACLE_LIB, prefix="orc_", returns_status=True)
    eps = np.finfo(np.float32).eps
    for s, c0 in zip(specs[-len(edges):], initial[-len(edges):]):
        n, k = s["n"], s["k"]
        a = np.asfortranarray(s["A"].cpu().numpy().T)
        b = np.asfortranarray(s["B"].cpu().numpy().T)
        ref = np.asfortranarray(c0.cpu().numpy().T)
        c00 = ref.copy(order="F")
        assert orc.call(np.float32, "syr2k", 102, 121 if s["upper"] else 122,
                        112 if s["trans"] else 111, n, k, 1.0, a, s["lda"], b, s["ldb"], 1.0,
                        ref, n) == 0
        got = s["C"].cpu().numpy().T
        scale = 2 * np.linalg.norm(a) * np.linalg.norm(b) + np.linalg.norm(c00)
        assert np.linalg.norm(got - ref) <= 2 * (k + 2) * eps * scale, (n, k)


def test_signed_zero_of_underflowing_products(lib):
    """Where every product of an element underflows the accumulator is -0,
    and an MFMA step that adds products of K's zero padding turns it into
    +0.  Every path gives the result of padding K to a multiple of the full
    slice depth (32 in FP32), so an element is +0 exactly when K is not a
    multiple of 32: single calls of every tile shape and depth (32-, 64-,
    64 half-depth, 128 half-depth) and the grouped kernel, whose last slice
    skips its steps of padding alone.  Compared bit for bit, signs of zero
    included."""
    torch = pytest.importorskip("torch")
    g = torch.Generator(device="cuda").manual_seed(5)
    ops, specs = [], []
    for (m, n) in ((40, 40), (200, 150), (2100, 2048), (3072, 3072)):
        for k in (16, 20, 28, 30, 32, 36, 64):
            A = -(torch.rand((k, m), device="cuda", generator=g) * 0.5 + 0.5) * 1e-30
            B = (torch.rand((n, k), device="cuda", generator=g) * 0.5 + 0.5) * 1e-30
            C = torch.ones((n, m), device="cuda")
            ops.append(M.SbatchOp(M.BATCH_GEMM, 0, 0, 0, m, n, k, 1.0, 0.0, A.data_ptr(), m,
                                  B.data_ptr(), k, C.data_ptr(), m))
            specs.append(dict(op="gemm", ta=0, tb=0, m=m, n=n, k=k, A=A, B=B, C=C, lda=m, ldb=k))
    initial = [s["C"].clone() for s in specs]
    want = single_calls(lib, torch, specs, initial, 1.0, 0.0)
    for s, w in zip(specs, want):
        k = s["k"]
        sign = 0 if k % 32 else -(2 ** 31)
        bits = w.view(torch.int32)
        assert bool((bits == sign).all()), (s["m"], s["n"], k, "single call")
    for s, c0 in zip(specs, initial):
        s["C"].copy_(c0)
    plan = M.GroupedPlan(ops)
    try:
        plan.execute()
        torch.cuda.synchronize()
    finally:
        plan.close()
    bad = [(s["m"], s["n"], s["k"]) for s, w in zip(specs, want)
           if not torch.equal(s["C"].view(torch.int32), w.view(torch.int32))]
    assert not bad, bad[:5]


def test_forced_tiles_and_replay(lib):
    """Every tile shape under every launch schedule gives the bits of the
    single calls, and a second replay accumulates again (beta = 1)."""
    torch = pytest.importorskip("torch")
    ops, keep, specs = grouped_problems(torch, seed=8, n_gemm=120, n_syrk=40,
                                        syrk_n=(32, 300), vary_flags=True)
    initial = [s["C"].clone() for s in specs]
    want = single_calls(lib, torch, specs, initial, 1.0, 1.0, times=2)
    for t in range(3):
        for v in range(2):
            for s, c0 in zip(specs, initial):
                s["C"].copy_(c0)
            plan = M.GroupedPlan(ops, flags=((t + 1) << 4) | ((v + 1) << 8))
            try:
                assert plan.info()["tiles"][t] == len(ops)
                plan.execute()
                torch.cuda.synchronize()
                plan.execute()  # beta = 1: the second replay accumulates again
                torch.cuda.synchronize()
            finally:
                plan.close()
            bad = [i for i, (s, w) in enumerate(zip(specs, want)) if not torch.equal(s["C"], w)]
            assert not bad, f"tile {t} schedule {v}: {len(bad)} problems differ"


def test_grouped_sample_vs_oracle(lib):
    torch = pytest.importorskip("torch")
    orc = L.Cblas(L.ORACLE_LIB, prefix="orc_", returns_status=True)
    ops, keep, specs = grouped_problems(torch, seed=9, n_gemm=40, n_syrk=12,
                                        syrk_n=(32, 256), alpha=0.5, beta=-1.0,
                                        vary_flags=True)
    initial = [s["C"].cpu().numpy().T.copy(order="F") for s in specs]
    assert lib.mublis_sgemm_grouped(ctypes.cast((M.SbatchOp * len(ops))(*ops),
                                                ctypes.c_void_p), len(ops)) == 0
    eps = np.finfo(np.float32).eps
    for s, c0 in zip(specs, initial):
        got = s["C"].cpu().numpy().T
        want = c0.copy(order="F")
        a = np.asfortranarray(s["A"].cpu().numpy().T)
        if s["op"] == "gemm":
            b = np.asfortranarray(s["B"].cpu().numpy().T)
            assert orc.call(np.float32, "gemm", 102, 112 if s["ta"] else 111,
                            112 if s["tb"] else 111, s["m"], s["n"], s["k"], 0.5, a, s["lda"],
                            b, s["ldb"], -1.0, want, s["m"]) == 0
            scale = 0.5 * np.linalg.norm(a) * np.linalg.norm(b) + np.linalg.norm(c0)
        else:
            assert orc.call(np.float32, "syrk", 102, 121 if s["upper"] else 122,
                            112 if s["trans"] else 111, s["n"], s["k"], 0.5, a, s["lda"], -1.0,
                            want, s["n"]) == 0
            scale = 0.5 * np.linalg.norm(a) ** 2 + np.linalg.norm(c0)
        assert np.linalg.norm(got - want) <= 2 * s["k"] * eps * scale


def _table(path):
    """{key: (choice, near_tie)} of a tuning table."""
    out = {}
    for ln in open(path):
        if ln.startswith("#") or not ln.strip():
            continue
        f = ln.split()
        tie = f[-1] == "~"
        if tie:
            f = f[:-1]
        out[tuple(f[:-1])] = (int(f[-1]), tie)
    return out


def test_autotuner_writes_and_uses_table(lib, tmp_path):
    """The table holds a schedule per tile class, a tile per GEMM bucket and a
    tile per SYRK bucket; plans follow it (GEMM and SYRK problems alike)."""
    torch = pytest.importorskip("torch")
    path = str(tmp_path / "tune.txt").encode()
    assert lib.mublis_grouped_autotune(path, 4, 5, 4, 2) == 0
    tab = _table(path.decode())
    gemm = {k[1:]: v[0] for k, v in tab.items() if k[0] == "gemm"}
    syrk = {k[1:]: v[0] for k, v in tab.items() if k[0] == "syrk"}
    sched = {k[1:]: v[0] for k, v in tab.items() if k[0] == "sched"}
    assert len(gemm) == 8 and len(syrk) == 4 and len(sched) == 3
    assert lib.mublis_grouped_load_table(path) == 12
    ops, keep, specs = grouped_problems(torch, seed=10, n_gemm=50, n_syrk=20,
                                        gemm_range=(16, 63), syrk_n=(16, 63),
                                        syrk_k=(16, 63))
    plan = M.GroupedPlan(ops)
    try:
        hist = plan.info()["tiles"]
    finally:
        plan.close()
    lg = lambda v: str(min(max(int(np.floor(np.log2(v))), 4), 13))  # noqa: E731
    expect = [0, 0, 0]
    for s in specs:
        if s["op"] == "gemm":
            expect[gemm[(lg(s["m"]), lg(s["n"]), lg(s["k"]))]] += 1
        else:
            expect[syrk[(lg(s["n"]), lg(s["k"]))]] += 1
    assert hist == expect
    lib.mublis_grouped_load_table(None)


def test_autotuner_is_repeatable(lib, tmp_path):
    """Two tuning runs over the same buckets choose the same schedule and tile
    for every entry that neither run marks as a near-tie (more than one
    candidate within the tuner's 3 % margin of the fastest; a near-tie takes
    the heuristic's choice if it is among them, else the first).  Tile
    entries are timed under the schedules the same run chose, so they are
    compared only when both runs chose the same schedules."""
    paths = [str(tmp_path / f"t{i}.txt").encode() for i in range(2)]
    for p in paths:
        assert lib.mublis_grouped_autotune(p, 5, 6, 4, 3) == 0
    a, b = (_table(p.decode()) for p in paths)
    assert a.keys() == b.keys()
    same_sched = all(a[k][0] == b[k][0] for k in a if k[0] == "sched")
    differ = [k for k in a if a[k][0] != b[k][0] and not (a[k][1] or b[k][1])
              and (k[0] == "sched" or same_sched)]
    assert not differ, ([(k, a[k], b[k]) for k in differ],
                        [open(p.decode()).read() for p in paths])
    lib.mublis_grouped_load_table(None)


# Grouped TRSM problems: (side, uplo, trans_a, diag, m, n, alpha).  Orders
# (m for Left, n for Right) cover one partial 64-row block, the three panel
# classes (32-column panels up to order 256, 16-column panels up to 512 and
# 1024) at their limits, problems solved by themselves (order > 1024) and
# alpha = 0 (B := 0, A unread).
TRSM_CASES = [
    ("L", "lo", 0, 0, 1, 1, 1.0),
    ("L", "up", 1, 1, 17, 33, -1.5),
    ("R", "lo", 0, 0, 40, 64, 0.75),
    ("R", "up", 1, 0, 100, 130, 1.0),
    ("L", "lo", 1, 0, 511, 16, 2.0),
    ("L", "up", 0, 1, 512, 70, 1.0),
    ("L", "lo", 0, 0, 513, 20, -1.0),
    ("R", "up", 0, 0, 90, 1024, 0.5),
    ("L", "up", 1, 0, 1100, 40, 1.0),
    ("R", "lo", 1, 1, 30, 1500, 1.0),
    ("L", "up", 0, 0, 200, 50, 0.0),
]


def trsm_batch(torch, seed):
    """SbatchOp list for TRSM_CASES on device operands (NaN outside each
    triangle and on unit diagonals), plus specs for checking."""
    rng = np.random.default_rng(seed)
    ops, specs = [], []
    for side, uplo, trans, diag, m, n, alpha in TRSM_CASES:
        left, lower = side == "L", uplo == "lo"
        order = m if left else n
        a = L.conditioned_triangular(rng, order, np.float32, "F", lower, diag == 1)
        b0 = L.rand(rng, (m, n), np.float32, "F")
        A = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
        B = torch.from_numpy(np.ascontiguousarray(b0.T)).cuda()
        ops.append(M.SbatchOp(M.BATCH_TRSM, trans, 0, 1 if lower else 2, m, n, 0, alpha, 0.0,
                              M.ptr(A), order, None, 1, M.ptr(B), m, 0 if left else 1, diag))
        specs.append(dict(left=left, lower=lower, trans=trans, unit=diag == 1, m=m, n=n,
                          alpha=alpha, a=a, b0=b0, A=A, B=B, order=order))
    return ops, specs


def trsm_single(lib, torch, s):
    """The same solve as one cblas_strsm call, on a copy of B0."""
    X = torch.from_numpy(np.ascontiguousarray(s["b0"].T)).cuda()
    lib.cblas_strsm(102, 141 if s["left"] else 142, 122 if s["lower"] else 121,
                    112 if s["trans"] else 111, 132 if s["unit"] else 131, s["m"], s["n"],
                    s["alpha"], M.ptr(s["A"]), s["order"], M.ptr(X), s["m"])
    torch.cuda.synchronize()
    return X.cpu().numpy().T


def check_trsm(lib, torch, specs):
    """Each grouped result: residual probe in long double within
    2 s eps ||A|| ||X|| ||v|| (s the triangle's order), and within a relative
    forward difference of 50 s eps of the single cblas_strsm call."""
    from test_gpu_trsm_pipe import residual
    eps = np.finfo(np.float32).eps
    bad = []
    for i, s in enumerate(specs):
        got = s["B"].cpu().numpy().T
        want = trsm_single(lib, torch, s)
        tag = f"case {i} {TRSM_CASES[i]}"
        if s["alpha"] == 0.0:
            if not (np.all(got == 0) and np.all(want == 0)):
                bad.append(f"{tag}: alpha = 0 must give B = 0")
            continue
        res, lim = residual(s["a"], got, s["b0"], s["alpha"], not s["left"], s["lower"],
                            s["trans"], s["unit"])
        if not np.isfinite(got).all() or not res <= lim:
            bad.append(f"{tag}: residual {res:.3e} > {lim:.3e}")
        diff = np.linalg.norm(got.astype(np.float64) - want) / np.linalg.norm(want)
        if not diff <= 50 * s["order"] * eps:
            bad.append(f"{tag}: grouped vs cblas_strsm {diff:.3e}")
    assert not bad, "\n".join(bad)


def test_grouped_trsm_mixed_with_gemm_and_syrk(lib):
    """TRSM problems in one plan with GEMM and SYRK problems: the TRSM
    results meet the residual and single-call bars, the GEMM/SYRK results
    stay bit-identical to single calls, and three more graph replays, each
    after eager work that restores the operands and a synchronisation, solve
    again to the same bits."""
    torch = pytest.importorskip("torch")
    gops, keep, gspecs = grouped_problems(torch, seed=11, n_gemm=60, n_syrk=16,
                                          syrk_n=(32, 300), vary_flags=True)
    initial = [s["C"].clone() for s in gspecs]
    want = single_calls(lib, torch, gspecs, initial, 1.0, 1.0)
    for s, c0 in zip(gspecs, initial):
        s["C"].copy_(c0)
    tops, tspecs = trsm_batch(torch, 5)
    ops = gops[:30] + tops + gops[30:]
    plan = M.GroupedPlan(ops)
    try:
        plan.execute()
        torch.cuda.synchronize()
        info = plan.info()
        first = [s["B"].clone() for s in tspecs]
        check_trsm(lib, torch, tspecs)
        bad = [i for i, (s, w) in enumerate(zip(gspecs, want)) if not torch.equal(s["C"], w)]
        assert not bad, f"{len(bad)} GEMM/SYRK problems differ"
        gfirst = [s["C"].clone() for s in gspecs]
        for rep in range(3):
            for s in tspecs:
                s["B"].copy_(torch.from_numpy(np.ascontiguousarray(s["b0"].T)).cuda())
            for s, c0 in zip(gspecs, initial):
                s["C"].copy_(c0)
            torch.cuda.synchronize()
            plan.execute()
            torch.cuda.synchronize()
            assert all(torch.equal(s["B"], f) for s, f in zip(tspecs, first)), f"replay {rep}"
            assert all(torch.equal(s["C"], f) for s, f in zip(gspecs, gfirst)), f"replay {rep}"
    finally:
        plan.close()
    assert sum(info["tiles"]) == len(gops)
    panels = sum((s["n"] if s["left"] else s["m"]) // 16 for s in tspecs)
    assert info["items"] >= panels


def test_grouped_trsm_workspace_launch(lib):
    """mublis_sgrouped_launch_ws (the serving loop's form) on the TRSM batch:
    the same bits as the captured plan."""
    torch = pytest.importorskip("torch")
    tops, tspecs = trsm_batch(torch, 6)
    arr = (M.SbatchOp * len(tops))(*tops)
    plan = M.GroupedPlan(tops)
    try:
        plan.execute()
        torch.cuda.synchronize()
    finally:
        plan.close()
    via_plan = [s["B"].clone() for s in tspecs]
    for s in tspecs:
        s["B"].copy_(torch.from_numpy(np.ascontiguousarray(s["b0"].T)).cuda())
    nbytes = lib.mublis_sgrouped_ws_bytes(ctypes.cast(arr, ctypes.c_void_p), len(tops))
    assert nbytes > 0
    host = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
    dev = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    assert lib.mublis_sgrouped_launch_ws(ctypes.cast(arr, ctypes.c_void_p), len(tops),
                                         host.data_ptr(), dev.data_ptr(), nbytes, None) == 0
    torch.cuda.synchronize()
    assert lib.mublis_rt_sync(lib.mublis_rt_stream()) == 0
    assert all(torch.equal(s["B"], w) for s, w in zip(tspecs, via_plan))
    check_trsm(lib, torch, tspecs)


def test_grouped_trsm_rejects_bad_arguments(lib):
    torch = pytest.importorskip("torch")
    A = torch.zeros((8, 8), device="cuda")
    B = torch.zeros((8, 8), device="cuda")
    good = dict(op=M.BATCH_TRSM, trans_a=0, trans_b=0, uplo=1, m=8, n=8, k=0, alpha=1.0,
                beta=0.0, a=M.ptr(A), lda=8, b=None, ldb=1, c=M.ptr(B), ldc=8, side=0, diag=0)
    for field, value in [("uplo", 0), ("side", 2), ("diag", 3), ("lda", 4), ("ldc", 2),
                         ("m", -1)]:
        o = M.SbatchOp(**dict(good, **{field: value}))
        h = ctypes.c_void_p()
        st = lib.mublis_sgrouped_plan_create(ctypes.byref(o), 1, 0, ctypes.byref(h))
        assert st == M.EINVAL, (field, st)
