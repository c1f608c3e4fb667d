"""GPU parity of the gfx950 L3 engine, called through its C ABI.

Bars (stated here so every result has one):
  * reference-harness and ragged golden cases: the reference harness's own
    element tolerance (f32 rtol 5e-4 / atol 5e-5, f64 rtol 2e-11 / atol
    1e-12, test/correctness/check_cblas.py:24-37) against the outputs of the
    reference's compiled linux_avx2 build, and bit equality of everything a
    routine must not write;
  * larger problems against the oracle, normwise:
      ||C - C*||_F <= 2 k eps (|alpha| ||A||_F ||B||_F + |beta| ||C0||_F)
    and for TRSM the residual ||op(A) X - alpha B||_F <= 2 m eps ||A||_F ||X||_F;
  * full-size workloads (DGEMM 8192 with op(A) = A', DTRSM 8192 x 4096,
    DSYRK 8192 / 2048): sampled entries against the oracle in long double with
    the elementwise bound |c - c*| <= 2 k eps sum_p |a_ip b_pj|, a residual
    probe for TRSM, and the untouched triangle for SYRK.
"""
import ctypes
import threading

import numpy as np
import pytest

import l3cases as L
import mublis_amd as M

pytestmark = pytest.mark.gpu

EPS = {np.dtype(np.float32): np.finfo(np.float32).eps,
       np.dtype(np.float64): np.finfo(np.float64).eps}
C_NORM = 2.0


@pytest.fixture(scope="module")
def lib():
    lib = M.load()
    st = lib.mublis_rt_init()
    assert st == 0, lib.mublis_rt_last_error()
    assert lib.mublis_rt_arch() == b"gfx950"
    return lib


@pytest.fixture(scope="module")
def prod(lib):
    return L.Cblas(lib)


@pytest.fixture(scope="module")
def orc():
    return L.Cblas(L.ORACLE_LIB, prefix="orc_", returns_status=True)


def normwise(actual, expected, bound_scale, k, dtype):
    err = np.linalg.norm(np.asarray(actual, np.float64) - np.asarray(expected, np.float64))
    lim = C_NORM * max(k, 1) * EPS[np.dtype(dtype)] * bound_scale
    return err, lim


# ------------------------------------------------ exact integer products --
@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("order,co", L.LAYOUTS)
def test_exact_small_integer_gemm(prod, dt, order, co):
    """Small-integer operands make every partial sum exact, so the MFMA
    tiles' operand and output lane layouts, edge tiles and the epilogue are
    checked bit for bit (shapes straddle the 16 x 16 MFMA and 128 x 128
    workgroup tiles)."""
    rng = np.random.default_rng(12)
    for (m, n, k) in [(16, 16, 4), (20, 18, 7), (130, 129, 33), (257, 131, 70)]:
        for ta in (0, 1):
            for tb in (0, 1):
                a = np.array(rng.integers(-3, 4, (k, m) if ta else (m, k)), dt, order=order)
                b = np.array(rng.integers(-3, 4, (n, k) if tb else (k, n)), dt, order=order)
                c0 = np.array(rng.integers(-3, 4, (m, n)), dt, order=order)
                c = c0.copy(order=order)
                prod.call(dt, "gemm", co, L.TR if ta else L.NOTR, L.TR if tb else L.NOTR, m,
                          n, k, 2.0, a, L.ld_of(a, order), b, L.ld_of(b, order), -1.0, c,
                          L.ld_of(c, order))
                A = a.astype(np.int64).T if ta else a.astype(np.int64)
                B = b.astype(np.int64).T if tb else b.astype(np.int64)
                np.testing.assert_array_equal(c, (2 * A @ B - c0.astype(np.int64)).astype(dt),
                                              err_msg=f"{m}x{n}x{k} ta{ta} tb{tb}")


# ---------------------------------------------------------- golden cases --
@pytest.mark.parametrize("name", ["l3_harness", "l3_ragged"])
def test_golden_vectors_of_the_reference_build(prod, name):
    failures = []
    for c, arr in L.load_golden(name):
        out = np.array(arr["c0"], copy=True, order=c["order"])
        prod.call(c["dtype"], c["routine"], *L.args_from_golden(c, arr, out))
        mask = L.golden_mask(c, out.shape[0])
        try:
            L.assert_ref_close(out, arr["avx2"], c["dtype"], mask, arr["c0"])
        except AssertionError as e:
            failures.append(f"{L.case_id(c)}: {str(e)[:300]}")
    assert not failures, "\n".join(failures[:10]) + f"\n({len(failures)} failing)"


# --------------------------------------------- larger problems vs oracle --
def _gemm_like_cases():
    out = []
    for dt in L.DTYPES:
        for (order, co) in L.LAYOUTS:
            for ta in (0, 1):
                for tb in (0, 1):
                    out.append(("gemm", dt, order, co, dict(ta=ta, tb=tb)))
            for right in (0, 1):
                for lower in (0, 1):
                    out.append(("symm", dt, order, co, dict(right=right, lower=lower)))
            for r in ("syrk", "syr2k"):
                for trans in (0, 1):
                    for lower in (0, 1):
                        out.append((r, dt, order, co, dict(trans=trans, lower=lower)))
    return out


@pytest.mark.parametrize("routine,dt,order,co,flags", _gemm_like_cases(),
                         ids=lambda v: str(v) if not isinstance(v, dict) else
                         "-".join(f"{k}{x}" for k, x in v.items()))
def test_products_vs_oracle(prod, orc, routine, dt, order, co, flags):
    rng = np.random.default_rng(2024)
    m, n, k = 300, 257, 190
    alpha, beta = 0.7, -1.3
    if routine == "gemm":
        a = L.rand(rng, (k, m) if flags["ta"] else (m, k), dt, order)
        b = L.rand(rng, (n, k) if flags["tb"] else (k, n), dt, order)
        c0 = L.rand(rng, (m, n), dt, order)
        args = lambda out: (co, L.TR if flags["ta"] else L.NOTR,  # noqa: E731
                            L.TR if flags["tb"] else L.NOTR, m, n, k, alpha, a,
                            L.ld_of(a, order), b, L.ld_of(b, order), beta, out,
                            L.ld_of(out, order))
        scale, kk, mask = abs(alpha) * np.linalg.norm(a) * np.linalg.norm(b), k, None
    elif routine == "symm":
        s = n if flags["right"] else m
        a = L.symmetric_storage(rng, s, dt, order, flags["lower"])
        b = L.rand(rng, (m, n), dt, order)
        c0 = L.rand(rng, (m, n), dt, order)
        args = lambda out: (co, L.RIGHT if flags["right"] else L.LEFT,  # noqa: E731
                            L.LO if flags["lower"] else L.UP, m, n, alpha, a,
                            L.ld_of(a, order), b, L.ld_of(b, order), beta, out,
                            L.ld_of(out, order))
        scale = abs(alpha) * np.linalg.norm(L.sym_dense(a, flags["lower"])) * np.linalg.norm(b)
        kk, mask = s, None
    else:
        nn = 300
        shape = (k, nn) if flags["trans"] else (nn, k)
        a = L.rand(rng, shape, dt, order)
        b = L.rand(rng, shape, dt, order)
        c0 = L.rand(rng, (nn, nn), dt, order)
        uplo = L.LO if flags["lower"] else L.UP
        tr = L.TR if flags["trans"] else L.NOTR
        if routine == "syrk":
            args = lambda out: (co, uplo, tr, nn, k, alpha, a, L.ld_of(a, order),  # noqa
                                beta, out, L.ld_of(out, order))
            scale = abs(alpha) * np.linalg.norm(a) ** 2
        else:
            args = lambda out: (co, uplo, tr, nn, k, alpha, a, L.ld_of(a, order), b,  # noqa
                                L.ld_of(b, order), beta, out, L.ld_of(out, order))
            scale = 2 * abs(alpha) * np.linalg.norm(a) * np.linalg.norm(b)
        kk = k
        mask = np.tril(np.ones((nn, nn), bool)) if flags["lower"] else \
            np.triu(np.ones((nn, nn), bool))
    scale += abs(beta) * np.linalg.norm(c0)
    want = np.array(c0, copy=True, order=order)
    assert orc.call(dt, routine, *args(want)) == 0
    got = np.array(c0, copy=True, order=order)
    prod.call(dt, routine, *args(got))
    if mask is not None:
        np.testing.assert_array_equal(got[~mask], c0[~mask])
        got, want = got[mask], want[mask]
    err, lim = normwise(got, want, scale, kk, dt)
    print(f"{routine} {dt} normwise error {err:.3e} (bound {lim:.3e})")
    assert err <= lim


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("routine", ["trmm", "trsm"])
def test_triangular_vs_oracle(prod, orc, dt, routine):
    rng = np.random.default_rng(99)
    m, n = 300, 129
    failures = []
    for (order, co) in L.LAYOUTS:
        for right, lower, trans, unit in __import__("itertools").product((0, 1), repeat=4):
            s = n if right else m
            a = L.conditioned_triangular(rng, s, dt, order, lower, unit)
            b0 = L.rand(rng, (m, n), dt, order)
            alpha = -1.5
            args = lambda out: (co, L.RIGHT if right else L.LEFT,  # noqa: E731
                                L.LO if lower else L.UP, L.TR if trans else L.NOTR,
                                L.UNIT if unit else L.NONUNIT, m, n, alpha, a,
                                L.ld_of(a, order), out, L.ld_of(out, order))
            want = np.array(b0, copy=True, order=order)
            assert orc.call(dt, routine, *args(want)) == 0
            got = np.array(b0, copy=True, order=order)
            prod.call(dt, routine, *args(got))
            T = L.tri_dense(a, lower, unit).astype(np.longdouble)
            op = T.T if trans else T
            if routine == "trmm":
                err, lim = normwise(got, want, abs(alpha) * np.linalg.norm(np.float64(T)) *
                                    np.linalg.norm(b0), s, dt)
            else:
                X = got.astype(np.longdouble)
                res = (X @ op if right else op @ X) - alpha * b0.astype(np.longdouble)
                err = float(np.linalg.norm(np.float64(res)))
                lim = C_NORM * s * EPS[np.dtype(dt)] * np.linalg.norm(np.float64(T)) * \
                    np.linalg.norm(got.astype(np.float64))
                fwd = np.linalg.norm(got.astype(np.float64) - want) / np.linalg.norm(want)
                if fwd > 100 * s * EPS[np.dtype(dt)]:
                    failures.append(f"{order} r{right} l{lower} t{trans} u{unit}: "
                                    f"forward error {fwd:.2e}")
            if not err <= lim:
                failures.append(f"{order} r{right} l{lower} t{trans} u{unit}: "
                                f"{err:.3e} > {lim:.3e}")
    assert not failures, "\n".join(failures)


# ---------------------------------------------------- contract corners --
@pytest.mark.parametrize("dt", L.DTYPES)
def test_special_cases(prod, dt):
    m, n, k = 70, 50, 40
    rng = np.random.default_rng(1)
    nan_a = np.full((m, k), np.nan, dt, order="F")
    nan_b = np.full((k, n), np.nan, dt, order="F")
    c0 = L.rand(rng, (m, n), dt, "F")
    c = c0.copy(order="F")
    prod.call(dt, "gemm", L.COL, L.NOTR, L.NOTR, m, n, k, 0.0, nan_a, m, nan_b, k, 2.0, c, m)
    np.testing.assert_array_equal(c, (2 * c0).astype(dt))
    a = L.rand(rng, (m, k), dt, "F")
    b = L.rand(rng, (k, n), dt, "F")
    c = np.full((m, n), np.nan, dt, order="F")
    prod.call(dt, "gemm", L.COL, L.NOTR, L.NOTR, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    assert np.isfinite(c).all()
    c = c0.copy(order="F")
    prod.call(dt, "gemm", L.COL, L.NOTR, L.NOTR, m, n, 0, 1.0, a, m, b, 1, 0.5, c, m)
    np.testing.assert_array_equal(c, (np.float64(0.5) * c0).astype(dt))
    c = c0.copy(order="F")
    prod.call(dt, "gemm", L.COL, L.NOTR, L.NOTR, 0, n, k, 1.0, a, 1, b, k, 0.5, c, 1)
    np.testing.assert_array_equal(c, c0)
    for r in ("trmm", "trsm"):
        y = c0.copy(order="F")
        tri = np.full((m, m), np.nan, dt, order="F")
        prod.call(dt, r, L.COL, L.LEFT, L.LO, L.NOTR, L.NONUNIT, m, n, 0.0, tri, m, y, m)
        assert np.all(y == 0)
    # ones in, k out (test/performance/benchmark_cblas.py:124-141)
    ones_a = np.ones((m, k), dt, order="F")
    ones_b = np.ones((k, n), dt, order="F")
    c = np.zeros((m, n), dt, order="F")
    prod.call(dt, "gemm", L.COL, L.NOTR, L.NOTR, m, n, k, 1.0, ones_a, m, ones_b, k, 0.0, c, m)
    assert np.all(c == k)


@pytest.mark.parametrize("dt", L.DTYPES)
def test_padded_leading_dimensions(prod, orc, dt):
    """ld larger than the matrix: gaps are neither read into the result nor
    overwritten."""
    rng = np.random.default_rng(3)
    m, n, k = 133, 77, 61
    A = L.rand(rng, (m + 9, k), dt, "F")
    B = L.rand(rng, (k + 3, n), dt, "F")
    C = L.rand(rng, (m + 5, n), dt, "F")
    want = C.copy(order="F")
    got = C.copy(order="F")
    args = lambda out: (L.COL, L.NOTR, L.NOTR, m, n, k, 1.0, A, m + 9, B, k + 3, 0.5, out,  # noqa
                        m + 5)
    orc.call(dt, "gemm", *args(want))
    prod.call(dt, "gemm", *args(got))
    np.testing.assert_array_equal(got[m:], C[m:])
    err, lim = normwise(got[:m], want[:m], np.linalg.norm(A) * np.linalg.norm(B) +
                        0.5 * np.linalg.norm(C), k, dt)
    assert err <= lim


def test_device_pointers_match_host_staging(lib, prod):
    """Device-resident operands are used in place and give the same bits as
    the staged host path (the kernels are deterministic)."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    m, n, k = 257, 300, 190
    a = L.rand(rng, (k, m), np.float64, "F")
    b = L.rand(rng, (k, n), np.float64, "F")
    c = np.zeros((m, n), np.float64, order="F")
    prod.call(np.float64, "gemm", L.COL, L.TR, L.NOTR, m, n, k, 1.0, a, k, b, k, 0.0, c, m)
    # column-major device storage: a torch tensor of shape (cols, rows)
    da = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
    dc = torch.zeros((n, m), dtype=torch.float64, device="cuda")
    assert lib.mublis_rt_is_device_ptr(M.ptr(da)) == 1
    prod.call(np.float64, "gemm", L.COL, L.TR, L.NOTR, m, n, k, 1.0, da, k, db, k, 0.0, dc, m)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dc.cpu().numpy().T, c)


def test_many_host_threads(prod):
    """Concurrent calls from several threads (one stream each) give the same
    bits as serial calls."""
    rng = np.random.default_rng(6)
    probs = []
    for t in range(6):
        m, n, k = 150 + 10 * t, 140, 130 + 7 * t
        probs.append((m, n, k, L.rand(rng, (m, k), np.float64, "F"),
                      L.rand(rng, (k, n), np.float64, "F")))
    serial = []
    for (m, n, k, a, b) in probs:
        c = np.zeros((m, n), order="F")
        prod.call(np.float64, "gemm", L.COL, L.NOTR, L.NOTR, m, n, k, 1.0, a, m, b, k, 0.0,
                  c, m)
        serial.append(c)
    outs = [np.zeros((p[0], p[1]), order="F") for p in probs]

    def work(i):
        m, n, k, a, b = probs[i]
        for _ in range(3):
            prod.call(np.float64, "gemm", L.COL, L.NOTR, L.NOTR, m, n, k, 1.0, a, m, b, k,
                      0.0, outs[i], m)

    th = [threading.Thread(target=work, args=(i,)) for i in range(len(probs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for o, s in zip(outs, serial):
        np.testing.assert_array_equal(o, s)


# ------------------------------------------------- plugin micro-kernels --
class _Sctx(ctypes.Structure):
    _fields_ = [("gemm_ukr", ctypes.c_void_p), ("gemmtrsml_ukr", ctypes.c_void_p),
                ("gemmtrsmu_ukr", ctypes.c_void_p), ("mr", ctypes.c_int),
                ("nr", ctypes.c_int), ("mc", ctypes.c_int), ("kc", ctypes.c_int),
                ("nc", ctypes.c_int)]


class _Ctx(ctypes.Structure):
    _fields_ = [("s", _Sctx), ("d", _Sctx)]


def test_context_micro_kernels(lib):
    """The gfx950 context honours the packed-panel micro-kernel contracts
    (include/mublis_instantiate.h:34-111), checked against numpy."""
    ctx = _Ctx()
    assert lib.mublis_get_context(ctypes.byref(ctx)) == 0
    d = ctx.d
    mr, nr = d.mr, d.nr
    assert d.mc % mr == 0 and d.nc % nr == 0 and d.kc % mr == 0 and d.kc % nr == 0
    VP, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    gemm = ctypes.CFUNCTYPE(None, I, D, VP, VP, D, VP, I, I, VP)(d.gemm_ukr)
    rng = np.random.default_rng(8)
    k = 37
    A = rng.standard_normal((mr, k))
    B = rng.standard_normal((k, nr))
    a = np.ascontiguousarray(A.T).ravel()          # a[p*MR + i]
    b = np.ascontiguousarray(B).ravel()            # b[p*NR + j]
    ldc = mr + 3
    C0 = rng.standard_normal((ldc, nr))
    cbuf = np.asfortranarray(C0.copy())
    gemm(k, 1.5, a.ctypes.data, b.ctypes.data, -0.25, cbuf.ctypes.data, 1, ldc, None)
    want = 1.5 * A @ B - 0.25 * C0[:mr]
    np.testing.assert_allclose(cbuf[:mr], want, rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(cbuf[mr:], C0[mr:])
    for lower, fp in ((1, d.gemmtrsml_ukr), (0, d.gemmtrsmu_ukr)):
        f = ctypes.CFUNCTYPE(None, I, D, VP, VP, VP, VP, VP, I, I, VP)(fp)
        T = np.tril(rng.standard_normal((mr, mr)) / mr, -1) if lower else \
            np.triu(rng.standard_normal((mr, mr)) / mr, 1)
        diag = 1.0 + rng.random(mr)
        a11 = T + np.diag(1.0 / diag)                 # diagonal pre-inverted
        a11c = np.asfortranarray(a11).ravel(order="F")
        B11 = rng.standard_normal((mr, nr))
        R = 0.5 * B11 - A @ B
        b11 = B11.ravel().copy()                    # row-major MR x NR, overwritten
        c11 = np.zeros((mr, nr), order="F")
        f(k, 0.5, a.ctypes.data, a11c.ctypes.data, b.ctypes.data, b11.ctypes.data,
          c11.ctypes.data, 1, mr, None)
        X = np.linalg.solve(T + np.diag(diag), R)
        np.testing.assert_allclose(c11, X, rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(b11.reshape(mr, nr), X, rtol=1e-11, atol=1e-11)


# ---------------------------------------------------- full-size workloads --
def _orc_fn(name, argtypes):
    lib = ctypes.CDLL(L.ORACLE_LIB)
    f = getattr(lib, name)
    f.argtypes = argtypes
    return f


def test_full_size_dgemm_8192_transA(lib, prod):
    """DGEMM M=N=K=8192, op(A)=A', alpha=1, beta=0, N(0,1) entries."""
    torch = pytest.importorskip("torch")
    n = 8192
    g = torch.Generator(device="cuda").manual_seed(314159)
    # column-major A (K x M) is the row-major tensor (M, K); same for B, C
    A = torch.randn((n, n), dtype=torch.float64, device="cuda", generator=g)
    B = torch.randn((n, n), dtype=torch.float64, device="cuda", generator=g)
    C = torch.full((n, n), float("nan"), dtype=torch.float64, device="cuda")
    prod.call(np.float64, "gemm", L.COL, L.TR, L.NOTR, n, n, n, 1.0, A, n, B, n, 0.0, C, n)
    torch.cuda.synchronize()
    assert torch.isfinite(C).all().item()
    rng = np.random.default_rng(0)
    I = np.array(rng.integers(0, n, 128), np.int32)
    J = np.array(rng.integers(0, n, 128), np.int32)
    # entries: C(i,j) = sum_p A(p,i) B(p,j); A(p,i) is tensor A[i, p]
    rowsA = A[torch.from_numpy(I).long().cuda()].cpu().numpy()   # 128 x K
    colsB = B[torch.from_numpy(J).long().cuda()].cpu().numpy()
    got = C[torch.from_numpy(J).long().cuda(), torch.from_numpy(I).long().cuda()].cpu().numpy()
    want = np.array([np.sum(rowsA[t].astype(np.longdouble) * colsB[t]) for t in range(128)],
                    dtype=np.float64)
    bound = C_NORM * n * EPS[np.dtype(np.float64)] * \
        np.array([np.sum(np.abs(rowsA[t] * colsB[t])) for t in range(128)])
    assert np.all(np.abs(got - want) <= bound), np.max(np.abs(got - want) / bound)


def test_full_size_dtrsm_8192x4096(lib, prod):
    """DTRSM Left/Lower/NoTrans/NonUnit, m=8192, nrhs=4096; off-diagonal
    N(0,1)/sqrt(m), diagonal 1+|N(0,1)|, B N(0,1).  Residual probe in long
    double: ||A (X v) - B v|| <= 2 m eps ||A||_F ||X||_F ||v||."""
    torch = pytest.importorskip("torch")
    m, nrhs = 8192, 4096
    g = torch.Generator(device="cuda").manual_seed(314159)
    At = torch.randn((m, m), dtype=torch.float64, device="cuda", generator=g) / np.sqrt(m)
    # column-major A(i,p) is At[p, i]; keep the lower triangle of A, i.e. the
    # upper triangle of At, and set the diagonal
    d = 1.0 + torch.randn(m, dtype=torch.float64, device="cuda", generator=g).abs()
    At = torch.triu(At, 1) + torch.diag(d)
    Bt = torch.randn((nrhs, m), dtype=torch.float64, device="cuda", generator=g)
    B0 = Bt.clone()
    prod.call(np.float64, "trsm", L.COL, L.LEFT, L.LO, L.NOTR, L.NONUNIT, m, nrhs, 1.0, At, m,
              Bt, m)
    torch.cuda.synchronize()
    A_np = At.cpu().numpy()    # row-major (p, i) = column-major A of ld m
    X_np = Bt.cpu().numpy()    # row-major (j, i) = column-major X
    B_np = B0.cpu().numpy()
    v = np.random.default_rng(1).standard_normal(nrhs)
    xv = np.zeros(m)
    bv = np.zeros(m)
    ay = np.zeros(m)
    gemv = _orc_fn("orc_dgemv_ld", [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_void_p])
    trmv = _orc_fn("orc_dtrmv_ld", [ctypes.c_int] * 4 + [ctypes.c_void_p, ctypes.c_int,
                                                         ctypes.c_void_p, ctypes.c_void_p])
    gemv(m, nrhs, X_np.ctypes.data, m, v.ctypes.data, xv.ctypes.data)
    gemv(m, nrhs, B_np.ctypes.data, m, v.ctypes.data, bv.ctypes.data)
    trmv(L.LO, L.NOTR, L.NONUNIT, m, A_np.ctypes.data, m, xv.ctypes.data, ay.ctypes.data)
    res = np.linalg.norm(ay - bv)
    lim = C_NORM * m * EPS[np.dtype(np.float64)] * np.linalg.norm(A_np) * \
        np.linalg.norm(X_np) * np.linalg.norm(v)
    print(f"DTRSM residual {res:.3e} bound {lim:.3e}")
    assert res <= lim


def test_full_size_dsyrk_8192_2048(lib, prod):
    """DSYRK lower, n=8192, k=2048: sampled lower entries against long
    double sums; the strict upper triangle keeps its bits."""
    torch = pytest.importorskip("torch")
    n, k = 8192, 2048
    g = torch.Generator(device="cuda").manual_seed(314159)
    At = torch.randn((k, n), dtype=torch.float64, device="cuda", generator=g)  # A n x k col-major
    Ct = torch.randn((n, n), dtype=torch.float64, device="cuda", generator=g)
    C0 = Ct.clone()
    prod.call(np.float64, "syrk", L.COL, L.LO, L.NOTR, n, k, 1.0, At, n, 0.0, Ct, n)
    torch.cuda.synchronize()
    # column-major C(i,j) is Ct[j, i]; lower (j <= i) is the upper triangle of Ct
    upper_untouched = torch.equal(torch.tril(Ct, -1), torch.tril(C0, -1))
    assert upper_untouched
    rng = np.random.default_rng(2)
    I = rng.integers(0, n, 256)
    J = rng.integers(0, n, 256)
    I, J = np.maximum(I, J), np.minimum(I, J)
    Anp = At.cpu().numpy()    # (p, i) -> A(i, p)
    got = Ct.cpu().numpy()[J, I]
    want = np.array([np.sum(Anp[:, i].astype(np.longdouble) * Anp[:, j]) for i, j in zip(I, J)],
                    dtype=np.float64)
    bound = C_NORM * k * EPS[np.dtype(np.float64)] * \
        np.array([np.sum(np.abs(Anp[:, i] * Anp[:, j])) for i, j in zip(I, J)])
    assert np.all(np.abs(got - want) <= bound)
