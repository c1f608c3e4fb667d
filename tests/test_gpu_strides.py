"""GPU parity of the mublis_* stride API (include/mublis.h:85-129) when no
operand has a unit stride: every matrix is stored with row stride 2 and
column stride 2 * rows + 3 (or, for the second storage kind, row stride
3 * cols + 2 and column stride 3; both map every element to its own
address), so the library packs each operand densely
before its MFMA tiles read it (gemm.hip densify, pack.hip).  Host pointers
(staged over the whole strided footprint) and device pointers (used in
place) are both covered.

Bars: C (or B) within 2 (k + 2) eps (|alpha| ||op(A)|| ||op(B)|| + |beta|
||C0||) of the float64 value; GEMM / SYMM / SYRK also within twice that of
the oracle's generic driver run with the reference shim's descriptor
(frame/l3/gemm.c:14-42, symm.c:36-56, syrk.c:19-51); TRSM through the
residual ||op(A) X - alpha B|| <= 2 m eps ||A|| ||X||; every storage element
outside the referenced part (gaps, the unwritten triangle) bit-identical.
"""
import ctypes

import numpy as np
import pytest

import l3cases as L
import mublis_amd as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    lib = M.load()
    assert lib.mublis_rt_init() == 0
    return lib


@pytest.fixture(scope="module")
def orc_lib():
    return ctypes.CDLL(L.ORACLE_LIB)


def strided(rng, rows, cols, kind, dt, fill=None):
    """Flat storage of a rows x cols matrix with no unit stride; NaN in the
    gaps.  Returns (buf, rs, cs, dense float64 values)."""
    if kind == 0:
        rs, cs = 2, 2 * rows + 3
    else:
        rs, cs = 3 * cols + 2, 3
    size = (rows - 1) * rs + (cols - 1) * cs + 1
    buf = np.full(size, np.nan)
    vals = rng.standard_normal((rows, cols)) if fill is None else fill
    idx = np.add.outer(np.arange(rows) * rs, np.arange(cols) * cs)
    assert np.unique(idx).size == idx.size  # no two elements share storage
    buf[idx] = vals
    buf = buf.astype(dt)
    return buf, rs, cs, buf[idx].astype(np.float64), idx


def run(lib, fn, dev, bufs, args_fn):
    """Call with host buffers, or with device copies of them; returns the
    (possibly device-updated) buffers as numpy."""
    if not dev:
        return fn(*args_fn([b.ctypes.data for b in bufs])), bufs
    import torch
    d = [torch.from_numpy(b).cuda() for b in bufs]
    st = fn(*args_fn([x.data_ptr() for x in d]))
    torch.cuda.synchronize()
    return st, [x.cpu().numpy() for x in d]


def oracle_driver(orc_lib, dt, m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, beta, c, rs_c,
                  cs_c, desc):
    ct = ctypes.c_float if np.dtype(dt) == np.float32 else ctypes.c_double
    f = getattr(orc_lib, "orc_sdriver" if ct is ctypes.c_float else "orc_ddriver")
    f.argtypes = [L._I] * 3 + [ct] + [L._VP, L._I, L._I] * 2 + [ct, L._VP, L._I, L._I] + \
        [L._I] * 9
    out = c.copy()
    assert f(m, n, k, alpha, a.ctypes.data, rs_a, cs_a, b.ctypes.data, rs_b, cs_b, beta,
             out.ctypes.data, rs_c, cs_c, *desc) == 0
    return out


def check(got, c_in, idx, write, exact, want_orc, bound, who):
    res = got[idx].astype(np.float64)
    err = np.linalg.norm(np.where(write, res - exact, 0))
    assert np.isfinite(err) and err <= bound, (who, err, bound)
    if want_orc is not None:
        d = np.linalg.norm(np.where(write, res - want_orc[idx].astype(np.float64), 0))
        assert d <= 2 * bound, (who, d, bound)
    keep = np.ones(got.size, bool)
    keep[idx[write]] = False
    np.testing.assert_array_equal(got[keep], c_in[keep], err_msg=who)


EPS = {np.dtype(np.float32): np.finfo(np.float32).eps,
       np.dtype(np.float64): np.finfo(np.float64).eps}


@pytest.mark.parametrize("dev", [False, True], ids=["host", "device"])
@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_gemm_no_unit_stride(lib, orc_lib, dt, ta, tb, dev):
    rng = np.random.default_rng(40 + 2 * ta + tb)
    m, n, k = 131, 97, 75
    a, rsa, csa, av, _ = strided(rng, k if ta else m, m if ta else k, ta, dt)
    b, rsb, csb, bv, _ = strided(rng, n if tb else k, k if tb else n, 1 - tb, dt)
    c, rsc, csc, cv, idx = strided(rng, m, n, 0, dt)
    A = av.T if ta else av
    B = bv.T if tb else bv
    alpha, beta = 1.25, -0.5
    f = getattr(lib, "mublis_%sgemm" % ("s" if dt == np.float32 else "d"))
    st, (_, _, got) = run(lib, f, dev, [a, b, c.copy()], lambda p: (
        ta, tb, m, n, k, alpha, p[0], rsa, csa, p[1], rsb, csb, beta, p[2], rsc, csc))
    assert st == 0
    # the shim's stride swap (frame/l3/gemm.c:14-24) and an all-GENERAL descriptor
    ra, ca = (csa, rsa) if ta else (rsa, csa)
    rb, cb = (csb, rsb) if tb else (rsb, csb)
    want = oracle_driver(orc_lib, dt, m, n, k, alpha, a, ra, ca, b, rb, cb, beta, c, rsc, csc,
                         [0] * 9)
    exact = alpha * A @ B + beta * cv
    bound = 2 * (k + 2) * EPS[np.dtype(dt)] * (abs(alpha) * np.linalg.norm(A) *
                                               np.linalg.norm(B) + abs(beta) * np.linalg.norm(cv))
    check(got, c, idx, np.ones((m, n), bool), exact, want, bound, f"gemm {ta}{tb}")


@pytest.mark.parametrize("dev", [False, True], ids=["host", "device"])
@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("right,upper", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_symm_no_unit_stride(lib, orc_lib, dt, right, upper, dev):
    rng = np.random.default_rng(50 + 2 * right + upper)
    m, n = 109, 83
    s = n if right else m
    full = rng.standard_normal((s, s))
    sym = np.triu(full) + np.triu(full, 1).T if upper else np.tril(full) + np.tril(full, -1).T
    stored = np.where(np.triu(np.ones((s, s), bool)) if upper else
                      np.tril(np.ones((s, s), bool)), sym, np.nan)
    a, rsa, csa, _, _ = strided(rng, s, s, right, dt, fill=stored)
    sym = np.where(np.isnan(stored), 0, stored)
    sym = sym + (np.triu(sym, 1).T if upper else np.tril(sym, -1).T)
    sym = sym.astype(dt).astype(np.float64)
    b, rsb, csb, bv, _ = strided(rng, m, n, 1 - right, dt)
    c, rsc, csc, cv, idx = strided(rng, m, n, 0, dt)
    alpha, beta = 0.75, 0.25
    f = getattr(lib, "mublis_%ssymm" % ("s" if dt == np.float32 else "d"))
    st, (_, _, got) = run(lib, f, dev, [a, b, c.copy()], lambda p: (
        right, 2 if upper else 1, m, n, alpha, p[0], rsa, csa, p[1], rsb, csb, beta, p[2],
        rsc, csc))
    assert st == 0
    uplo = 2 if upper else 1
    if right:   # driver(m, n, n) with the symmetric operand in the B role
        want = oracle_driver(orc_lib, dt, m, n, n, alpha, b, rsb, csb, a, rsa, csa, beta, c,
                             rsc, csc, [0, 0, 0, 0, 0, 0, 1, uplo, 0])
        exact = alpha * bv @ sym + beta * cv
    else:
        want = oracle_driver(orc_lib, dt, m, n, m, alpha, a, rsa, csa, b, rsb, csb, beta, c,
                             rsc, csc, [0, 0, 0, 1, uplo, 0, 0, 0, 0])
        exact = alpha * sym @ bv + beta * cv
    bound = 2 * (s + 2) * EPS[np.dtype(dt)] * (abs(alpha) * np.linalg.norm(sym) *
                                               np.linalg.norm(bv) + abs(beta) * np.linalg.norm(cv))
    check(got, c, idx, np.ones((m, n), bool), exact, want, bound, f"symm {right}{upper}")


@pytest.mark.parametrize("dev", [False, True], ids=["host", "device"])
@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("two", [0, 1], ids=["syrk", "syr2k"])
@pytest.mark.parametrize("upper,trans", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_syrk_no_unit_stride(lib, orc_lib, dt, upper, trans, two, dev):
    rng = np.random.default_rng(60 + 4 * two + 2 * upper + trans)
    n, k = 143, 61
    a, rsa, csa, av, _ = strided(rng, k if trans else n, n if trans else k, trans, dt)
    b, rsb, csb, bv, _ = strided(rng, k if trans else n, n if trans else k, 1 - trans, dt)
    c, rsc, csc, cv, idx = strided(rng, n, n, 0, dt)
    A = av.T if trans else av
    B = bv.T if trans else bv
    alpha, beta = 1.0, 0.5
    p_ = "s" if dt == np.float32 else "d"
    uplo = 2 if upper else 1
    if two:
        f = getattr(lib, f"mublis_{p_}syr2k")
        st, (_, _, got) = run(lib, f, dev, [a, b, c.copy()], lambda p: (
            uplo, trans, n, k, alpha, p[0], rsa, csa, p[1], rsb, csb, beta, p[2], rsc, csc))
        exact = alpha * (A @ B.T + B @ A.T) + beta * cv
        want = None
        scale = 2 * np.linalg.norm(A) * np.linalg.norm(B)
    else:
        f = getattr(lib, f"mublis_{p_}syrk")
        st, (_, got) = run(lib, f, dev, [a, c.copy()], lambda p: (
            uplo, trans, n, k, alpha, p[0], rsa, csa, beta, p[1], rsc, csc))
        ra, ca = (csa, rsa) if trans else (rsa, csa)
        # frame/l3/syrk.c:19-51: ji from uplo, B = the stride-swapped A
        want = oracle_driver(orc_lib, dt, n, n, k, alpha, a, ra, ca, a, ca, ra, beta, c, rsc,
                             csc, [0, uplo, 0, 0, 0, 0, 0, 0, 0])
        exact = alpha * A @ A.T + beta * cv
        scale = np.linalg.norm(A) ** 2
    assert st == 0
    write = np.triu(np.ones((n, n), bool)) if upper else np.tril(np.ones((n, n), bool))
    bound = 2 * (k + 2) * EPS[np.dtype(dt)] * (abs(alpha) * scale +
                                               abs(beta) * np.linalg.norm(cv[write]))
    check(got, c, idx, write, exact, want, bound, f"syr{'2' if two else ''}k {upper}{trans}")


@pytest.mark.parametrize("dev", [False, True], ids=["host", "device"])
@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("routine", ["trmm", "trsm"])
@pytest.mark.parametrize("right,upper,trans,unit", [(0, 0, 0, 0), (0, 1, 1, 1), (1, 0, 1, 0),
                                                    (1, 1, 0, 1)])
def test_triangular_no_unit_stride(lib, dt, routine, right, upper, trans, unit, dev):
    rng = np.random.default_rng(70 + 8 * right + 4 * upper + 2 * trans + unit)
    m, n = 117, 71
    s = n if right else m
    tri = rng.uniform(-0.25, 0.25, (s, s)) / np.sqrt(s) * 4
    np.fill_diagonal(tri, rng.uniform(1.5, 2.5, s))
    keep = np.triu(np.ones((s, s), bool)) if upper else np.tril(np.ones((s, s), bool))
    stored = np.where(keep, tri, np.nan)
    if unit:
        np.fill_diagonal(stored, np.nan)
    a, rsa, csa, _, _ = strided(rng, s, s, trans, dt, fill=stored)
    T = np.where(keep, a.astype(np.float64)[np.add.outer(np.arange(s) * rsa,
                                                          np.arange(s) * csa)], 0)
    T = np.nan_to_num(T)
    if unit:
        np.fill_diagonal(T, 1.0)
    opT = T.T if trans else T
    b, rsb, csb, bv, idx = strided(rng, m, n, 1 - right, dt)
    alpha = 0.5
    f = getattr(lib, "mublis_%s%s" % ("s" if dt == np.float32 else "d", routine))
    st, (_, got) = run(lib, f, dev, [a, b.copy()], lambda p: (
        right, 2 if upper else 1, trans, unit, m, n, alpha, p[0], rsa, csa, p[1], rsb, csb))
    assert st == 0
    eps = EPS[np.dtype(dt)]
    x = got[idx].astype(np.float64)
    if routine == "trmm":
        exact = alpha * (bv @ opT if right else opT @ bv)
        bound = 2 * (s + 2) * eps * abs(alpha) * np.linalg.norm(T) * np.linalg.norm(bv)
        check(got, b, idx, np.ones((m, n), bool), exact, None, bound, "trmm")
    else:
        res = np.linalg.norm((x @ opT if right else opT @ x) - alpha * bv)
        assert res <= 2 * s * eps * np.linalg.norm(T) * np.linalg.norm(x), res
        keep_b = np.ones(got.size, bool)
        keep_b[idx] = False
        np.testing.assert_array_equal(got[keep_b], b[keep_b])
