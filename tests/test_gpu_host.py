"""GPU tests of host-pointer operands: the pinned staging pool and the GEMM
that streams C by column panels (runtime/hostxfer.hip).

Neither changes an element's arithmetic -- the panels cut only N, which no
tile's k order depends on -- so every host-pointer call must give the same
bits as the same call on device copies of its operands (itself checked
against the oracle by the other GPU tests), and must leave every host
storage element outside the referenced part (column gaps, the unwritten
triangle) bit-identical.  Sizes put the operands above the 16 MiB at which
the panel path starts (host_gemm_ok in frame/l3.c) and the 4 MiB at which
staging goes through the pinned pool, with ragged panels (N not a multiple
of the panel width, M not a multiple of the tile).  A float64 matrix
product checks one case outright.
"""
import numpy as np
import pytest
import torch

import mublis_amd as M

pytestmark = pytest.mark.gpu

COL, ROW, NOTR, TR, UP, LO, NONUNIT, LEFT = 102, 101, 111, 112, 121, 122, 131, 141


@pytest.fixture(scope="module")
def lib():
    lib = M.load()
    assert lib.mublis_rt_init() == 0
    return lib


def host(rng, rows, cols, ld, dt):
    """Column-major rows x cols in a flat buffer of leading dimension ld;
    NaN sentinels in the gap rows."""
    buf = np.full(ld * cols, np.nan, dtype=dt)
    v = buf.reshape(cols, ld)
    v[:, :rows] = rng.uniform(-1, 1, (cols, rows)).astype(dt)
    return buf


def dev_copy(buf):
    return torch.from_numpy(buf.copy()).cuda()


def same_bits(x, y):
    return np.array_equal(x.view(np.uint8), y.view(np.uint8))


GEMM_CASES = [
    # dtype, order, ta, tb, m, n, k, beta, ldc pad, which operands are host
    (np.float64, COL, NOTR, NOTR, 1500, 1300, 1100, 0.0, 0, "abc"),
    (np.float64, COL, TR, NOTR, 2100, 1999, 700, 0.5, 7, "abc"),
    (np.float64, COL, NOTR, TR, 1024, 3000, 900, 1.0, 0, "abc"),
    (np.float64, ROW, NOTR, NOTR, 1700, 1200, 1000, -0.5, 3, "abc"),
    (np.float32, COL, NOTR, NOTR, 2500, 2300, 1500, 0.5, 5, "abc"),
    (np.float32, COL, TR, TR, 1900, 2100, 1300, 0.0, 0, "abc"),
    (np.float64, COL, NOTR, NOTR, 1500, 1300, 1100, 0.5, 2, "c"),
    (np.float64, COL, NOTR, NOTR, 1500, 1300, 1100, 0.5, 0, "bc"),
]


@pytest.mark.parametrize("dt,order,ta,tb,m,n,k,beta,pad,hosts", GEMM_CASES)
def test_host_gemm_matches_device_call(lib, dt, order, ta, tb, m, n, k, beta, pad, hosts):
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    col = order == COL
    # stored shapes (rows x cols in the column-major sense of the storage)
    if col:
        ar, ac = (k, m) if ta == TR else (m, k)
        br, bc = (n, k) if tb == TR else (k, n)
        cr, cc = m, n
    else:
        ar, ac = (m, k) if ta == TR else (k, m)
        br, bc = (k, n) if tb == TR else (n, k)
        cr, cc = n, m
    lda, ldb, ldc = ar + 1, br + 2, cr + pad
    a = host(rng, ar, ac, lda, dt)
    b = host(rng, br, bc, ldb, dt)
    c = host(rng, cr, cc, ldc, dt)
    f = lib.cblas_dgemm if dt == np.float64 else lib.cblas_sgemm
    da, db, dc = dev_copy(a), dev_copy(b), dev_copy(c)
    f(order, ta, tb, m, n, k, 1.0, M.ptr(da), lda, M.ptr(db), ldb, beta, M.ptr(dc), ldc)
    torch.cuda.synchronize()
    want = dc.cpu().numpy()
    pa = a.ctypes.data if "a" in hosts else M.ptr(da)
    pb = b.ctypes.data if "b" in hosts else M.ptr(db)
    c0 = c.copy()
    f(order, ta, tb, m, n, k, 1.0, pa, lda, pb, ldb, beta, c.ctypes.data, ldc)
    assert same_bits(c, want)
    # gap rows untouched
    gap = c.reshape(cc, ldc)[:, cr:]
    assert same_bits(gap, c0.reshape(cc, ldc)[:, cr:])


def test_host_gemm_against_float64(lib):
    rng = np.random.default_rng(5)
    m, n, k = 1337, 2051, 1111
    a = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    b = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    c = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    want = a @ b + 0.5 * c
    lib.cblas_dgemm(COL, NOTR, NOTR, m, n, k, 1.0, a.ctypes.data, m, b.ctypes.data, k, 0.5,
                    c.ctypes.data, m)
    err = np.abs(c - want).max()
    bound = 2 * (k + 2) * np.finfo(np.float64).eps * (k + 1)
    assert err <= bound, (err, bound)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_pinned_staging_syrk_trsm(lib, dt):
    """SYRK (C's triangle and gaps) and TRSM (B in place) with host operands
    whose footprints go through the pinned pool: same bits as on device
    copies, the unwritten triangle and the gaps unchanged."""
    rng = np.random.default_rng(11)
    n, k = 1800, 900
    lda, ldc = n + 3, n + 5
    a = host(rng, n, k, lda, dt)
    c = host(rng, n, n, ldc, dt)
    f = lib.cblas_dsyrk if dt == np.float64 else lib.cblas_ssyrk
    da, dc = dev_copy(a), dev_copy(c)
    f(COL, LO, NOTR, n, k, 1.0, M.ptr(da), lda, 0.5, M.ptr(dc), ldc)
    torch.cuda.synchronize()
    want = dc.cpu().numpy()
    f(COL, LO, NOTR, n, k, 1.0, a.ctypes.data, lda, 0.5, c.ctypes.data, ldc)
    assert same_bits(c, want)
    # TRSM: diagonally dominant lower triangle
    m, nrhs = 2000, 700
    t = host(rng, m, m, m, dt)
    tv = t.reshape(m, m)
    tv[np.diag_indices(m)] = float(m)
    bb = host(rng, m, nrhs, m + 1, dt)
    g = lib.cblas_dtrsm if dt == np.float64 else lib.cblas_strsm
    dtt, dbb = dev_copy(t), dev_copy(bb)
    g(COL, LEFT, LO, NOTR, NONUNIT, m, nrhs, 1.0, M.ptr(dtt), m, M.ptr(dbb), m + 1)
    torch.cuda.synchronize()
    want = dbb.cpu().numpy()
    t0 = t.copy()
    g(COL, LEFT, LO, NOTR, NONUNIT, m, nrhs, 1.0, t.ctypes.data, m, bb.ctypes.data, m + 1)
    assert same_bits(bb, want)
    assert same_bits(t, t0)


def test_staging_2d_round_trip(lib):
    """mublis_rt_h2d_2d / _d2h_2d on blocks taller than a pinned block (row
    pieces), wider than one (column pieces) and ragged: what comes back is
    what was sent, host gap rows untouched."""
    import ctypes
    rng = np.random.default_rng(3)
    h2d, d2h = lib.mublis_rt_h2d_2d, lib.mublis_rt_d2h_2d
    for f in (h2d, d2h):
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                      ctypes.c_int64, ctypes.c_size_t, ctypes.c_void_p]
        f.restype = ctypes.c_int
    for rows, cols, ld in ((5_000_000, 2, 5_000_003), (1000, 9000, 1003), (3, 5, 4)):
        src = rng.standard_normal(ld * cols)
        dev = torch.empty(rows * cols, dtype=torch.float64, device="cuda")
        s = lib.mublis_rt_stream()
        assert h2d(M.ptr(dev), src.ctypes.data, rows, cols, ld, 8, s) == 0
        lib.mublis_rt_sync(s)
        want = src.reshape(cols, ld)[:, :rows]
        assert np.array_equal(dev.cpu().numpy().reshape(cols, rows), want)
        back = np.full(ld * cols, -7.0)
        assert d2h(back.ctypes.data, M.ptr(dev), rows, cols, ld, 8, s) == 0
        bv = back.reshape(cols, ld)
        assert np.array_equal(bv[:, :rows], want)
        assert (bv[:, rows:] == -7.0).all()
