"""GPU checks of the pipelined GEMMTRSM schedule (narrow B: one launch, block
rows chained by device-side flags) against the recursive schedule and the
oracle, through the cblas_?trsm C ABI.

Bars:
* residual probe in long double (oracle): for a random vector v,
  ||op(A) (X v) - alpha B v|| <= 2 s eps ||A||_F ||X||_F ||v|| (Left), and
  ||X (op(A) v) - alpha B v|| likewise (Right), s the order of A and eps of
  the data type;
* the pipelined and the recursive schedule agree to a relative forward
  difference of 50 s eps (both solve the same well-conditioned system);
* ragged shapes: m not a multiple of the 64-row block, n not a multiple of
  the column group (16, 32 or 64 columns), a single right-hand side; the
  SHAPES are ones the default dispatch sends to the pipeline (n <= 128,
  n <= 2048 and m n <= 2^22, or n <= 768 and m n <= 2^23 for double; float
  past 128 columns with m n >= 2^21 takes GEMM leaves wherever they apply,
  so 1536 x 200 and 3000 x 300 are float pipelines too, while float
  2100 x 1500 takes the GEMM leaves); the wider ones go through the
  recursive schedule with GEMM leaves (1024-row diagonal blocks inverted by
  doubling the 64-row inverses with batched GEMMs).
"""
import ctypes
import itertools

import numpy as np
import pytest

import l3cases as L
import mublis_amd as M

pytestmark = pytest.mark.gpu

EPS = {np.dtype(np.float32): np.finfo(np.float32).eps,
       np.dtype(np.float64): np.finfo(np.float64).eps}


@pytest.fixture(scope="module")
def lib():
    lib = M.load()
    assert lib.mublis_rt_init() == 0
    return lib


@pytest.fixture(scope="module")
def prod(lib):
    return L.Cblas(lib)


def _orc(name, argtypes):
    f = getattr(ctypes.CDLL(L.ORACLE_LIB), name)
    f.argtypes = argtypes
    return f


_VP, _I = ctypes.c_void_p, ctypes.c_int
GEMV = _orc("orc_dgemv_ld", [_I, _I, _VP, _I, _VP, _VP])
TRMV = _orc("orc_dtrmv_ld", [_I, _I, _I, _I, _VP, _I, _VP, _VP])


def residual(a, x, b0, alpha, right, lower, trans, unit, seed=0):
    """||probe|| and its bound, in long double through the oracle."""
    m, n = x.shape
    s = a.shape[0]
    A = np.asfortranarray(a, dtype=np.float64)
    X = np.asfortranarray(x, dtype=np.float64)
    B = np.asfortranarray(b0, dtype=np.float64)
    uplo, ta, dg = (L.LO if lower else L.UP), (L.TR if trans else L.NOTR), \
        (L.UNIT if unit else L.NONUNIT)
    v = np.random.default_rng(seed).standard_normal(n)
    bv = np.zeros(m)
    GEMV(m, n, B.ctypes.data, m, v.ctypes.data, bv.ctypes.data)
    lhs = np.zeros(m)
    if right:
        w = np.zeros(n)
        TRMV(uplo, ta, dg, n, A.ctypes.data, s, v.ctypes.data, w.ctypes.data)
        GEMV(m, n, X.ctypes.data, m, w.ctypes.data, lhs.ctypes.data)
    else:
        xv = np.zeros(m)
        GEMV(m, n, X.ctypes.data, m, v.ctypes.data, xv.ctypes.data)
        TRMV(uplo, ta, dg, m, A.ctypes.data, s, xv.ctypes.data, lhs.ctypes.data)
    res = np.linalg.norm(lhs - alpha * bv)
    T = L.tri_dense(a, lower, unit)
    lim = 2 * s * EPS[np.dtype(x.dtype)] * np.linalg.norm(T) * np.linalg.norm(X) * \
        np.linalg.norm(v)
    return res, lim


SHAPES = [(65, 1), (200, 16), (257, 17), (1000, 33), (640, 128), (1100, 72), (2100, 100),
          (1536, 200), (3000, 300), (2100, 1500)]


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("m,n", SHAPES)
def test_pipelined_vs_recursive_left(lib, prod, dt, m, n):
    rng = np.random.default_rng(m * 1000 + n)
    failures = []
    for lower, unit in itertools.product((True, False), repeat=2):
        trans = (m + n + lower) % 2 == 1
        a = L.conditioned_triangular(rng, m, dt, "F", lower, unit)
        b0 = L.rand(rng, (m, n), dt, "F")
        alpha = 0.75
        outs = []
        for max_n in (-1, 0):  # default (pipelined for this n), then recursive only
            old = lib.mublis_gfx950_trsm_pipe_max_n(max_n)
            try:
                x = np.array(b0, copy=True, order="F")
                prod.call(dt, "trsm", L.COL, L.LEFT, L.LO if lower else L.UP,
                          L.TR if trans else L.NOTR, L.UNIT if unit else L.NONUNIT, m, n,
                          alpha, a, m, x, m)
            finally:
                lib.mublis_gfx950_trsm_pipe_max_n(old)
            outs.append(x)
        pipe, rec = outs
        res, lim = residual(a, pipe, b0, alpha, False, lower, trans, unit)
        diff = np.linalg.norm(pipe.astype(np.float64) - rec) / np.linalg.norm(rec)
        tag = f"l{int(lower)} u{int(unit)} t{int(trans)}"
        if not np.isfinite(pipe).all() or not res <= lim:
            failures.append(f"{tag}: residual {res:.3e} > {lim:.3e}")
        if not diff <= 50 * m * EPS[np.dtype(dt)]:
            failures.append(f"{tag}: pipelined vs recursive {diff:.3e}")
    assert not failures, "\n".join(failures)


@pytest.mark.parametrize("dt", L.DTYPES)
def test_pipelined_right_side(lib, prod, dt):
    """Right side: the frame transposes B's view (unit stride along the
    columns of the solve), which the pipeline stages along that stride."""
    rng = np.random.default_rng(4)
    m, n = 96, 700
    assert lib.mublis_gfx950_trsm_pipe_max_n(-1) >= m
    failures = []
    for lower, trans, unit in itertools.product((True, False), repeat=3):
        a = L.conditioned_triangular(rng, n, dt, "F", lower, unit)
        b0 = L.rand(rng, (m, n), dt, "F")
        x = np.array(b0, copy=True, order="F")
        prod.call(dt, "trsm", L.COL, L.RIGHT, L.LO if lower else L.UP,
                  L.TR if trans else L.NOTR, L.UNIT if unit else L.NONUNIT, m, n, -1.25, a,
                  n, x, m)
        res, lim = residual(a, x, b0, -1.25, True, lower, trans, unit)
        if not np.isfinite(x).all() or not res <= lim:
            failures.append(f"l{lower} t{trans} u{unit}: {res:.3e} > {lim:.3e}")
    assert not failures, "\n".join(failures)


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("m,n", [(2000, 2200), (1100, 2100)])
def test_recursive_schedule_with_gemm_leaves_wide(lib, prod, dt, m, n):
    """Wide B (n >= 256): the 1024-row diagonal blocks are inverted by
    doubling the 64-row inverses with batched GEMMs and the recursion, split
    at multiples of 1024 (2000 -> 1024 + 976; 1100 -> 1024 + 76; B wider
    than 2048 columns, so that double solves take this schedule too), ends
    in GEMM leaves; compared with the recursive schedule down to 64-row leaves
    (pipeline disabled) and checked by the residual probe."""
    rng = np.random.default_rng(21)
    failures = []
    for lower, trans in ((True, False), (False, True)):
        a = L.conditioned_triangular(rng, m, dt, "F", lower, False)
        b0 = L.rand(rng, (m, n), dt, "F")
        outs = []
        for max_n in (-1, 0):
            old = lib.mublis_gfx950_trsm_pipe_max_n(max_n)
            try:
                x = np.array(b0, copy=True, order="F")
                prod.call(dt, "trsm", L.COL, L.LEFT, L.LO if lower else L.UP,
                          L.TR if trans else L.NOTR, L.NONUNIT, m, n, 2.0, a, m, x, m)
            finally:
                lib.mublis_gfx950_trsm_pipe_max_n(old)
            outs.append(x)
        res, lim = residual(a, outs[0], b0, 2.0, False, lower, trans, False)
        diff = np.linalg.norm(outs[0].astype(np.float64) - outs[1]) / np.linalg.norm(outs[1])
        if not np.isfinite(outs[0]).all() or not res <= lim:
            failures.append(f"l{lower} t{trans}: residual {res:.3e} > {lim:.3e}")
        if not diff <= 50 * m * EPS[np.dtype(dt)]:
            failures.append(f"l{lower} t{trans}: vs 64-row leaves {diff:.3e}")
    assert not failures, "\n".join(failures)


def test_pipe_threshold_knob(lib):
    old = lib.mublis_gfx950_trsm_pipe_max_n(100000)
    try:
        assert lib.mublis_gfx950_trsm_pipe_max_n(-1) == 4096  # clamped
    finally:
        lib.mublis_gfx950_trsm_pipe_max_n(old)
    assert lib.mublis_gfx950_trsm_pipe_max_n(-1) == old


GEMM_LEAF_CASES = [
    # (m, n, order, right): the solve's order is n on the right side; the
    # inverted blocks have 1024 rows, so the doubling's last pairs are short.
    # Both types take GEMM leaves in the last two (float: B wider than 128
    # columns with m n >= 2^21 and order above 1024; double: m n > 2^22
    # beyond 768 columns) and the pipeline in the others.
    (1500, 300, "F", False),   # leaves 1024 + 476
    (2600, 256, "F", False),   # 2048 + 552
    (777, 513, "C", False),    # row-major B, pipelined leaves (m < 1024)
    (300, 1500, "F", True),    # right side: B's transposed view, 1500-row solve
    (257, 1111, "C", True),    # row-major, 1024 + 87
    (128, 2100, "F", True),    # 1024 + 1024 + 52: a 52-row last block
    (1100, 64, "F", False),    # pipeline for both types (64 columns)
    (2050, 100, "C", False),   # pipeline for both types, 2048 + 2 rows
    (2600, 1700, "F", False),  # GEMM leaves for both: 2048 + 552
    (2100, 2100, "C", True),   # right side, row-major: 2048 + 52
]


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("m,n,order,right", GEMM_LEAF_CASES)
def test_gemm_leaves_ragged(lib, prod, dt, m, n, order, right):
    """The cases above on ragged shapes: short last 1024-row blocks (the
    doubling's short last pairs), row-major storage and the right side;
    every uplo / trans / diag combination, against the recursive schedule
    down to 64-row leaves and the residual probe."""
    co = L.ROW if order == "C" else L.COL
    s_ = n if right else m
    rng = np.random.default_rng(m * 7 + n)
    failures = []
    for lower, trans, unit in itertools.product((True, False), repeat=3):
        a = L.conditioned_triangular(rng, s_, dt, order, lower, unit)
        b0 = L.rand(rng, (m, n), dt, order)
        alpha = -0.5
        lda = s_
        ldb = n if order == "C" else m
        outs = []
        for max_n in (-1, 0):
            old = lib.mublis_gfx950_trsm_pipe_max_n(max_n)
            try:
                x = np.array(b0, copy=True, order=order)
                prod.call(dt, "trsm", co, L.RIGHT if right else L.LEFT,
                          L.LO if lower else L.UP, L.TR if trans else L.NOTR,
                          L.UNIT if unit else L.NONUNIT, m, n, alpha, a, lda, x, ldb)
            finally:
                lib.mublis_gfx950_trsm_pipe_max_n(old)
            outs.append(x)
        # logical matrices: the probe copies them to column-major itself
        res, lim = residual(a, outs[0], b0, alpha, right, lower, trans, unit)
        diff = np.linalg.norm(outs[0].astype(np.float64) - outs[1]) / np.linalg.norm(outs[1])
        tag = f"l{int(lower)} t{int(trans)} u{int(unit)}"
        if not np.isfinite(outs[0]).all() or not res <= lim:
            failures.append(f"{tag}: residual {res:.3e} > {lim:.3e}")
        if not diff <= 50 * s_ * EPS[np.dtype(dt)]:
            failures.append(f"{tag}: vs 64-row leaves {diff:.3e}")
    assert not failures, "\n".join(failures)


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("lower", [True, False])
def test_gemm_leaves_without_unit_stride(lib, dt, lower):
    """mublis_?trsm on an A with no unit stride (every other row of a
    2m x m column-major array: rs = 2, cs = 2m): the GEMM leaves' inverses
    then come from the batched pipelined identity solves instead of the
    doubling GEMMs; against the 64-row-leaf schedule and the residual
    probe."""
    m, n = 1100, 300
    rng = np.random.default_rng(1100 + lower)
    a = L.conditioned_triangular(rng, m, dt, "F", lower, False)
    big = np.zeros((2 * m, m), dtype=dt, order="F")
    big[0::2, :] = a
    b0 = L.rand(rng, (m, n), dt, "F")
    f = lib.mublis_dtrsm if np.dtype(dt) == np.float64 else lib.mublis_strsm
    outs = []
    for max_n in (-1, 0):
        old = lib.mublis_gfx950_trsm_pipe_max_n(max_n)
        try:
            x = np.array(b0, copy=True, order="F")
            st = f(0, 1 if lower else 2, 0, 0, m, n, 1.5, M.ptr(big), 2, 2 * m, M.ptr(x), 1, m)
        finally:
            lib.mublis_gfx950_trsm_pipe_max_n(old)
        assert st == 0
        outs.append(x)
    res, lim = residual(a, outs[0], b0, 1.5, False, lower, False, False)
    diff = np.linalg.norm(outs[0].astype(np.float64) - outs[1]) / np.linalg.norm(outs[1])
    assert np.isfinite(outs[0]).all() and res <= lim, (res, lim)
    assert diff <= 50 * m * EPS[np.dtype(dt)], diff


# Shapes at the edges of the whole-solve column-group rule (16 columns when
# B has at most 32 columns, or at most 128 with a 16-column grid of more than
# 256 and at most 1024 workgroups; then 32 columns up to 1024 with a grid of
# at most 1024; else 64) and of the widened pipeline regions: (m, n, width
# the rule picks).
WIDTH_CASES = [
    (4096, 64, 32),    # 16-column grid of exactly 256: 32 columns
    (4096, 80, 16),    # 320 workgroups at 16 columns
    (8192, 128, 16),   # exactly 1024 at 16 columns
    (8192, 130, 32),   # past 128 columns: 32 (640 workgroups)
    (8192, 300, 64),   # 32-column grid of 1280: 64
    (4096, 512, 32),   # 32-column grid of exactly 1024
    (2100, 1500, 64),  # widened double pipeline (m n <= 2^22, n <= 2048)
    (3000, 555, 32),   # float below m n = 2^21 but wider than 128: pipeline
]


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("m,n,width", WIDTH_CASES)
def test_column_group_widths(lib, prod, dt, m, n, width):
    """One uplo / trans / diag combination per shape (rotating), against the
    recursive schedule down to 64-row leaves and the residual probe."""
    i = WIDTH_CASES.index((m, n, width))
    lower, trans, unit = bool(i & 1), bool(i & 2), bool(i & 4)
    rng = np.random.default_rng(m + 7 * n)
    a = L.conditioned_triangular(rng, m, dt, "F", lower, unit)
    b0 = L.rand(rng, (m, n), dt, "F")
    outs = []
    for max_n in (-1, 0):
        old = lib.mublis_gfx950_trsm_pipe_max_n(max_n)
        try:
            x = np.array(b0, copy=True, order="F")
            prod.call(dt, "trsm", L.COL, L.LEFT, L.LO if lower else L.UP,
                      L.TR if trans else L.NOTR, L.UNIT if unit else L.NONUNIT, m, n, 1.5,
                      a, m, x, m)
        finally:
            lib.mublis_gfx950_trsm_pipe_max_n(old)
        outs.append(x)
    res, lim = residual(a, outs[0], b0, 1.5, False, lower, trans, unit)
    diff = np.linalg.norm(outs[0].astype(np.float64) - outs[1]) / np.linalg.norm(outs[1])
    assert np.isfinite(outs[0]).all() and res <= lim, f"residual {res:.3e} > {lim:.3e}"
    assert diff <= 50 * m * EPS[np.dtype(dt)], f"vs 64-row leaves {diff:.3e}"


def test_timed_out_wait_leaves_nan(lib, prod):
    """Failure handling of the pipeline's bounded waits, forced from the host:
    with the spin limit at 2 polls the 64-step chain of a 4096 x 64 solve
    gives up (every wait still ends and the kernel exits normally), the
    call completes and B holds NaN instead of plausible values; the next
    solve with the default limit meets the residual bound again."""
    m, n = 4096, 64
    rng = np.random.default_rng(99)
    a = L.conditioned_triangular(rng, m, np.float64, "F", True, False)
    b0 = L.rand(rng, (m, n), np.float64, "F")

    def solve():
        x = np.array(b0, copy=True, order="F")
        prod.call(np.float64, "trsm", L.COL, L.LEFT, L.LO, L.NOTR, L.NONUNIT, m, n, 1.0, a,
                  m, x, m)
        return x

    old = lib.mublis_gfx950_trsm_pipe_spin_limit(2)
    try:
        x = solve()
    finally:
        assert lib.mublis_gfx950_trsm_pipe_spin_limit(old) == 2
    assert np.isnan(x).any(), "a given-up wait must show in the result"
    x = solve()
    res, lim = residual(a, x, b0, 1.0, False, True, False, False)
    assert np.isfinite(x).all() and res <= lim
