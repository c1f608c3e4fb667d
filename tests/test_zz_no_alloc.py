"""GPU check of the allocation-free DTRSM entry point
(mublis_gfx950_dtrsm_noalloc_dev), through the C ABI.

It runs the schedule of mublis_gfx950_dtrsm_ws_dev given the minimal
workspace (mublis_gfx950_trsm_ws_bytes(m) bytes), with that workspace in a
static device array of the library.  Bar: bit-identical to the workspace
entry point on the same inputs -- two calls in a row, the second with a
smaller problem (the array keeps the first call's contents) -- plus the
TRSM residual bound ||op(A) X - alpha B|| <= 2 m eps ||A|| ||X|| against
float64; sizes beyond the array and operands without a unit stride are
refused with MUBLIS_EINVAL before anything is written.
"""
import ctypes

import numpy as np
import pytest

import mublis_amd as M

pytestmark = pytest.mark.gpu

_I, _D, _VP, _I64, _SZ = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_int64, \
    ctypes.c_size_t


@pytest.fixture(scope="module")
def lib():
    lib = M.load()
    assert lib.mublis_rt_init() == 0
    f = lib.mublis_gfx950_dtrsm_noalloc_dev
    f.argtypes = [_I, _I, _I, _I, _D, _VP, _I64, _I64, _VP, _I64, _I64, _VP]
    f.restype = _I
    return lib


def problem(torch, m, n, lower, seed):
    """A well-conditioned m x m triangle (off-diagonal N(0,1)/sqrt(m), diagonal
    1 + |N(0,1)|), NaN in the other triangle, and B m x n; column-major, as
    (cols, ld) tensors on the device."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    a = torch.randn((m, m), dtype=torch.float64, device="cuda", generator=g) / m ** 0.5
    d = 1.0 + torch.randn(m, dtype=torch.float64, device="cuda", generator=g).abs()
    # tensor (col, row): element (i, j) of A is a[j, i]
    keep = torch.ones((m, m), dtype=torch.bool, device="cuda")
    keep = torch.tril(keep) if not lower else torch.triu(keep)   # (j, i) with i >= j: lower
    a = torch.where(keep, a, torch.full_like(a, float("nan")))
    a.diagonal().copy_(d)
    b = torch.randn((n, m), dtype=torch.float64, device="cuda", generator=g)
    return a, b


def solve_ws(lib, torch, lower, unit, m, n, alpha, a, b):
    ws = torch.empty(int(lib.mublis_gfx950_trsm_ws_bytes(m)), dtype=torch.uint8, device="cuda")
    x = b.clone()
    st = lib.mublis_gfx950_dtrsm_ws_dev(lower, unit, m, n, alpha, a.data_ptr(), 1, m,
                                        x.data_ptr(), 1, m, ws.data_ptr(), ws.numel(), None)
    assert st == 0
    torch.cuda.synchronize()
    return x


def solve_noalloc(lib, torch, lower, unit, m, n, alpha, a, b):
    x = b.clone()
    torch.cuda.synchronize()
    st = lib.mublis_gfx950_dtrsm_noalloc_dev(lower, unit, m, n, alpha, a.data_ptr(), 1, m,
                                             x.data_ptr(), 1, m, None)
    assert st == 0, lib.mublis_rt_last_error()
    torch.cuda.synchronize()
    return x


def residual_ok(a, x, b, alpha, lower, unit):
    A = np.nan_to_num(a.cpu().numpy().T)
    A = np.tril(A) if lower else np.triu(A)
    if unit:
        np.fill_diagonal(A, 1.0)
    X = x.cpu().numpy().T
    B = b.cpu().numpy().T
    res = np.linalg.norm(A @ X - alpha * B)
    m = A.shape[0]
    return res <= 2 * m * np.finfo(np.float64).eps * np.linalg.norm(A) * np.linalg.norm(X)


@pytest.mark.parametrize("lower,unit", [(1, 0), (0, 1)])
def test_two_calls_match_the_workspace_entry(lib, lower, unit):
    """A large solve, then a smaller one on the same static workspace: both
    bit-identical to the workspace entry point and within the residual
    bound.  Shapes cover the pipelined (narrow) and recursive (wide)
    schedules."""
    torch = pytest.importorskip("torch")
    for (m, n, alpha, seed) in ((3000, 700, 0.75, 1), (1000, 96, -1.5, 2)):
        a, b = problem(torch, m, n, lower, seed)
        want = solve_ws(lib, torch, lower, unit, m, n, alpha, a, b)
        got = solve_noalloc(lib, torch, lower, unit, m, n, alpha, a, b)
        assert torch.equal(got, want), (m, n)
        assert residual_ok(a, got, b, alpha, lower, unit), (m, n)


def test_limits_are_refused(lib):
    """An order above the static workspace, and an operand without a unit
    stride, return MUBLIS_EINVAL and leave B untouched."""
    torch = pytest.importorskip("torch")
    b = torch.ones((4, 16385), dtype=torch.float64, device="cuda")
    a = torch.ones((8, 8), dtype=torch.float64, device="cuda")
    f = lib.mublis_gfx950_dtrsm_noalloc_dev
    assert f(1, 0, 16385, 4, 1.0, a.data_ptr(), 1, 16385, b.data_ptr(), 1, 16385,
             None) == M.EINVAL
    assert b"16384" in lib.mublis_rt_last_error()
    assert f(1, 0, 4, 2, 1.0, a.data_ptr(), 2, 16, b.data_ptr(), 1, 4, None) == M.EINVAL
    torch.cuda.synchronize()
    assert bool((b == 1).all())
