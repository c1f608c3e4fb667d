"""GPU checks of GEMMs with a triangular operand at sizes where the launch
takes the long-first, XCD-balanced tile walk on 128 x 128 tiles (at least
2048 of them) and on 64 x 64 grids whose line count is not a multiple of 8
(padding workgroups): cblas_?trmm on device operands, all four
side/uplo cases (triangular A on the left, triangular B on the right),
against an FP64 torch product of the materialised triangle.

Bar: ||B - B*||_F <= 2 s eps |alpha| ||T||_F ||B0||_F, s the triangle's order
(the normwise GEMM bound of the other tests, with K = s).
"""
import numpy as np
import pytest

import mublis_amd as M

pytestmark = pytest.mark.gpu
COL, LEFT, RIGHT, LO, UP, NOTR, NONUNIT = 102, 141, 142, 122, 121, 111, 131


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = M.load()
    assert lib.mublis_rt_init() == 0, lib.mublis_rt_last_error()
    return torch, lib


CASES = [
    # (dtype, right, lower, m, n)
    ("f64", 0, 1, 8192, 4096),
    ("f64", 0, 0, 8192, 4096),
    ("f64", 1, 1, 4096, 8192),
    ("f64", 1, 0, 4096, 8192),
    ("f32", 0, 1, 8192, 4096),
    ("f32", 1, 0, 4096, 8192),
    ("f64", 0, 1, 3000, 700),
    ("f32", 1, 1, 700, 3000),
]


@pytest.mark.parametrize("dt,right,lower,m,n", CASES)
def test_trmm_long_first_walk(env, dt, right, lower, m, n):
    torch, lib = env
    tdt = torch.float64 if dt == "f64" else torch.float32
    eps = float(np.finfo(np.float64 if dt == "f64" else np.float32).eps)
    g = torch.Generator(device="cuda")
    g.manual_seed(1000 * m + 10 * n + 2 * right + lower)
    s = n if right else m
    A = torch.rand((s, s), dtype=tdt, device="cuda", generator=g) * 2 - 1
    B0 = torch.rand((m, n), dtype=tdt, device="cuda", generator=g) * 2 - 1
    alpha = -1.25
    T = torch.tril(A) if lower else torch.triu(A)
    want = alpha * ((B0.double() @ T.double()) if right else (T.double() @ B0.double()))
    a_cm = A.t().contiguous()  # column-major storage of A
    b_cm = B0.t().contiguous()
    f = lib.cblas_dtrmm if dt == "f64" else lib.cblas_strmm
    f(COL, RIGHT if right else LEFT, LO if lower else UP, NOTR, NONUNIT, m, n, alpha,
      M.ptr(a_cm), s, M.ptr(b_cm), m)
    torch.cuda.synchronize()
    got = b_cm.t().double()
    err = (got - want).norm().item()
    lim = 2 * s * eps * abs(alpha) * T.double().norm().item() * B0.double().norm().item()
    print(f"trmm {dt} r{right} l{lower} {m}x{n}: normwise error {err:.3e} (bound {lim:.3e})")
    assert err <= lim
