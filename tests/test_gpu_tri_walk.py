"""GPU checks of the triangle-aware launches at the sizes that take them:

* GEMMs with a triangular operand, whose launch takes the long-first,
  XCD-balanced tile walk on 128 x 128 tiles (at least 2048 of them) and on
  64 x 64 grids whose line count is not a multiple of 8 (padding
  workgroups): cblas_?trmm on device operands, all four side/uplo cases,
  against an FP64 torch product of the materialised triangle.
  Bar: ||B - B*||_F <= 2 s eps |alpha| ||T||_F ||B0||_F, s the triangle's
  order (the normwise GEMM bound of the other tests, with K = s).
* SYRK / SYR2K at n >= 4096, launched as 128-tiles strictly inside the ji
  triangle plus the diagonal blocks as ten 32-tiles each (n = 4100: a short
  last diagonal block), against FP64 torch on the stored triangle; the other
  triangle must come back bit for bit.
  Bar: ||C - C*||_F <= 2 k eps (|alpha| ||A|| ||B|| + |beta| ||C0||)
  (twice the product term for SYR2K).
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


SYRK_CASES = [
    # (dtype, routine, lower, n, k)
    ("f64", "syrk", 1, 8192, 256),
    ("f64", "syrk", 0, 4100, 300),
    ("f32", "syrk", 1, 4100, 129),
    ("f32", "syr2k", 0, 8192, 96),
    ("f64", "syr2k", 1, 4100, 64),
]


@pytest.mark.parametrize("dt,routine,lower,n,k", SYRK_CASES)
def test_syrk_split_triangle(env, dt, routine, lower, n, k):
    torch, lib = env
    tdt = torch.float64 if dt == "f64" else torch.float32
    eps = float(np.finfo(np.float64 if dt == "f64" else np.float32).eps)
    g = torch.Generator(device="cuda")
    g.manual_seed(n + k + lower)
    A = torch.rand((n, k), dtype=tdt, device="cuda", generator=g) * 2 - 1
    B = torch.rand((n, k), dtype=tdt, device="cuda", generator=g) * 2 - 1
    C0 = torch.rand((n, n), dtype=tdt, device="cuda", generator=g) * 2 - 1
    alpha, beta = 0.75, -0.5
    a_cm, b_cm = A.t().contiguous(), B.t().contiguous()  # column-major n x k
    c_cm = C0.t().contiguous()
    uplo = LO if lower else UP
    if routine == "syrk":
        f = lib.cblas_dsyrk if dt == "f64" else lib.cblas_ssyrk
        f(COL, uplo, NOTR, n, k, alpha, M.ptr(a_cm), n, beta, M.ptr(c_cm), n)
        P = A.double() @ A.double().t()
        pn = A.double().norm().item() ** 2
    else:
        f = lib.cblas_dsyr2k if dt == "f64" else lib.cblas_ssyr2k
        f(COL, uplo, NOTR, n, k, alpha, M.ptr(a_cm), n, M.ptr(b_cm), n, beta, M.ptr(c_cm), n)
        P = A.double() @ B.double().t() + B.double() @ A.double().t()
        pn = 2 * A.double().norm().item() * B.double().norm().item()
    torch.cuda.synchronize()
    got = c_cm.t()
    want = alpha * P + beta * C0.double()
    mask = torch.ones((n, n), dtype=torch.bool, device="cuda")
    mask = torch.tril(mask) if lower else torch.triu(mask)
    err = ((got.double() - want) * mask).norm().item()
    lim = 2 * k * eps * (abs(alpha) * pn + abs(beta) * C0.double().norm().item())
    print(f"{routine} {dt} lower={lower} n={n} k={k}: error {err:.3e} (bound {lim:.3e})")
    assert err <= lim
    assert torch.equal(got[~mask], C0[~mask]), "the other triangle was written"
