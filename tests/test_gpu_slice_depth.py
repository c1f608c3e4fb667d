"""GPU checks of the 64-tile GEMM launches that stage K in half-depth slices
(grids of at least 1024 FP64 / more than 1024 FP32 workgroups without a
triangular operand, 6 workgroups per CU) against the deeper-slice launches
of smaller grids:

* the full product must equal, bit for bit, the same product computed as
  two calls on column halves (half the grid: deep slices), since the slice
  depth changes only where K is cut, never the order of the MFMA k steps;
* SYRK on such a grid must equal the GEMM A * A^T on its stored triangle,
  bit for bit, and leave the other triangle untouched;
* both against an FP64 torch product.
  Bar: ||C - C*||_F <= 2 k eps (|alpha| ||A||_F ||B||_F + |beta| ||C0||_F).
"""
import numpy as np
import pytest

import mublis_amd as M

pytestmark = pytest.mark.gpu
COL, LO, UP, NOTR, TR = 102, 122, 121, 111, 112


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = M.load()
    assert lib.mublis_rt_init() == 0, lib.mublis_rt_last_error()
    return torch, lib


def colmajor(torch, rows, cols, dt, g):
    # column-major rows x cols with ld = rows: a (cols, rows) row-major tensor
    return torch.rand((cols, rows), dtype=dt, device="cuda", generator=g) * 2 - 1


def bound(k, eps, alpha, beta, A, B, C0):
    return 2 * k * eps * (abs(alpha) * A.double().norm().item() * B.double().norm().item()
                          + abs(beta) * C0.double().norm().item())


GEMM_CASES = [
    # (dtype, ta, m, n, k): 64-tile grids of 1024 (f64) / 1536 (f32) tiles
    ("f64", 0, 2048, 2048, 2048),
    ("f64", 1, 1536, 4096, 1000),
    ("f32", 0, 1536, 4096, 1536),
    ("f32", 1, 2048, 3072, 1000),
]


@pytest.mark.parametrize("dt,ta,m,n,k", GEMM_CASES)
def test_shallow_gemm_matches_deep_halves(env, dt, ta, m, n, k):
    torch, lib = env
    tdt = torch.float64 if dt == "f64" else torch.float32
    eps = float(np.finfo(np.float64 if dt == "f64" else np.float32).eps)
    f = lib.cblas_dgemm if dt == "f64" else lib.cblas_sgemm
    g = torch.Generator(device="cuda")
    g.manual_seed(7 * m + n + 3 * k + ta)
    A = colmajor(torch, k, m, tdt, g) if ta else colmajor(torch, m, k, tdt, g)
    lda = k if ta else m
    B = colmajor(torch, k, n, tdt, g)
    C0 = colmajor(torch, m, n, tdt, g)
    alpha, beta = -0.75, 0.5
    C = C0.clone()
    f(COL, TR if ta else NOTR, NOTR, m, n, k, alpha, M.ptr(A), lda, M.ptr(B), k, beta,
      M.ptr(C), m)
    H = C0.clone()
    h = n // 2
    for j0, w in ((0, h), (h, n - h)):
        # columns j0 .. j0 + w - 1 of B and C (column-major: rows of the tensors)
        f(COL, TR if ta else NOTR, NOTR, m, w, k, alpha, M.ptr(A), lda, M.ptr(B[j0:]), k,
          beta, M.ptr(H[j0:]), m)
    torch.cuda.synchronize()
    assert torch.equal(C.view(torch.int64 if dt == "f64" else torch.int32),
                       H.view(torch.int64 if dt == "f64" else torch.int32))
    Am = A.double().t() if not ta else A.double()   # m x k
    Bm = B.double().t()                              # k x n
    Cm = C0.double().t()                             # m x n
    want = alpha * (Am @ Bm) + beta * Cm
    err = (C.double().t() - want).norm().item()
    assert err <= bound(k, eps, alpha, beta, A, B, C0), err


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("upper", [0, 1])
def test_shallow_syrk_matches_gemm(env, dt, upper):
    torch, lib = env
    tdt = torch.float64 if dt == "f64" else torch.float32
    eps = float(np.finfo(np.float64 if dt == "f64" else np.float32).eps)
    n, k = 3072, 1000  # 48 x 49 / 2 = 1176 64-tiles touching the triangle
    g = torch.Generator(device="cuda")
    g.manual_seed(31 * n + k + upper)
    A = colmajor(torch, n, k, tdt, g)
    C0 = colmajor(torch, n, n, tdt, g)
    alpha, beta = 1.5, -0.25
    C = C0.clone()
    fs = lib.cblas_dsyrk if dt == "f64" else lib.cblas_ssyrk
    fg = lib.cblas_dgemm if dt == "f64" else lib.cblas_sgemm
    fs(COL, UP if upper else LO, NOTR, n, k, alpha, M.ptr(A), n, beta, M.ptr(C), n)
    G = C0.clone()
    fg(COL, NOTR, TR, n, n, k, alpha, M.ptr(A), n, M.ptr(A), n, beta, M.ptr(G), n)
    torch.cuda.synchronize()
    # column-major C(i, j) is tensor[j, i]: the lower triangle (j <= i) is the
    # tensor's upper triangle
    keep = torch.triu(torch.ones((n, n), dtype=torch.bool, device="cuda"))
    stored = keep.t() if upper else keep
    it = torch.int64 if dt == "f64" else torch.int32
    assert torch.equal(C.view(it)[stored], G.view(it)[stored])
    assert torch.equal(C.view(it)[~stored], C0.view(it)[~stored])
    Am = A.double().t()
    want = alpha * (Am @ Am.t()) + beta * C0.double().t()
    diff = (C.double().t() - want)[stored.t()]
    assert diff.norm().item() <= bound(k, eps, alpha, beta, A, A, C0)
