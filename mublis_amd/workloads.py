"""Synthetic workloads of the benchmark configurations, built on the device.

`grouped_problems` draws the batched configuration: GEMM problems with M,
N, K log-uniform in a range and SYRK problems with n and k log-uniform,
entries uniform[-1, 1], column-major device tensors, plus the
`SbatchOp` descriptors that point at them.  Tensors are stored transposed
(shape (cols, ld)) so that a row-major torch tensor is a column-major
matrix.
"""
import numpy as np

from . import BATCH_GEMM, BATCH_SYRK, SbatchOp


def _logu(rng, lo, hi, size):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size)).astype(np.int64).clip(lo, hi)


def grouped_problems(torch, seed=314159, n_gemm=4096, n_syrk=1024, gemm_range=(16, 512),
                     syrk_n=(32, 1024), syrk_k=(16, 512), alpha=1.0, beta=1.0,
                     vary_flags=False, device="cuda"):
    """Returns (ops, keep, specs): SbatchOp list, the tensors they point at
    (kept alive by the caller) and per-problem dicts for checking."""
    rng = np.random.default_rng(seed)
    g = torch.Generator(device=device).manual_seed(seed)
    ops, keep, specs = [], [], []

    def urand(rows, cols):
        # column-major rows x cols = tensor (cols, rows)
        return torch.rand((cols, rows), device=device, generator=g) * 2 - 1

    ms = _logu(rng, *gemm_range, n_gemm)
    ns = _logu(rng, *gemm_range, n_gemm)
    ks = _logu(rng, *gemm_range, n_gemm)
    for i in range(n_gemm):
        m, n, k = int(ms[i]), int(ns[i]), int(ks[i])
        ta = int(rng.integers(0, 2)) if vary_flags else 0
        tb = int(rng.integers(0, 2)) if vary_flags else 0
        A = urand(k, m) if ta else urand(m, k)
        B = urand(n, k) if tb else urand(k, n)
        C = urand(m, n)
        lda, ldb = A.shape[1], B.shape[1]
        ops.append(SbatchOp(BATCH_GEMM, ta, tb, 0, m, n, k, alpha, beta, A.data_ptr(), lda,
                            B.data_ptr(), ldb, C.data_ptr(), m))
        keep += [A, B, C]
        specs.append(dict(op="gemm", m=m, n=n, k=k, ta=ta, tb=tb, A=A, B=B, C=C, lda=lda,
                          ldb=ldb))
    sn = _logu(rng, *syrk_n, n_syrk)
    sk = _logu(rng, *syrk_k, n_syrk)
    for i in range(n_syrk):
        n, k = int(sn[i]), int(sk[i])
        tr = int(rng.integers(0, 2)) if vary_flags else 0
        up = int(rng.integers(0, 2)) if vary_flags else 0
        A = urand(k, n) if tr else urand(n, k)
        C = urand(n, n)
        ops.append(SbatchOp(BATCH_SYRK, tr, 0, 2 if up else 1, n, n, k, alpha, beta,
                            A.data_ptr(), A.shape[1], None, 1, C.data_ptr(), n))
        keep += [A, C]
        specs.append(dict(op="syrk", n=n, k=k, trans=tr, upper=up, A=A, C=C, lda=A.shape[1]))
    return ops, keep, specs
