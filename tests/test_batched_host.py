"""Host-side validation of grouped batched problem lists (CPU only).

mublis_sgrouped_ws_bytes expands a list of mublis_sbatch_op_t into the work
list on the host and returns 0 when any problem is rejected, so the
argument rules of the grouped interface are checked here without a device:
the CBLAS leading-dimension minima per op and transpose
(frame/compat/cblas/cblas.c:150-171, 306-325, 380-404, column-major), the
transpose and triangle codes, negative dimensions and unknown op codes.
Nothing here dereferences the operand pointers, which stay NULL.
"""
import pytest

import mublis_amd as M


@pytest.fixture(scope="module")
def lib():
    return M.load()


def op(kind, m=8, n=6, k=5, ta=0, tb=0, uplo=1, lda=None, ldb=None, ldc=None):
    o = M.SbatchOp()
    o.op, o.trans_a, o.trans_b, o.uplo = kind, ta, tb, uplo
    o.m, o.n, o.k, o.alpha, o.beta = m, n, k, 1.0, 1.0
    if kind == M.BATCH_GEMM:
        o.lda = lda if lda is not None else (k if ta else m)
        o.ldb = ldb if ldb is not None else (n if tb else k)
        o.ldc = ldc if ldc is not None else m
    else:
        o.lda = lda if lda is not None else (k if ta else n)
        o.ldb = ldb if ldb is not None else (k if ta else n)
        o.ldc = ldc if ldc is not None else n
    return o


def ws(lib, ops):
    arr = (M.SbatchOp * len(ops))(*ops)
    return lib.mublis_sgrouped_ws_bytes(arr, len(ops))


GOOD = [op(M.BATCH_GEMM), op(M.BATCH_GEMM, ta=1, tb=1), op(M.BATCH_SYRK),
        op(M.BATCH_SYRK, ta=1, uplo=2), op(M.BATCH_SYR2K), op(M.BATCH_SYR2K, ta=1)]

BAD = {
    "gemm_lda_lt_m": op(M.BATCH_GEMM, lda=7),
    "gemm_T_lda_lt_k": op(M.BATCH_GEMM, ta=1, lda=4),
    "gemm_ldb_lt_k": op(M.BATCH_GEMM, ldb=4),
    "gemm_T_ldb_lt_n": op(M.BATCH_GEMM, tb=1, ldb=5),
    "gemm_ldc_lt_m": op(M.BATCH_GEMM, ldc=7),
    "gemm_ld_zero_empty": op(M.BATCH_GEMM, m=0, lda=0),
    "syrk_lda_lt_n": op(M.BATCH_SYRK, lda=5),
    "syrk_T_lda_lt_k": op(M.BATCH_SYRK, ta=1, lda=4),
    "syrk_ldc_lt_n": op(M.BATCH_SYRK, ldc=5),
    "syr2k_ldb_lt_n": op(M.BATCH_SYR2K, ldb=5),
    "syr2k_T_ldb_lt_k": op(M.BATCH_SYR2K, ta=1, ldb=4),
    "bad_trans_a": op(M.BATCH_GEMM, ta=2),
    "bad_trans_b": op(M.BATCH_GEMM, tb=3),
    "bad_uplo": op(M.BATCH_SYRK, uplo=0),
    "negative_k": op(M.BATCH_GEMM, k=-1),
    "unknown_op": op(7),
}


def test_valid_lists_are_accepted(lib):
    assert ws(lib, GOOD) > 0
    for o in GOOD:
        assert ws(lib, [o]) > 0


@pytest.mark.parametrize("name", sorted(BAD))
def test_invalid_problem_rejects_the_list(lib, name):
    """One bad problem anywhere in the list rejects the whole list."""
    assert ws(lib, [BAD[name]]) == 0
    assert ws(lib, GOOD[:3] + [BAD[name]] + GOOD[3:]) == 0


def test_syr2k_with_unit_leading_dimension_is_accepted(lib):
    """n = 1 lets lda or ldb be 1 while the other is larger: the two products
    of SYR2K still share one staging layout (the extent-1 dimension's stride
    is normalised), so such a list is valid."""
    for lda, ldb in ((1, 5), (5, 1)):
        assert ws(lib, [op(M.BATCH_SYR2K, n=1, k=7, lda=lda, ldb=ldb)]) > 0
    assert ws(lib, [op(M.BATCH_SYR2K, n=1, k=7, ta=1, lda=7, ldb=9)]) > 0
