"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against
it.  CPU only.

  * every golden vector written from the reference's own C build (both the
    linux_avx2 and the reference config) -- tests/golden/*.npz;
  * scipy's BLAS, the reference harness's own expected values
    (test/correctness/check_cblas.py:179-190), on the 192 harness cases;
  * exact closed forms: small-integer operands (every product and sum is an
    exactly representable integer, so a dropped term, a wrong sign, index or
    transposed operand changes the bits), identity/ones inputs of the
    reference benchmark (test/performance/benchmark_cblas.py:124-141), and
    TRSM on integer systems with power-of-two diagonals;
  * the argument-check parameter numbers against the reference build's
    cblas_xerbla messages (frame/compat/cblas/cblas.c:114-516);
  * the special cases of the driver contract: alpha == 0, beta == 0 with NaN
    in C, k == 0, m == 0 / n == 0.
"""
import ctypes
import itertools
import json
import os
import subprocess
import sys

import numpy as np
import pytest
from scipy.linalg import blas as sblas

import l3cases as L



@pytest.fixture(scope="module")
def orc():
    if not os.path.exists(L.ORACLE_LIB):
        subprocess.check_call(["make", "-C", os.path.join(L.ROOT, "oracle"), "oracle"])
    return L.Cblas(L.ORACLE_LIB, prefix="orc_", returns_status=True)


def rel_err(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)


@pytest.mark.parametrize("name", ["l3_harness", "l3_ragged"])
def test_oracle_matches_reference_build(orc, name):
    """Oracle vs the reference's compiled outputs, both configs."""
    for c, arr in L.load_golden(name):
        out = np.array(arr["c0"], copy=True, order=c["order"])
        st = orc.call(c["dtype"], c["routine"], *L.args_from_golden(c, arr, out))
        assert st == 0, L.case_id(c)
        mask = L.golden_mask(c, out.shape[0])
        for key in ("avx2", "ref"):
            try:
                L.assert_ref_close(out, arr[key], c["dtype"], mask, arr["c0"])
            except AssertionError as e:
                raise AssertionError(f"{name} {L.case_id(c)} vs {key}: {e}")


@pytest.mark.parametrize("case", L.all_cases(), ids=L.case_id)
def test_oracle_matches_scipy_and_closed_form(orc, case):
    """The reference harness's own check (scipy BLAS), plus float64 numpy."""
    c = case
    prob = L.build_problem(c, np.random.default_rng(L.SEED))
    out = np.array(prob["c0"], copy=True, order=c["order"])
    assert orc.call(c["dtype"], c["routine"], *prob["args"](out)) == 0
    L.assert_ref_close(out, prob["exact"], c["dtype"], prob["mask"], prob["c0"])
    # scipy BLAS, as check_cblas.py:179-190 computes the expected value
    f = sblas.get_blas_funcs(c["routine"], dtype=c["dtype"])
    F = lambda x: np.array(x, dtype=c["dtype"], order="F", copy=True)  # noqa: E731
    a, b = prob["inputs"]["a"], prob["inputs"].get("b")
    r = c["routine"]
    if r == "gemm":
        exp = f(1.25, F(a), F(b), beta=-0.5, c=F(prob["c0"]), trans_a=int(c["ta"]),
                trans_b=int(c["tb"]))
    elif r == "symm":
        exp = f(0.75, F(a), F(b), beta=-0.25, c=F(prob["c0"]), side=int(c["right"]),
                lower=int(c["lower"]))
    elif r == "syrk":
        exp = f(1.25, F(a), beta=-0.5, c=F(prob["c0"]), trans=int(c["trans"]),
                lower=int(c["lower"]))
    elif r == "syr2k":
        exp = f(0.75, F(a), F(b), beta=-0.25, c=F(prob["c0"]), trans=int(c["trans"]),
                lower=int(c["lower"]))
    else:
        exp = f(0.75, F(a), F(prob["c0"]), side=int(c["right"]), lower=int(c["lower"]),
                trans_a=int(c["trans"]), diag=int(c["unit"]))
    L.assert_ref_close(out, exp, c["dtype"], prob["mask"], prob["c0"])


# ------------------------------------------------------ exact closed forms --
def ints(rng, shape, dt, order="F", lo=-4, hi=5):
    return np.array(rng.integers(lo, hi, size=shape), dtype=dt, order=order)


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("order,co", L.LAYOUTS)
def test_exact_integer_products(orc, dt, order, co):
    rng = np.random.default_rng(7)
    m, n, k = 9, 7, 6
    for ta, tb in itertools.product((0, 1), repeat=2):
        a = ints(rng, (k, m) if ta else (m, k), dt, order)
        b = ints(rng, (n, k) if tb else (k, n), dt, order)
        c0 = ints(rng, (m, n), dt, order)
        out = c0.copy(order=order)
        assert orc.call(dt, "gemm", co, L.TR if ta else L.NOTR, L.TR if tb else L.NOTR, m, n,
                        k, 2.0, a, L.ld_of(a, order), b, L.ld_of(b, order), -1.0, out,
                        L.ld_of(out, order)) == 0
        A = a.astype(np.int64).T if ta else a.astype(np.int64)
        B = b.astype(np.int64).T if tb else b.astype(np.int64)
        np.testing.assert_array_equal(out, (2 * A @ B - c0.astype(np.int64)).astype(dt))
    # SYMM: only the stored triangle is read; the rest is NaN
    for right, lower in itertools.product((0, 1), repeat=2):
        s = n if right else m
        full = ints(rng, (s, s), np.int64)
        full = np.tril(full) + np.tril(full, -1).T
        store = full.astype(dt)
        store[np.triu_indices(s, 1) if lower else np.tril_indices(s, -1)] = np.nan
        store = np.array(store, order=order)
        b = ints(rng, (m, n), dt, order)
        c0 = ints(rng, (m, n), dt, order)
        out = c0.copy(order=order)
        assert orc.call(dt, "symm", co, L.RIGHT if right else L.LEFT, L.LO if lower else L.UP,
                        m, n, 3.0, store, L.ld_of(store, order), b, L.ld_of(b, order), 0.0,
                        out, L.ld_of(out, order)) == 0
        B = b.astype(np.int64)
        np.testing.assert_array_equal(out, (3 * (B @ full if right else full @ B)).astype(dt))
    # SYRK / SYR2K: untouched triangle keeps its bits
    for trans, lower in itertools.product((0, 1), repeat=2):
        a = ints(rng, (k, n) if trans else (n, k), dt, order)
        b = ints(rng, (k, n) if trans else (n, k), dt, order)
        c0 = ints(rng, (n, n), dt, order)
        A = a.astype(np.int64).T if trans else a.astype(np.int64)
        B = b.astype(np.int64).T if trans else b.astype(np.int64)
        mask = np.tril(np.ones((n, n), bool)) if lower else np.triu(np.ones((n, n), bool))
        out = c0.copy(order=order)
        assert orc.call(dt, "syrk", co, L.LO if lower else L.UP, L.TR if trans else L.NOTR, n,
                        k, 1.0, a, L.ld_of(a, order), 1.0, out, L.ld_of(out, order)) == 0
        exp = A @ A.T + c0.astype(np.int64)
        np.testing.assert_array_equal(out[mask], exp[mask].astype(dt))
        np.testing.assert_array_equal(out[~mask], c0[~mask])
        out = c0.copy(order=order)
        assert orc.call(dt, "syr2k", co, L.LO if lower else L.UP, L.TR if trans else L.NOTR, n,
                        k, 1.0, a, L.ld_of(a, order), b, L.ld_of(b, order), 2.0, out,
                        L.ld_of(out, order)) == 0
        exp = A @ B.T + B @ A.T + 2 * c0.astype(np.int64)
        np.testing.assert_array_equal(out[mask], exp[mask].astype(dt))
        np.testing.assert_array_equal(out[~mask], c0[~mask])


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("order,co", L.LAYOUTS)
def test_exact_triangular(orc, dt, order, co):
    """TRMM on integers and TRSM on integer systems whose diagonal is a power
    of two, so substitution with reciprocal diagonals is exact."""
    rng = np.random.default_rng(11)
    m, n = 8, 5
    for right, lower, trans, unit in itertools.product((0, 1), repeat=4):
        s = n if right else m
        t = ints(rng, (s, s), np.int64, lo=-2, hi=3)
        t = np.tril(t, -1) if lower else np.triu(t, 1)
        d = np.ones(s) if unit else 2.0 ** rng.integers(-1, 2, size=s)
        dense = t + np.diag(d)
        store = dense.astype(dt)
        store[np.triu_indices(s, 1) if lower else np.tril_indices(s, -1)] = np.nan
        if unit:
            np.fill_diagonal(store, np.nan)
        store = np.array(store, order=order)
        op = dense.T if trans else dense
        x = ints(rng, (m, n), np.float64, lo=-3, hi=4)
        flags = (co, L.RIGHT if right else L.LEFT, L.LO if lower else L.UP,
                 L.TR if trans else L.NOTR, L.UNIT if unit else L.NONUNIT, m, n)
        # TRMM: B := 2 op(A) X (or 2 X op(A)), exact in both precisions
        b = np.array(x, dtype=dt, order=order)
        assert orc.call(dt, "trmm", *flags, 2.0, store, L.ld_of(store, order), b,
                        L.ld_of(b, order)) == 0
        np.testing.assert_array_equal(b, (2 * (x @ op if right else op @ x)).astype(dt))
        # TRSM: B = op(A) X exactly; solving with alpha 1 recovers X bit for bit
        rhs = x @ op if right else op @ x
        if np.abs(rhs).max() > 2 ** 20:
            continue
        b = np.array(rhs, dtype=dt, order=order)
        assert orc.call(dt, "trsm", *flags, 1.0, store, L.ld_of(store, order), b,
                        L.ld_of(b, order)) == 0
        np.testing.assert_array_equal(b, x.astype(dt))


@pytest.mark.parametrize("dt", L.DTYPES)
def test_benchmark_known_answers(orc, dt):
    """benchmark_cblas.py:124-141: ones give k in every entry; identity
    leaves B unchanged under TRMM and TRSM."""
    m = n = k = 33
    a = np.ones((m, k), dt, order="F")
    b = np.ones((k, n), dt, order="F")
    c = np.zeros((m, n), dt, order="F")
    orc.call(dt, "gemm", L.COL, L.NOTR, L.NOTR, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    assert np.all(c == k)
    eye = np.eye(m, dtype=dt, order="F")
    x = np.array(np.random.default_rng(3).standard_normal((m, n)), dt, order="F")
    for r in ("trmm", "trsm"):
        y = x.copy(order="F")
        orc.call(dt, r, L.COL, L.LEFT, L.LO, L.NOTR, L.NONUNIT, m, n, 1.0, eye, m, y, m)
        np.testing.assert_array_equal(y, x)


@pytest.mark.parametrize("dt", L.DTYPES)
def test_driver_special_cases(orc, dt):
    """alpha == 0 reads neither A nor B; beta == 0 never reads C
    (include/mublis_instantiate.h:51, frame/l3/l3_driver.c:104-124);
    k == 0 gives beta*C; m == 0 or n == 0 touches nothing."""
    m, n, k = 6, 5, 4
    nan_a = np.full((m, k), np.nan, dt, order="F")
    nan_b = np.full((k, n), np.nan, dt, order="F")
    c0 = np.array(np.arange(m * n).reshape(m, n), dt, order="F")
    c = c0.copy(order="F")
    orc.call(dt, "gemm", L.COL, L.NOTR, L.NOTR, m, n, k, 0.0, nan_a, m, nan_b, k, 2.0, c, m)
    np.testing.assert_array_equal(c, 2 * c0)
    a = np.ones((m, k), dt, order="F")
    b = np.ones((k, n), dt, order="F")
    c = np.full((m, n), np.nan, dt, order="F")
    orc.call(dt, "gemm", L.COL, L.NOTR, L.NOTR, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    assert np.all(c == k)
    c = c0.copy(order="F")
    orc.call(dt, "gemm", L.COL, L.NOTR, L.NOTR, m, n, 0, 1.0, a, m, b, 1, 0.5, c, m)
    np.testing.assert_array_equal(c, 0.5 * c0)
    c = c0.copy(order="F")
    orc.call(dt, "gemm", L.COL, L.NOTR, L.NOTR, 0, n, k, 1.0, a, 1, b, k, 0.5, c, 1)
    orc.call(dt, "gemm", L.COL, L.NOTR, L.NOTR, m, 0, k, 1.0, a, m, b, k, 0.5, c, m)
    np.testing.assert_array_equal(c, c0)
    for r in ("trmm", "trsm"):
        y = c0.copy(order="F")
        tri = np.full((m, m), np.nan, dt, order="F")
        orc.call(dt, r, L.COL, L.LEFT, L.LO, L.NOTR, L.NONUNIT, m, n, 0.0, tri, m, y, m)
        assert np.all(y == 0)


# -------------------------------------------------- argument-check numbers --
BAD_CALLS = [
    ("gemm", (0, 111, 111, 2, 2, 2, 1.0, "A", 2, "B", 2, 0.0, "C", 2)),
    ("gemm", (102, 0, 111, 2, 2, 2, 1.0, "A", 2, "B", 2, 0.0, "C", 2)),
    ("gemm", (102, 111, 0, 2, 2, 2, 1.0, "A", 2, "B", 2, 0.0, "C", 2)),
    ("gemm", (102, 111, 111, -1, 2, 2, 1.0, "A", 2, "B", 2, 0.0, "C", 2)),
    ("gemm", (102, 111, 111, 2, -1, 2, 1.0, "A", 2, "B", 2, 0.0, "C", 2)),
    ("gemm", (102, 111, 111, 2, 2, -1, 1.0, "A", 2, "B", 2, 0.0, "C", 2)),
    ("gemm", (102, 111, 111, 3, 2, 2, 1.0, "A", 2, "B", 2, 0.0, "C", 3)),
    ("gemm", (101, 111, 112, 2, 3, 2, 1.0, "A", 2, "B", 1, 0.0, "C", 3)),
    ("gemm", (102, 111, 111, 3, 2, 2, 1.0, "A", 3, "B", 2, 0.0, "C", 2)),
    ("symm", (102, 0, 122, 2, 2, 1.0, "A", 2, "B", 2, 0.0, "C", 2)),
    ("symm", (102, 141, 0, 2, 2, 1.0, "A", 2, "B", 2, 0.0, "C", 2)),
    ("symm", (102, 142, 122, 2, 3, 1.0, "A", 2, "B", 2, 0.0, "C", 2)),
    ("symm", (101, 141, 122, 2, 3, 1.0, "A", 2, "B", 2, 0.0, "C", 3)),
    ("symm", (101, 141, 122, 2, 3, 1.0, "A", 2, "B", 3, 0.0, "C", 2)),
    ("syrk", (102, 0, 111, 2, 2, 1.0, "A", 2, 0.0, "C", 2)),
    ("syrk", (102, 122, 0, 2, 2, 1.0, "A", 2, 0.0, "C", 2)),
    ("syrk", (102, 122, 111, -2, 2, 1.0, "A", 2, 0.0, "C", 2)),
    ("syrk", (102, 122, 111, 2, -2, 1.0, "A", 2, 0.0, "C", 2)),
    ("syrk", (101, 122, 111, 2, 3, 1.0, "A", 2, 0.0, "C", 2)),
    ("syrk", (102, 122, 111, 3, 2, 1.0, "A", 3, 0.0, "C", 2)),
    ("syr2k", (102, 122, 112, 3, 2, 1.0, "A", 2, "B", 1, 0.0, "C", 3)),
    ("syr2k", (102, 122, 111, 3, 2, 1.0, "A", 3, "B", 3, 0.0, "C", 2)),
    ("trsm", (102, 0, 122, 111, 131, 2, 2, 1.0, "A", 2, "B", 2)),
    ("trsm", (102, 141, 0, 111, 131, 2, 2, 1.0, "A", 2, "B", 2)),
    ("trsm", (102, 141, 122, 0, 131, 2, 2, 1.0, "A", 2, "B", 2)),
    ("trsm", (102, 141, 122, 111, 0, 2, 2, 1.0, "A", 2, "B", 2)),
    ("trsm", (102, 141, 122, 111, 131, -1, 2, 1.0, "A", 2, "B", 2)),
    ("trsm", (102, 141, 122, 111, 131, 2, -1, 1.0, "A", 2, "B", 2)),
    ("trmm", (102, 142, 122, 111, 131, 2, 3, 1.0, "A", 2, "B", 2)),
    ("trmm", (101, 141, 122, 111, 131, 2, 3, 1.0, "A", 2, "B", 2)),
]

_XERBLA_PROBE = r"""
import ctypes, sys, numpy as np
sys.path.insert(0, %r)
import l3cases as L
lib = L.Cblas(%r)
routine = %r
args = %r
buf = {k: np.zeros(64, np.float64) for k in "ABC"}
conv = [buf[a] if isinstance(a, str) else a for a in args]
lib.call(np.float64, routine, *conv)
print("RETURNED")
"""


@pytest.mark.parametrize("routine,args", BAD_CALLS, ids=[f"{r}{i}" for i, (r, _) in
                                                         enumerate(BAD_CALLS)])
def test_argument_numbers_match_reference_xerbla(orc, routine, args):
    """The oracle reports the parameter number that the reference build's
    cblas_xerbla prints ("Parameter %d to routine %s was incorrect") before
    it exits (frame/compat/cblas/cblas.c:14-32)."""
    if not os.path.exists(L.REF_AVX2_LIB):
        pytest.skip("reference build absent")
    code = _XERBLA_PROBE % (os.path.dirname(os.path.abspath(__file__)), L.REF_AVX2_LIB,
                            routine, args)
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=60)
    assert p.returncode != 0 and "RETURNED" not in p.stdout
    line = [ln for ln in p.stderr.splitlines() if ln.startswith("Parameter")][0]
    ref_num = int(line.split()[1])
    buf = {k: np.zeros(64, np.float64) for k in "ABC"}
    conv = [buf[a] if isinstance(a, str) else a for a in args]
    assert orc.call(np.float64, routine, *conv) == ref_num


_XERBLA_TYPED = r"""
import ctypes, sys, numpy as np
sys.path.insert(0, %r)
import l3cases as L
lib = L.Cblas(%r)
routine, args, dt = %r, %r, np.dtype(%r)
buf = {k: np.zeros(64, dt) for k in "ABC"}
conv = [buf[a] if isinstance(a, str) else a for a in args]
lib.call(dt, routine, *conv)
print("RETURNED")
"""


def _xerbla_run(lib_path, routine, args, dt):
    code = _XERBLA_TYPED % (os.path.dirname(os.path.abspath(__file__)), lib_path, routine,
                            args, np.dtype(dt).name)
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=120, env=dict(os.environ, HIP_VISIBLE_DEVICES=""))
    return p.returncode, p.stdout, p.stderr


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("routine,args", BAD_CALLS, ids=[f"{r}{i}" for i, (r, _) in
                                                         enumerate(BAD_CALLS)])
def test_product_argument_checks_match_reference(routine, args, dt):
    """The product's cblas_* (libmublis.so) rejects each invalid call exactly
    as the reference build does: the same stderr text from the default
    cblas_xerbla ("Parameter %d to routine %s was incorrect", the routine's
    own name), the same exit status, and nothing returned
    (frame/compat/cblas/cblas.c:14-32,92-112).  The checks run before any
    device work, so no GPU is needed (HIP_VISIBLE_DEVICES is empty)."""
    if not os.path.exists(L.REF_AVX2_LIB):
        pytest.skip("reference build absent")
    prod_lib = os.path.join(L.ROOT, "mublis_amd", "lib", "libmublis.so")
    r_rc, r_out, r_err = _xerbla_run(L.REF_AVX2_LIB, routine, args, dt)
    p_rc, p_out, p_err = _xerbla_run(prod_lib, routine, args, dt)
    assert r_rc == 1 and "RETURNED" not in r_out
    assert p_rc == r_rc and "RETURNED" not in p_out
    ref_lines = [ln for ln in r_err.splitlines() if ln.startswith("Parameter")]
    prod_lines = [ln for ln in p_err.splitlines() if ln.startswith("Parameter")]
    assert len(ref_lines) == 1 and prod_lines == ref_lines, (p_err, r_err)
    p_name = "s" if np.dtype(dt) == np.float32 else "d"
    assert f"routine cblas_{p_name}{routine} " in ref_lines[0]


# ------------------------------------------------ large-problem probes --
# orc_{d,s}gemm_entries, orc_{d,s}gemv_ld and orc_{d,s}trmv_ld compute the
# sampled entries and residual probes of the full-size GPU checks.  Each is
# pinned against float64 numpy on operands where every plausible mistake
# shows: outputs are pre-filled with NaN (an untouched output fails), the
# storage the helper must not read (padding rows beyond m, the unstored
# triangle, a unit diagonal) is NaN, the matrices are far from symmetric (a
# transposed index fails) and rectangular where the helper allows it.
_CI, _CD, _CF, _VP = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_void_p


def _probe(lib, name, argtypes):
    f = getattr(lib, name)
    f.argtypes = argtypes
    f.restype = None
    return f


def _padded(rng, rows, cols, ld, dt):
    """rows x cols column-major in ld >= rows storage; padding rows are NaN."""
    v = np.full((ld, cols), np.nan)
    v[:rows] = rng.standard_normal((rows, cols)) + np.arange(cols)[None, :] * 0.01
    return np.asfortranarray(v.astype(dt))


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("ta,tb", list(itertools.product((L.NOTR, L.TR), repeat=2)))
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_probe_gemm_entries(orc, dt, ta, tb, beta):
    """Sampled entries of alpha*op(A)op(B) + beta*C0 for every transpose pair;
    beta == 0 must not read C0 (it is NaN then)."""
    p = "d" if dt == np.float64 else "s"
    ct = _CD if p == "d" else _CF
    f = _probe(orc.lib, f"orc_{p}gemm_entries",
               [_CI, _CI, _CI, ct, _VP, _CI, _VP, _CI, ct, _VP, _CI, _CI, _VP, _VP, _VP])
    rng = np.random.default_rng(11)
    m, n, k = 37, 29, 23
    ra, ca = (m, k) if ta == L.NOTR else (k, m)
    rb, cb = (k, n) if tb == L.NOTR else (n, k)
    A = _padded(rng, ra, ca, ra + 3, dt)
    B = _padded(rng, rb, cb, rb + 2, dt)
    C0 = _padded(rng, m, n, m + 1, dt)
    if beta == 0:
        C0[:] = np.nan
    I = np.array(rng.integers(0, m, 80), np.int32)
    J = np.array(rng.integers(0, n, 80), np.int32)
    out = np.full(80, np.nan)
    f(ta, tb, k, 1.5, A.ctypes.data, ra + 3, B.ctypes.data, rb + 2, beta, C0.ctypes.data,
      m + 1, 80, I.ctypes.data, J.ctypes.data, out.ctypes.data)
    opa = A[:ra].astype(np.float64)
    opa = opa if ta == L.NOTR else opa.T
    opb = B[:rb].astype(np.float64)
    opb = opb if tb == L.NOTR else opb.T
    full = 1.5 * (opa @ opb)
    if beta:
        full += beta * C0[:m].astype(np.float64)
    np.testing.assert_allclose(out, full[I, J], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("dt", L.DTYPES)
def test_probe_gemv(orc, dt):
    """y := M x for a rectangular column-major M in padded storage."""
    p = "d" if dt == np.float64 else "s"
    f = _probe(orc.lib, f"orc_{p}gemv_ld", [_CI, _CI, _VP, _CI, _VP, _VP])
    rng = np.random.default_rng(12)
    for m, n in ((41, 17), (17, 41), (1, 9), (9, 1)):
        Mx = _padded(rng, m, n, m + 5, dt)
        x = rng.standard_normal(n)
        y = np.full(m, np.nan)
        f(m, n, Mx.ctypes.data, m + 5, x.ctypes.data, y.ctypes.data)
        want = Mx[:m].astype(np.float64) @ x
        np.testing.assert_allclose(y, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("uplo,ta,diag", list(itertools.product(
    (L.LO, L.UP), (L.NOTR, L.TR), (L.NONUNIT, L.UNIT))))
def test_probe_trmv(orc, dt, uplo, ta, diag):
    """y := op(tri(A)) x in all eight (uplo, trans, diag) branches; the unstored
    triangle, a unit diagonal and the padding rows are NaN."""
    p = "d" if dt == np.float64 else "s"
    f = _probe(orc.lib, f"orc_{p}trmv_ld", [_CI, _CI, _CI, _CI, _VP, _CI, _VP, _VP])
    rng = np.random.default_rng(13)
    m, ld = 45, 49
    A = _padded(rng, m, m, ld, np.float64)
    A[:m] += 3 * np.tril(np.ones((m, m)), -1)   # far from symmetric
    lower = uplo == L.LO
    stor = A.copy()
    mask = np.triu(np.ones((m, m), bool), 1) if lower else np.tril(np.ones((m, m), bool), -1)
    stor[:m][mask] = np.nan
    if diag == L.UNIT:
        d = np.arange(m)
        stor[d, d] = np.nan
    stor = np.asfortranarray(stor.astype(dt))
    T = np.where(np.isnan(stor[:m].astype(np.float64)), 0.0, stor[:m].astype(np.float64))
    if diag == L.UNIT:
        np.fill_diagonal(T, 1.0)
    x = rng.standard_normal(m)
    y = np.full(m, np.nan)
    f(uplo, ta, diag, m, stor.ctypes.data, ld, x.ctypes.data, y.ctypes.data)
    want = (T if ta == L.NOTR else T.T) @ x
    np.testing.assert_allclose(y, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # the other transpose gives a clearly different vector: the check above
    # would fail a helper that ignored `ta`
    other = (T.T if ta == L.NOTR else T) @ x
    assert np.linalg.norm(other - want) > 1e-3 * np.linalg.norm(want)


def test_every_oracle_export_is_pinned():
    """Every function liboracle.so exports is exercised by a CPU test of this
    file, either by name or through the CBLAS-shaped calls (prefix orc_)."""
    out = subprocess.run(["nm", "-D", "--defined-only", L.ORACLE_LIB], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " orc_" in ln}
    by_cblas = {f"orc_{p}{r}" for p in "sd"
                for r in ("gemm", "symm", "syrk", "syr2k", "trmm", "trsm")}
    src = open(__file__).read()
    by_name = {s for s in exported if s in src or s.replace("orc_d", "orc_{p}") in src
               or s.replace("orc_s", "orc_{p}") in src}
    # the SYRK / SYR2K entry points are thin calls of these two
    impl = {f"orc_{p}{r}_impl" for p in "sd" for r in ("syrk", "syr2k")}
    missing = exported - by_cblas - by_name - impl
    assert not missing, missing


# ------------------------------------------ generic driver, custom descriptors --
def _orc_driver(lib, dt):
    ct = ctypes.c_float if np.dtype(dt) == np.float32 else ctypes.c_double
    f = getattr(lib, "orc_sdriver" if ct is ctypes.c_float else "orc_ddriver")
    f.argtypes = [_CI, _CI, _CI, ct, _VP, _CI, _CI, _VP, _CI, _CI, ct, _VP, _CI, _CI] + \
        [_CI] * 9
    f.restype = _CI
    return f


def _check_driver_out(prob, got, dt, who):
    """Written elements within the normwise bound of the float64 value;
    everything else in C's storage (the ji-excluded part, column gaps)
    bit-identical to the input."""
    idx, w = prob["idx"], prob["write"]
    res = got[idx].astype(np.float64)
    err = np.linalg.norm(np.where(w, res - prob["exact"], 0))
    assert np.isfinite(err) and err <= L.driver_bound(prob, dt), (who, err)
    keep = np.ones(got.size, bool)
    keep[idx[w]] = False
    np.testing.assert_array_equal(got[keep], prob["c"][keep], err_msg=who)


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("case", L.DRIVER_CASES + L.DRIVER_MASK_ONLY, ids=lambda c: c[0])
def test_driver_descriptors(orc, dt, case):
    """orc_?driver with caller-built descriptors against the float64 closed
    form, and -- for every descriptor whose result the reference defines
    independently of its block sizes -- against mublis_l3_?driver of both
    reference builds (frame/l3/l3_driver.c:49-249, frame/l1m/packm.c:18-84):
    trapezoidal TRIANGULAR operands with NONUNIT / UNIT / INVERT diagonals,
    offset SYMMETRIC operands, ji masks on rectangular C, jp / pi equal to
    the triangular operand's uplo, strides without a unit stride."""
    prob = L.driver_problem(case, dt)
    f = _orc_driver(orc.lib, dt)
    p = prob
    out = p["c"].copy()
    st = f(p["m"], p["n"], p["k"], p["alpha"], p["a"].ctypes.data, p["rs_a"], p["cs_a"],
           p["b"].ctypes.data, p["rs_b"], p["cs_b"], p["beta"], out.ctypes.data, p["rs_c"],
           p["cs_c"], *p["desc"])
    assert st == 0
    _check_driver_out(p, out, dt, "oracle")
    if case in L.DRIVER_MASK_ONLY:
        return
    for lib_path in (L.REF_AVX2_LIB, L.REF_REF_LIB):
        if not os.path.exists(lib_path):
            pytest.skip("reference build absent")
        ref = L.driver_fn(ctypes.CDLL(lib_path), dt)
        got = p["c"].copy()
        desc = L.Product(*p["desc"])
        st = ref(p["m"], p["n"], p["k"], p["alpha"], p["a"].ctypes.data, p["rs_a"],
                 p["cs_a"], p["b"].ctypes.data, p["rs_b"], p["cs_b"], p["beta"],
                 got.ctypes.data, p["rs_c"], p["cs_c"], ctypes.byref(desc))
        assert st == 0
        _check_driver_out(p, got, dt, os.path.basename(lib_path))
        diff = np.linalg.norm(np.where(p["write"], got[p["idx"]].astype(np.float64) -
                                       out[p["idx"]].astype(np.float64), 0))
        assert diff <= 2 * L.driver_bound(p, dt), (lib_path, diff)


def test_driver_mask_is_per_term(orc):
    """The read domains are applied per term: with pi = LOWER on a GENERAL A
    (and jp = UPPER on a GENERAL B) the excluded elements of storage are NaN,
    so an oracle that read them, or multiplied them by zero, would return
    NaN; a finite result within the bound of the masked product pins it."""
    for case in L.DRIVER_MASK_ONLY:
        p = L.driver_problem(case, np.float64)
        jp, _, pi = p["desc"][:3]
        m, n, k = p["m"], p["n"], p["k"]
        a_ex = [p["a"][i * p["rs_a"] + q * p["cs_a"]] for i in range(m) for q in range(k)
                if pi and not L._rel(pi, q, i)]
        b_ex = [p["b"][q * p["rs_b"] + j * p["cs_b"]] for q in range(k) for j in range(n)
                if jp and not L._rel(jp, j, q)]
        assert (a_ex or b_ex) and np.all(np.isnan(a_ex + b_ex))
        f = _orc_driver(orc.lib, np.float64)
        out = p["c"].copy()
        f(m, n, k, p["alpha"], p["a"].ctypes.data, p["rs_a"], p["cs_a"], p["b"].ctypes.data,
          p["rs_b"], p["cs_b"], p["beta"], out.ctypes.data, p["rs_c"], p["cs_c"], *p["desc"])
        _check_driver_out(p, out, np.float64, case[0])
