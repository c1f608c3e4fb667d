"""Shared helpers for the L3 parity tests.

`Cblas` wraps any library with the CBLAS L3 signatures (the product's
``cblas_*``, the reference build's ``cblas_*`` or the oracle's ``orc_*``),
taking numpy arrays.  The case generators follow the reference's own
correctness harness (test/correctness/check_cblas.py:157-473): seed 314159,
5x4x3-sized problems, NaN in every element the routine must not read, and
the untouched triangle checked bit for bit.
"""
import ctypes
import itertools
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_AVX2_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_avx2.so")
REF_REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_ref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")

ROW, COL = 101, 102
NOTR, TR = 111, 112
UP, LO = 121, 122
NONUNIT, UNIT = 131, 132
LEFT, RIGHT = 141, 142

# check_cblas.py:24-37
TOL = {np.dtype(np.float32): (5e-4, 5e-5), np.dtype(np.float64): (2e-11, 1e-12)}
SEED = 314159
DTYPES = [np.dtype(np.float32), np.dtype(np.float64)]
LAYOUTS = [("C", ROW), ("F", COL)]

_VP = ctypes.c_void_p
_I = ctypes.c_int


def ld_of(x, order):
    axis = 0 if order == "C" else 1
    return x.strides[axis] // x.itemsize


def ptr(x):
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return _VP(x.data_ptr())
    return _VP(x.ctypes.data)


class Cblas:
    """Call CBLAS-shaped routines of a ctypes library by prefix."""

    def __init__(self, path_or_lib, prefix="cblas_", returns_status=False):
        self.lib = (ctypes.CDLL(path_or_lib) if isinstance(path_or_lib, str)
                    else path_or_lib)
        self.prefix = prefix
        self.returns_status = returns_status
        for p, ct in (("s", ctypes.c_float), ("d", ctypes.c_double)):
            sig = {
                "gemm": [_I, _I, _I, _I, _I, _I, ct, _VP, _I, _VP, _I, ct, _VP, _I],
                "symm": [_I, _I, _I, _I, _I, ct, _VP, _I, _VP, _I, ct, _VP, _I],
                "syrk": [_I, _I, _I, _I, _I, ct, _VP, _I, ct, _VP, _I],
                "syr2k": [_I, _I, _I, _I, _I, ct, _VP, _I, _VP, _I, ct, _VP, _I],
                "trmm": [_I, _I, _I, _I, _I, _I, _I, ct, _VP, _I, _VP, _I],
                "trsm": [_I, _I, _I, _I, _I, _I, _I, ct, _VP, _I, _VP, _I],
            }
            for r, args in sig.items():
                f = getattr(self.lib, f"{prefix}{p}{r}")
                f.argtypes = args
                f.restype = _I if returns_status else None

    def fn(self, dtype, routine):
        p = "s" if np.dtype(dtype) == np.float32 else "d"
        return getattr(self.lib, f"{self.prefix}{p}{routine}")

    def call(self, dtype, routine, *args):
        conv = [ptr(a) if (isinstance(a, np.ndarray) or hasattr(a, "data_ptr"))
                else a for a in args]
        return self.fn(dtype, routine)(*conv)


def make(values, dtype, order):
    return np.array(values, dtype=dtype, order=order, copy=True)


def rand(rng, shape, dtype, order):
    return make(rng.standard_normal(shape), dtype, order)


def symmetric_storage(rng, size, dtype, order, lower):
    """check_cblas.py:192-200: the unstored triangle is NaN."""
    v = rng.standard_normal((size, size))
    if lower:
        v[np.triu_indices(size, 1)] = np.nan
    else:
        v[np.tril_indices(size, -1)] = np.nan
    return make(v, dtype, order)


def triangular_storage(rng, size, dtype, order, lower, unit):
    """check_cblas.py:202-222: NaN outside the triangle and on a unit
    diagonal; off-diagonal U(-0.25, 0.25), diagonal U(1.5, 2.5)."""
    v = rng.uniform(-0.25, 0.25, size=(size, size))
    if lower:
        v[np.triu_indices(size, 1)] = np.nan
    else:
        v[np.tril_indices(size, -1)] = np.nan
    if unit:
        np.fill_diagonal(v, np.nan)
    else:
        np.fill_diagonal(v, rng.uniform(1.5, 2.5, size=size))
    return make(v, dtype, order)


def conditioned_triangular(rng, size, dtype, order, lower, unit):
    """Larger triangular operands: off-diagonal N(0,1)/sqrt(size), diagonal
    1 + |N(0,1)| (the well-conditioned DTRSM workload's distribution), NaN
    outside the triangle and on a unit diagonal."""
    v = rng.standard_normal((size, size)) / np.sqrt(size)
    if lower:
        v[np.triu_indices(size, 1)] = np.nan
    else:
        v[np.tril_indices(size, -1)] = np.nan
    if unit:
        np.fill_diagonal(v, np.nan)
    else:
        np.fill_diagonal(v, 1.0 + np.abs(rng.standard_normal(size)))
    return make(v, dtype, order)


def tri_dense(a, lower, unit):
    """Dense float64 triangular matrix from poisoned storage."""
    t = np.tril(np.nan_to_num(a.astype(np.float64), nan=0.0)) if lower else \
        np.triu(np.nan_to_num(a.astype(np.float64), nan=0.0))
    if unit:
        np.fill_diagonal(t, 1.0)
    return t


def sym_dense(a, lower):
    x = np.nan_to_num(a.astype(np.float64), nan=0.0)
    t = np.tril(x) if lower else np.triu(x)
    return t + t.T - np.diag(np.diag(t))


# ---------------------------------------------------------------- cases ----
def gemm_cases():
    for dt, (order, co), (ta, tb) in itertools.product(
            DTYPES, LAYOUTS, itertools.product((False, True), repeat=2)):
        yield dict(routine="gemm", dtype=dt, order=order, corder=co, ta=ta, tb=tb)


def symm_cases():
    for dt, (order, co), right, lower in itertools.product(
            DTYPES, LAYOUTS, (False, True), (False, True)):
        yield dict(routine="symm", dtype=dt, order=order, corder=co,
                   right=right, lower=lower)


def rk_cases(routine):
    for dt, (order, co), trans, lower in itertools.product(
            DTYPES, LAYOUTS, (False, True), (False, True)):
        yield dict(routine=routine, dtype=dt, order=order, corder=co,
                   trans=trans, lower=lower)


def tr_cases(routine):
    for dt, (order, co), (right, lower, trans, unit) in itertools.product(
            DTYPES, LAYOUTS, itertools.product((False, True), repeat=4)):
        yield dict(routine=routine, dtype=dt, order=order, corder=co,
                   right=right, lower=lower, trans=trans, unit=unit)


def all_cases():
    """The reference harness's 192 cases, in a fixed order."""
    out = list(gemm_cases()) + list(symm_cases()) + list(rk_cases("syrk")) + \
        list(rk_cases("syr2k")) + list(tr_cases("trmm")) + list(tr_cases("trsm"))
    return out


def case_id(c):
    keys = [k for k in c if k not in ("routine", "dtype", "order", "corder")]
    flags = "-".join(f"{k}{int(c[k])}" for k in keys)
    return f"{c['routine']}-{c['dtype'].name}-{c['order']}-{flags}"


def build_problem(c, rng):
    """Inputs for one case, drawn in the reference harness's order.  Returns
    (args-builder, outputs): `args(out)` gives the CBLAS argument list with
    `out` as the output array."""
    dt, order, co = c["dtype"], c["order"], c["corder"]
    r = c["routine"]
    m, n, k = c.get("m", 5), c.get("n", 4), c.get("k", 3)
    if r == "gemm":
        a = rand(rng, (k, m) if c["ta"] else (m, k), dt, order)
        b = rand(rng, (n, k) if c["tb"] else (k, n), dt, order)
        c0 = rand(rng, (m, n), dt, order)
        alpha, beta = 1.25, -0.5

        def args(out):
            return (co, TR if c["ta"] else NOTR, TR if c["tb"] else NOTR, m, n, k,
                    alpha, a, ld_of(a, order), b, ld_of(b, order), beta, out,
                    ld_of(out, order))
        ref = alpha * (a.astype(np.float64).T if c["ta"] else a.astype(np.float64)) @ \
            (b.astype(np.float64).T if c["tb"] else b.astype(np.float64)) + beta * c0.astype(np.float64)
        return dict(inputs=dict(a=a, b=b), c0=c0, args=args, exact=ref, mask=None)
    if r == "symm":
        mm, nn = c.get("m", 5), c.get("n", 3)
        size = nn if c["right"] else mm
        a = symmetric_storage(rng, size, dt, order, c["lower"])
        b = rand(rng, (mm, nn), dt, order)
        c0 = rand(rng, (mm, nn), dt, order)
        alpha, beta = 0.75, -0.25
        s = sym_dense(a, c["lower"])
        ref = alpha * (b.astype(np.float64) @ s if c["right"] else s @ b.astype(np.float64)) \
            + beta * c0.astype(np.float64)

        def args(out):
            return (co, RIGHT if c["right"] else LEFT, LO if c["lower"] else UP, mm, nn,
                    alpha, a, ld_of(a, order), b, ld_of(b, order), beta, out,
                    ld_of(out, order))
        return dict(inputs=dict(a=a, b=b), c0=c0, args=args, exact=ref, mask=None)
    if r in ("syrk", "syr2k"):
        nn, kk = c.get("n", 5), c.get("k", 3)
        shape = (kk, nn) if c["trans"] else (nn, kk)
        a = rand(rng, shape, dt, order)
        b = rand(rng, shape, dt, order) if r == "syr2k" else None
        c0 = rand(rng, (nn, nn), dt, order)
        alpha, beta = (1.25, -0.5) if r == "syrk" else (0.75, -0.25)
        A = a.astype(np.float64)
        A = A.T if c["trans"] else A
        if r == "syrk":
            ref = alpha * A @ A.T + beta * c0.astype(np.float64)
        else:
            B = b.astype(np.float64)
            B = B.T if c["trans"] else B
            ref = alpha * (A @ B.T + B @ A.T) + beta * c0.astype(np.float64)
        mask = np.tril(np.ones((nn, nn), bool)) if c["lower"] else np.triu(np.ones((nn, nn), bool))

        def args(out):
            lst = [co, LO if c["lower"] else UP, TR if c["trans"] else NOTR, nn, kk,
                   alpha, a, ld_of(a, order)]
            if r == "syr2k":
                lst += [b, ld_of(b, order)]
            lst += [beta, out, ld_of(out, order)]
            return tuple(lst)
        return dict(inputs=dict(a=a, b=b), c0=c0, args=args, exact=ref, mask=mask)
    if r in ("trmm", "trsm"):
        mm, nn = c.get("m", 5), c.get("n", 3)
        size = nn if c["right"] else mm
        if size <= 8:
            a = triangular_storage(rng, size, dt, order, c["lower"], c["unit"])
        else:
            a = conditioned_triangular(rng, size, dt, order, c["lower"], c["unit"])
        b0 = rand(rng, (mm, nn), dt, order)
        alpha = 0.75
        t = tri_dense(a, c["lower"], c["unit"])
        op = t.T if c["trans"] else t
        B = b0.astype(np.float64)
        if r == "trmm":
            ref = alpha * (B @ op if c["right"] else op @ B)
        else:
            ref = alpha * (np.linalg.solve(op.T, B.T).T if c["right"] else np.linalg.solve(op, B))

        def args(out):
            return (co, RIGHT if c["right"] else LEFT, LO if c["lower"] else UP,
                    TR if c["trans"] else NOTR, UNIT if c["unit"] else NONUNIT, mm, nn,
                    alpha, a, ld_of(a, order), out, ld_of(out, order))
        return dict(inputs=dict(a=a), c0=b0, args=args, exact=ref, mask=None)
    raise ValueError(r)


def assert_ref_close(actual, expected, dtype, mask=None, initial=None):
    """Reference tolerance (check_cblas.py:182-190) on the written part and
    bit equality on the untouched triangle (check_cblas.py:355-363)."""
    rtol, atol = TOL[np.dtype(dtype)]
    if mask is None:
        np.testing.assert_allclose(actual, expected, rtol=rtol, atol=atol)
    else:
        np.testing.assert_allclose(actual[mask], expected[mask], rtol=rtol, atol=atol)
        np.testing.assert_array_equal(actual[~mask], initial[~mask])


def ragged_cases():
    """Larger, ragged problems: several micro-tiles of the reference's avx2
    target (S 8x8, D 12x4; blocks D 72/252/512), edge tiles in every
    dimension, and for TRSM/TRMM more than one kc block, so the full-tile
    path and the trailing update (frame/l3/trsm.c:201-274) are exercised."""
    out = []
    for dt in DTYPES:
        for (ta, tb) in itertools.product((False, True), repeat=2):
            out.append(dict(routine="gemm", dtype=dt, order="F", corder=COL, ta=ta,
                            tb=tb, m=67, n=45, k=33))
        out.append(dict(routine="gemm", dtype=dt, order="C", corder=ROW, ta=False,
                        tb=True, m=100, n=60, k=80))
        for right, lower in itertools.product((False, True), repeat=2):
            out.append(dict(routine="symm", dtype=dt, order="F", corder=COL,
                            right=right, lower=lower, m=70, n=50))
        for r in ("syrk", "syr2k"):
            for trans, lower in itertools.product((False, True), repeat=2):
                out.append(dict(routine=r, dtype=dt, order="F", corder=COL,
                                trans=trans, lower=lower, n=70, k=45))
        for r in ("trmm", "trsm"):
            for right, lower, trans, unit in itertools.product((False, True), repeat=4):
                mm, nn = (76, 17) if not right else (17, 40)
                if dt == np.float32:
                    mm, nn = (48, 19) if not right else (19, 48)
                out.append(dict(routine=r, dtype=dt, order="F", corder=COL,
                                right=right, lower=lower, trans=trans, unit=unit,
                                m=mm, n=nn))
    return out


# ------------------------------------------------------- golden vectors --
def load_golden(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    meta = json.loads(str(z["meta"]))
    cases = []
    for i, m in enumerate(meta):
        c = dict(m)
        c["dtype"] = np.dtype(c["dtype"])
        arr = {k: z[f"{i}_{k}"] for k in ("a", "b", "c0", "avx2", "ref") if f"{i}_{k}" in z.files}
        cases.append((c, arr))
    return cases


def args_from_golden(c, arr, out):
    """Rebuild the CBLAS argument list for a golden case from its arrays."""
    r, co, order = c["routine"], c["corder"], c["order"]
    a, b = arr["a"], arr.get("b")
    lda = ld_of(a, order)
    if r == "gemm":
        al, be = (1.25, -0.5)
        m, n = out.shape
        k = a.shape[0] if c["ta"] else a.shape[1]
        return (co, TR if c["ta"] else NOTR, TR if c["tb"] else NOTR, m, n, k,
                al, a, lda, b, ld_of(b, order), be, out, ld_of(out, order))
    if r == "symm":
        m, n = out.shape
        return (co, RIGHT if c["right"] else LEFT, LO if c["lower"] else UP, m, n,
                0.75, a, lda, b, ld_of(b, order), -0.25, out, ld_of(out, order))
    if r in ("syrk", "syr2k"):
        n = out.shape[0]
        k = a.shape[0] if c["trans"] else a.shape[1]
        al, be = (1.25, -0.5) if r == "syrk" else (0.75, -0.25)
        lst = [co, LO if c["lower"] else UP, TR if c["trans"] else NOTR, n, k, al,
               a, lda]
        if r == "syr2k":
            lst += [b, ld_of(b, order)]
        return tuple(lst + [be, out, ld_of(out, order)])
    m, n = out.shape
    return (co, RIGHT if c["right"] else LEFT, LO if c["lower"] else UP,
            TR if c["trans"] else NOTR, UNIT if c["unit"] else NONUNIT, m, n,
            0.75, a, lda, out, ld_of(out, order))


def golden_mask(c, n):
    if c["routine"] in ("syrk", "syr2k"):
        return np.tril(np.ones((n, n), bool)) if c["lower"] else np.triu(np.ones((n, n), bool))
    return None




# ------------------------------------------------ caller-built descriptors --
# Generic-driver calls with descriptors the caller builds
# (include/mublis.h:41-83): trapezoidal TRIANGULAR operands (m != k), the
# UNIT and INVERT diagonals, SYMMETRIC operands on either side (offset when
# m != k), ji masks on rectangular C, jp / pi read domains, and strides with
# no unit stride.  Enum values are mublis.h's: struc G/S/T = 0/1/2, uplo
# dense/lower/upper = 0/1/2, diag nonunit/unit/invert = 0/1/2, relations
# all/lower/upper = 0/1/2.
G, S, T = 0, 1, 2
DENSE, MLO, MUP = 0, 1, 2
DN, DU, DI = 0, 1, 2
RALL, RLO, RUP = 0, 1, 2

# name, m, n, k, A (struc, uplo, diag), B (struc, uplo, diag),
# (jp, ji, pi), alpha, beta, storage ("col", "row" or "gen" per operand A, B, C)
DRIVER_CASES = [
    ("triA_lo_tall", 150, 70, 110, (T, MLO, DN), (G, DENSE, DN), (RALL, RALL, RLO), 1.25, -0.5,
     ("col", "col", "col")),
    ("triA_up_wide_unit", 90, 60, 170, (T, MUP, DU), (G, DENSE, DN), (RALL, RALL, RUP), 0.75,
     0.0, ("col", "row", "col")),
    ("triA_lo_invert", 101, 53, 101, (T, MLO, DI), (G, DENSE, DN), (RALL, RALL, RALL), 1.0, 1.0,
     ("row", "col", "col")),
    ("triA_up_tall_invert", 130, 40, 75, (T, MUP, DI), (G, DENSE, DN), (RALL, RALL, RUP), -1.5,
     0.25, ("col", "col", "row")),
    ("triB_lo_tall", 64, 90, 140, (G, DENSE, DN), (T, MLO, DN), (RLO, RALL, RALL), 1.0, 0.5,
     ("col", "col", "col")),
    ("triB_up_wide_unit", 77, 160, 100, (G, DENSE, DN), (T, MUP, DU), (RUP, RALL, RALL), 2.0,
     0.0, ("row", "row", "col")),
    ("triB_lo_invert", 50, 96, 96, (G, DENSE, DN), (T, MLO, DI), (RLO, RALL, RALL), 1.0, -1.0,
     ("col", "gen", "col")),
    ("symA_lo_tall", 140, 50, 90, (S, MLO, DN), (G, DENSE, DN), (RALL, RALL, RALL), 1.0, 0.5,
     ("col", "col", "col")),
    ("symA_up_wide", 70, 45, 130, (S, MUP, DN), (G, DENSE, DN), (RALL, RALL, RALL), -0.5, 1.0,
     ("row", "col", "col")),
    ("symB_lo", 60, 88, 88, (G, DENSE, DN), (S, MLO, DN), (RALL, RALL, RALL), 1.0, 0.0,
     ("col", "col", "col")),
    ("symB_up_offset", 66, 50, 120, (G, DENSE, DN), (S, MUP, DN), (RALL, RALL, RALL), 1.5, 0.5,
     ("col", "row", "col")),
    ("ji_lo_rect", 130, 95, 60, (G, DENSE, DN), (G, DENSE, DN), (RALL, RLO, RALL), 1.0, 0.5,
     ("col", "col", "col")),
    ("ji_up_rect", 80, 140, 70, (G, DENSE, DN), (G, DENSE, DN), (RALL, RUP, RALL), 1.0, 2.0,
     ("row", "col", "row")),
    ("both_tri", 120, 110, 115, (T, MLO, DU), (T, MUP, DI), (RUP, RALL, RLO), 1.0, 0.5,
     ("col", "col", "col")),
    ("tri_ji_gen_strides", 97, 83, 91, (T, MUP, DN), (S, MLO, DN), (RALL, RLO, RUP), 0.5, -0.5,
     ("gen", "gen", "gen")),
    ("alpha0_tri", 40, 30, 50, (T, MLO, DU), (G, DENSE, DN), (RALL, RUP, RLO), 0.0, 0.5,
     ("col", "col", "col")),
    ("deep_k_blocks", 60, 36, 300, (T, MLO, DN), (T, MLO, DN), (RLO, RALL, RLO), 1.0, 0.0,
     ("col", "col", "col")),
]

# Inconsistent read domains (jp / pi on GENERAL operands): the reference's
# result depends on its cache-block sizes there, so only the per-term
# definition (and the product against the oracle) is checked.
DRIVER_MASK_ONLY = [
    ("mask_pi_general", 90, 70, 80, (G, DENSE, DN), (G, DENSE, DN), (RALL, RALL, RLO), 1.0, 0.5,
     ("col", "col", "col")),
    ("mask_jp_general", 70, 90, 85, (G, DENSE, DN), (G, DENSE, DN), (RUP, RLO, RALL), 1.0, 0.0,
     ("row", "col", "col")),
]


def _extent(struc, uplo, rows, cols):
    """Storage box an operand view reads (the reflected reads of SYMMETRIC)."""
    if struc == S and rows != cols:
        lo, hi = min(rows, cols), max(rows, cols)
        if uplo == MLO:
            return hi, lo
        if uplo == MUP:
            return lo, hi
    return rows, cols


def _strided(rows, cols, kind):
    """(flat size, rs, cs) of rows x cols storage of a given kind."""
    if kind == "col":
        ld = rows + 3
        return ld * cols, 1, ld
    if kind == "row":
        ld = cols + 2
        return ld * rows, ld, 1
    rs, cs = 2, 2 * rows + 3  # no unit stride
    return (rows - 1) * rs + (cols - 1) * cs + 1, rs, cs


def _stored(uplo, d):
    return d >= 0 if uplo == MLO else (d <= 0 if uplo == MUP else True)


def _rel(rel, lhs, rhs):
    return lhs <= rhs if rel == RLO else (lhs >= rhs if rel == RUP else True)


def _operand(rng, rows, cols, desc, mask, kind, dt):
    """Storage for a rows x cols operand view (flat buffer + strides), NaN
    wherever the view must not read, and the view itself in float64."""
    struc, uplo, diag = desc
    er, ec = _extent(struc, uplo, rows, cols)
    size, rs, cs = _strided(er, ec, kind)
    buf = np.full(size, np.nan)
    vals = rng.standard_normal((er, ec))
    for i in range(er):
        for p in range(ec):
            d = i - p
            if struc == T and (not _stored(uplo, d) or (d == 0 and diag == DU)):
                continue
            if struc == S and not _stored(uplo, d):
                continue
            v = vals[i, p]
            if struc == T and d == 0 and diag == DI:
                v = rng.uniform(1.5, 2.5) * (1 if v >= 0 else -1)
            if struc == G and i < rows and p < cols and not _rel(mask, p, i):
                continue  # excluded by the read domain: never read
            buf[i * rs + p * cs] = v
    buf = buf.astype(dt)
    x = buf.astype(np.float64)
    view = np.zeros((rows, cols))
    for i in range(rows):
        for p in range(cols):
            d = i - p
            if not _rel(mask, p, i):
                continue
            if struc == S:
                view[i, p] = x[i * rs + p * cs] if _stored(uplo, d) else x[p * rs + i * cs]
            elif struc == T:
                if not _stored(uplo, d):
                    continue
                if d == 0 and diag == DU:
                    view[i, p] = 1.0
                elif d == 0 and diag == DI:
                    view[i, p] = 1.0 / x[i * rs + p * cs]
                else:
                    view[i, p] = x[i * rs + p * cs]
            else:
                view[i, p] = x[i * rs + p * cs]
    return buf, rs, cs, view


def driver_problem(case, dt, seed=SEED):
    """Inputs of one caller-built driver call, its expected value in float64,
    and the mask of C elements it may write."""
    name, m, n, k, da, db, (jp, ji, pi), alpha, beta, kinds = case
    rng = np.random.default_rng([seed, sum(map(ord, name))])
    a, rs_a, cs_a, av = _operand(rng, m, k, da, pi, kinds[0], dt)
    # jp(j, p) for B(p, j) is "column <= row" for LOWER, like pi for A(i, p)
    b, rs_b, cs_b, bv = _operand(rng, k, n, db, jp, kinds[1], dt)
    size, rs_c, cs_c = _strided(m, n, kinds[2])
    c = np.full(size, np.nan)
    idx = np.add.outer(np.arange(m) * rs_c, np.arange(n) * cs_c)
    c[idx] = rng.standard_normal((m, n))
    c = c.astype(dt)
    if beta == 0:
        c[idx] = np.nan  # beta == 0 never reads C
    write = np.array([[_rel(ji, j, i) for j in range(n)] for i in range(m)])
    c0 = c[idx].astype(np.float64)
    exact = alpha * (av @ bv) if alpha != 0 else np.zeros((m, n))
    if beta != 0:
        exact = exact + beta * c0
    return dict(m=m, n=n, k=k, alpha=alpha, beta=beta, a=a, rs_a=rs_a, cs_a=cs_a, b=b,
                rs_b=rs_b, cs_b=cs_b, c=c, rs_c=rs_c, cs_c=cs_c, idx=idx, write=write,
                exact=exact, av=av, bv=bv, c0=c0,
                desc=(jp, ji, pi) + tuple(da) + tuple(db))


class Product(ctypes.Structure):
    """mublis_l3_product_t (include/mublis.h:41-63): nine enum ints."""
    _fields_ = [(f, ctypes.c_int) for f in
                ("jp", "ji", "pi", "a_struc", "a_uplo", "a_diag", "b_struc", "b_uplo",
                 "b_diag")]


def driver_fn(lib, dt):
    """The library's mublis_l3_?driver with its argtypes set."""
    ct = ctypes.c_float if np.dtype(dt) == np.float32 else ctypes.c_double
    f = getattr(lib, "mublis_l3_sdriver" if ct is ctypes.c_float else "mublis_l3_ddriver")
    f.argtypes = [_I, _I, _I, ct, _VP, _I, _I, _VP, _I, _I, ct, _VP, _I, _I,
                  ctypes.POINTER(Product)]
    f.restype = _I
    return f


def driver_bound(prob, dt):
    """Normwise bound on the computed C: 2 (k + 2) eps (|alpha| ||Av|| ||Bv||
    + |beta| ||C0||) over the written elements."""
    eps = np.finfo(dt).eps
    nb = abs(prob["beta"]) * np.linalg.norm(np.where(prob["write"], prob["c0"], 0)) \
        if prob["beta"] else 0.0
    return 2 * (prob["k"] + 2) * eps * (abs(prob["alpha"]) * np.linalg.norm(prob["av"]) *
                                        np.linalg.norm(prob["bv"]) + nb)
