"""GPU checks of the serving loop (mublis_server_* / mublis_submit /
mublis_wait), through the C ABI.

Bars:
* each request's result is bit-identical whether it runs alone or inside a
  mix submitted from several threads at once (the grouped launch runs the
  same per-tile routine in the same K order as a launch of one problem, and
  a TRSM runs the same schedule on any stream);
* each result is bit-identical to the plain cblas_* call on the same device
  operands;
* each result is within the normwise bound of the oracle computed in FP64:
  ||C - C_ref||_F <= 4 (k + 2) eps (|alpha| ||op(A)||_F ||op(B)||_F
  + |beta| ||C0||_F) for GEMM/SYRK/SYR2K (twice the product term for SYR2K),
  and ||op(A) X - alpha B||_F <= 4 n_A eps ||A||_F ||X||_F for TRSM, with eps
  of the operation's precision; the triangle a SYRK/SYR2K must not write
  keeps its bits, and a TRSM never reads A's NaN-poisoned other triangle;
* requests on one matrix keep their submission order; matrix creation and
  workspace are admitted against the budget (waiting, or MUBLIS_ENOMEM).
"""
import ctypes
import threading

import numpy as np
import pytest

import l3cases as L
import mublis_amd as M
import mublis_amd.serving as S

pytestmark = pytest.mark.gpu

EPS = {np.dtype(np.float32): np.finfo(np.float32).eps,
       np.dtype(np.float64): np.finfo(np.float64).eps}


@pytest.fixture(scope="module")
def lib():
    lib = M.load()
    assert lib.mublis_rt_init() == 0
    return lib


@pytest.fixture(scope="module")
def oracle():
    return L.Cblas(L.ORACLE_LIB, prefix="orc_", returns_status=True)


def cb_trans(t):
    return L.TR if t else L.NOTR


def cb_uplo(u):
    return L.LO if u == S.LOWER else L.UP


def spec_for(rng, op, dims, flags):
    """One request: its arguments and its host inputs (Fortran order)."""
    m, n, k = dims
    ta, tb, upper, right, unit, alpha, beta = flags
    uplo = S.UPPER if upper else S.LOWER
    f64 = op in (S.DGEMM, S.DTRSM)
    dt = np.float64 if f64 else np.float32

    def rnd(r, c):
        return np.asfortranarray(rng.uniform(-1, 1, (r, c)).astype(dt))

    if op in (S.SGEMM, S.DGEMM):
        ins = {"a": rnd(k, m) if ta else rnd(m, k), "b": rnd(n, k) if tb else rnd(k, n),
               "c": rnd(m, n)}
        args = dict(m=m, n=n, k=k, trans_a=ta, trans_b=tb, alpha=alpha, beta=beta)
        out = "c"
    elif op in (S.SSYRK, S.SSYR2K):
        ins = {"a": rnd(k, n) if ta else rnd(n, k), "c": rnd(n, n)}
        if op == S.SSYR2K:
            ins["b"] = rnd(k, n) if ta else rnd(n, k)
        args = dict(n=n, k=k, trans_a=ta, uplo=uplo, alpha=alpha, beta=beta)
        out = "c"
    else:
        na = n if right else m
        a = L.conditioned_triangular(rng, na, dt, "F", uplo == S.LOWER, unit)
        ins = {"a": np.asfortranarray(a), "b": rnd(m, n)}
        args = dict(m=m, n=n, trans_a=ta, uplo=uplo, side=S.RIGHT if right else S.LEFT,
                    diag=S.UNIT if unit else S.NON_UNIT, alpha=alpha)
        out = "b"
    return dict(op=op, args=args, inputs=ins, out=out, dtype=np.dtype(dt))


def make_mix(seed, count, lo=16, hi=384):
    rng = np.random.default_rng(seed)
    ops = [S.SGEMM] * 6 + [S.SSYRK, S.SSYR2K] * 2 + [S.DGEMM, S.STRSM, S.DTRSM, S.DTRSM]
    specs = []
    for i in range(count):
        op = ops[i % len(ops)] if i < len(ops) else ops[rng.integers(len(ops))]
        dims = tuple(int(x) for x in np.exp(rng.uniform(np.log(lo), np.log(hi), 3)))
        flags = (int(rng.integers(2)), int(rng.integers(2)), bool(rng.integers(2)),
                 bool(rng.integers(2)), bool(rng.integers(4) == 0),
                 float(rng.choice([1.0, -0.5, 1.25, 0.0])),
                 float(rng.choice([0.0, 1.0, -0.75])))
        specs.append(spec_for(rng, op, dims, flags))
    return specs


def place(server, spec):
    """Matrices for one request, holding its inputs."""
    dt = S.F64 if spec["dtype"] == np.float64 else S.F32
    return {name: server.matrix(dt, *arr.shape, data=arr)
            for name, arr in spec["inputs"].items()}


def submit(server, spec, mats):
    return server.submit(spec["op"], a=mats.get("a"), b=mats.get("b"), c=mats.get("c"),
                         **spec["args"])


def reset_inputs(spec, mats):
    mats[spec["out"]].write(spec["inputs"][spec["out"]])


def run_alone(server, specs, mats):
    outs = []
    for s, ms in zip(specs, mats):
        st, _ = server.wait(submit(server, s, ms))
        assert st == 0
        outs.append(ms[s["out"]].read())
    return outs


def run_mix(server, specs, mats, threads=8):
    tickets = [None] * len(specs)
    errors = []

    def worker(idx):
        try:
            for i in idx:
                tickets[i] = submit(server, specs[i], mats[i])
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    ts = [threading.Thread(target=worker, args=(range(t, len(specs), threads),))
          for t in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    for t in tickets:
        st, lat = server.wait(t)
        assert st == 0 and lat >= 0
    return [ms[s["out"]].read() for s, ms in zip(specs, mats)]


def untouched_mask(spec):
    """True where SYRK/SYR2K must not write (the other triangle of C)."""
    if spec["op"] not in (S.SSYRK, S.SSYR2K):
        return None
    n = spec["args"]["n"]
    lower = spec["args"]["uplo"] == S.LOWER
    return np.triu(np.ones((n, n), bool), 1) if lower else np.tril(np.ones((n, n), bool), -1)


def check_oracle(oracle, spec, got):
    a = spec["args"]
    ins = {k: v.astype(np.float64, order="F") for k, v in spec["inputs"].items()}
    eps = EPS[spec["dtype"]]
    op = spec["op"]
    if op in (S.STRSM, S.DTRSM):
        right = a["side"] == S.RIGHT
        x = np.asfortranarray(got.astype(np.float64))
        t = x.copy(order="F")
        m, n = a["m"], a["n"]
        lower = a["uplo"] == S.LOWER
        unit = a["diag"] == S.UNIT
        assert oracle.call(np.float64, "trmm", L.COL, L.RIGHT if right else L.LEFT,
                           cb_uplo(a["uplo"]), cb_trans(a["trans_a"]),
                           L.UNIT if unit else L.NONUNIT, m, n, 1.0, ins["a"],
                           ins["a"].shape[0], t, m) == 0
        res = np.linalg.norm(t - a["alpha"] * ins["b"])
        na = ins["a"].shape[0]
        bound = 4 * (na + 2) * eps * np.linalg.norm(L.tri_dense(ins["a"], lower, unit)) * \
            np.linalg.norm(x) + 4 * eps * abs(a["alpha"]) * np.linalg.norm(ins["b"])
        assert np.isfinite(got).all()
        assert res <= bound, (res, bound)
        return
    c = ins["c"].copy(order="F")
    k = a["k"]
    if op in (S.SGEMM, S.DGEMM):
        m, n = a["m"], a["n"]
        assert oracle.call(np.float64, "gemm", L.COL, cb_trans(a["trans_a"]),
                           cb_trans(a["trans_b"]), m, n, k, a["alpha"], ins["a"],
                           ins["a"].shape[0], ins["b"], ins["b"].shape[0], a["beta"], c,
                           m) == 0
        prod = np.linalg.norm(ins["a"]) * np.linalg.norm(ins["b"])
    elif op == S.SSYRK:
        n = a["n"]
        assert oracle.call(np.float64, "syrk", L.COL, cb_uplo(a["uplo"]),
                           cb_trans(a["trans_a"]), n, k, a["alpha"], ins["a"],
                           ins["a"].shape[0], a["beta"], c, n) == 0
        prod = np.linalg.norm(ins["a"]) ** 2
    else:
        n = a["n"]
        assert oracle.call(np.float64, "syr2k", L.COL, cb_uplo(a["uplo"]),
                           cb_trans(a["trans_a"]), n, k, a["alpha"], ins["a"],
                           ins["a"].shape[0], ins["b"], ins["b"].shape[0], a["beta"], c,
                           n) == 0
        prod = 2 * np.linalg.norm(ins["a"]) * np.linalg.norm(ins["b"])
    mask = untouched_mask(spec)
    g = got.astype(np.float64)
    if mask is not None:
        np.testing.assert_array_equal(got[mask], spec["inputs"]["c"][mask])
        g = np.where(mask, c, g)
    err = np.linalg.norm(g - c)
    bound = 4 * (k + 2) * eps * (abs(a["alpha"]) * prod +
                                 abs(a["beta"]) * np.linalg.norm(ins["c"]))
    assert err <= bound, (S.OP_NAMES[op], err, bound)


def cblas_call(lib, spec, mats):
    """The same request through the plain cblas_* entry point, on the
    matrices' device memory (device pointers are used in place)."""
    cb = L.Cblas(lib, prefix="cblas_")
    a = spec["args"]
    dt = spec["dtype"]
    p = {k: ctypes.c_void_p(v.ptr) for k, v in mats.items()}
    ld = {k: int(v.ld) for k, v in mats.items()}
    op = spec["op"]
    if op in (S.SGEMM, S.DGEMM):
        cb.call(dt, "gemm", L.COL, cb_trans(a["trans_a"]), cb_trans(a["trans_b"]), a["m"],
                a["n"], a["k"], a["alpha"], p["a"], ld["a"], p["b"], ld["b"], a["beta"],
                p["c"], ld["c"])
    elif op == S.SSYRK:
        cb.call(dt, "syrk", L.COL, cb_uplo(a["uplo"]), cb_trans(a["trans_a"]), a["n"], a["k"],
                a["alpha"], p["a"], ld["a"], a["beta"], p["c"], ld["c"])
    elif op == S.SSYR2K:
        cb.call(dt, "syr2k", L.COL, cb_uplo(a["uplo"]), cb_trans(a["trans_a"]), a["n"],
                a["k"], a["alpha"], p["a"], ld["a"], p["b"], ld["b"], a["beta"], p["c"],
                ld["c"])
    else:
        cb.call(dt, "trsm", L.COL, L.RIGHT if a["side"] == S.RIGHT else L.LEFT,
                cb_uplo(a["uplo"]), cb_trans(a["trans_a"]),
                L.UNIT if a["diag"] == S.UNIT else L.NONUNIT, a["m"], a["n"], a["alpha"],
                p["a"], ld["a"], p["b"], ld["b"])
    lib.mublis_rt_sync(None)


@pytest.mark.parametrize("split_flop", [-1.0, 2e7])
def test_mix_equals_alone_and_oracle(lib, oracle, split_flop):
    specs = make_mix(seed=11 if split_flop < 0 else 12, count=56)
    with S.Server(hbm_budget=2 << 30, split_flop=split_flop) as srv:
        mats = [place(srv, s) for s in specs]
        alone = run_alone(srv, specs, mats)
        for s, ms in zip(specs, mats):
            reset_inputs(s, ms)
        srv.reset_stats()
        mixed = run_mix(srv, specs, mats)
        st = srv.stats()
        for i, (s, x, y) in enumerate(zip(specs, alone, mixed)):
            assert np.array_equal(x, y, equal_nan=True), (i, S.OP_NAMES[s["op"]], s["args"])
            check_oracle(oracle, s, y)
        # the plain cblas_* path gives the same bits
        for i, (s, ms, y) in enumerate(zip(specs, mats, mixed)):
            reset_inputs(s, ms)
            cblas_call(lib, s, ms)
            z = ms[s["out"]].read()
            assert np.array_equal(y, z, equal_nan=True), (i, S.OP_NAMES[s["op"]], s["args"])
    assert st["completed"] == len(specs) and st["failed"] == 0
    assert st["grouped_requests"] > 0 and st["launches"] <= len(specs)
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["flop"] > 0 and st["tflops"] > 0


def test_requests_on_one_matrix_keep_order(lib):
    """A chain that reads and writes the same matrices, submitted all at once
    from one thread, gives the bits of the same chain run one request at a
    time."""
    rng = np.random.default_rng(5)
    n = 192
    a = np.asfortranarray(rng.uniform(-1, 1, (n, n)).astype(np.float32))
    b = np.asfortranarray(rng.uniform(-1, 1, (n, n)).astype(np.float32))
    c = np.asfortranarray(rng.uniform(-1, 1, (n, n)).astype(np.float32))
    t = np.asfortranarray(L.conditioned_triangular(rng, n, np.float32, "F", True, False))

    def chain(srv, ma, mb, mc, mt, wait_each):
        steps = [
            dict(op=S.SGEMM, m=n, n=n, k=n, alpha=1.0, beta=1.0, a=ma, b=mb, c=mc),
            dict(op=S.STRSM, m=n, n=n, alpha=1.0, uplo=S.LOWER, a=mt, b=mc),
            dict(op=S.SGEMM, m=n, n=n, k=n, alpha=0.5, beta=0.0, a=mc, b=ma, c=mb),
            dict(op=S.SSYRK, n=n, k=n, alpha=1.0, beta=1.0, uplo=S.UPPER, a=mb, c=mc),
            dict(op=S.SGEMM, m=n, n=n, k=n, alpha=1.0, beta=-1.0, a=mb, b=mc, c=ma,
                 trans_b=S.TRANS),
        ]
        tickets = []
        for st in steps:
            op = st.pop("op")
            tickets.append(srv.submit(op, **st))
            if wait_each:
                assert srv.wait(tickets.pop())[0] == 0
        for tk in tickets:
            assert srv.wait(tk)[0] == 0
        return [x.read() for x in (ma, mb, mc)]

    outs = []
    for wait_each in (True, False):
        with S.Server(hbm_budget=256 << 20) as srv:
            ms = [srv.matrix(S.F32, n, n, data=x) for x in (a, b, c, t)]
            outs.append(chain(srv, *ms, wait_each))
    for x, y in zip(*outs):
        assert np.array_equal(x, y)


def test_waiting_requests_share_one_launch(lib):
    """Requests that arrive while the GPU is busy are admitted together
    (continuous batching).  With a single launch slot, 40 small SGEMMs that
    read the output of a SSYRK queued behind two long DGEMMs all wait, then
    run as one grouped launch."""
    rng = np.random.default_rng(9)
    big, small, count = 4096, 96, 40
    with S.Server(hbm_budget=2 << 30, streams=1, slots_per_stream=1) as srv:
        x = [srv.matrix(S.F64, big, big, data=rng.uniform(-1, 1, (big, big)))
             for _ in range(3)]

        def f32():
            return srv.matrix(S.F32, small, small, data=rng.uniform(-1, 1, (small, small)))

        src, gate_in = f32(), f32()
        ins = [f32() for _ in range(count)]
        outs = [f32() for _ in range(count)]
        srv.reset_stats()
        t0 = srv.submit(S.DGEMM, m=big, n=big, k=big, a=x[0], b=x[1], c=x[2], beta=1.0)
        t1 = srv.submit(S.DGEMM, m=big, n=big, k=big, a=x[1], b=x[0], c=x[2], beta=1.0)
        gate = srv.submit(S.SSYRK, n=small, k=small, a=gate_in, c=src, beta=1.0)
        tk = [srv.submit(S.SGEMM, m=small, n=small, k=small, a=src, b=ins[i], c=outs[i])
              for i in range(count)]
        for t in [t0, t1, gate] + tk:
            assert srv.wait(t)[0] == 0
        st = srv.stats()
    assert st["completed"] == 3 + count and st["max_pending"] >= count
    assert st["grouped_requests"] == 1 + count
    # the gate alone, then the readers in (almost always) one launch
    assert st["grouped_launches"] <= 4, st


def test_replay_equals_alone(lib):
    """Requests replayed at Poisson arrival offsets by the native replay loop
    (mublis_server_replay) give the bits of the same requests run alone; a
    rejected request stops the replay with its status after the requests
    before it were queued."""
    specs = make_mix(seed=13, count=40)
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(2e-4, len(specs)))
    with S.Server(hbm_budget=2 << 30) as srv:
        mats = [place(srv, s) for s in specs]
        alone = run_alone(srv, specs, mats)
        for s, ms in zip(specs, mats):
            reset_inputs(s, ms)
        reqs = [S.Server.request(s["op"], a=ms.get("a"), b=ms.get("b"), c=ms.get("c"),
                                 **s["args"]) for s, ms in zip(specs, mats)]
        tickets, lag = srv.replay(reqs, arrivals)
        assert len(tickets) == len(specs) and lag >= 0
        for t in tickets:
            assert srv.wait(t)[0] == 0
        for i, (s, ms, x) in enumerate(zip(specs, mats, alone)):
            y = ms[s["out"]].read()
            assert np.array_equal(x, y, equal_nan=True), (i, S.OP_NAMES[s["op"]], s["args"])
        # the third request is invalid (k beyond A): two are queued, then EINVAL
        a = srv.matrix(S.F32, 64, 32)
        c = [srv.matrix(S.F32, 64, 64) for _ in range(3)]
        bad = [S.Server.request(S.SSYRK, n=64, k=32, a=a, c=c[0], uplo=S.LOWER),
               S.Server.request(S.SSYRK, n=64, k=32, a=a, c=c[1], uplo=S.LOWER),
               S.Server.request(S.SGEMM, m=64, n=64, k=48, a=a, b=a, c=c[2])]
        n = len(bad)
        arr = (S.Request * n)(*bad)
        at = (ctypes.c_double * n)(0.0, 0.0, 0.0)
        tk = (ctypes.c_uint64 * n)()
        done, lg = ctypes.c_int(), ctypes.c_double()
        st = lib.mublis_server_replay(srv.handle, ctypes.cast(arr, ctypes.c_void_p),
                                      ctypes.cast(at, ctypes.c_void_p), n,
                                      ctypes.cast(tk, ctypes.c_void_p), ctypes.byref(done),
                                      ctypes.byref(lg))
        assert st == M.EINVAL and done.value == 2
        for i in range(2):
            assert srv.wait(tk[i])[0] == 0
        for x in [a] + c:
            x.destroy()


def test_budget_admission_and_rejections(lib):
    with S.Server(hbm_budget=64 << 20) as srv:
        big = 1536  # 9 MiB per FP32 matrix; 56 MiB are left for matrices
        held = []
        with pytest.raises(RuntimeError, match="status 2"):
            for _ in range(16):
                held.append(srv.matrix(S.F32, big, big, wait=False))
        assert 4 <= len(held) < 8
        with pytest.raises(RuntimeError, match="status 2"):
            srv.matrix(S.F32, 8192, 8192)  # larger than the budget: fails at once
        # a waiting creation is admitted when another thread releases memory
        got = []
        th = threading.Thread(target=lambda: got.append(srv.matrix(S.F32, big, big)))
        th.start()
        th.join(0.3)
        assert th.is_alive() and not got
        held.pop().destroy()
        th.join(10)
        assert got and not th.is_alive()
        held.append(got[0])
        # invalid requests are rejected at submission
        a = srv.matrix(S.F32, 64, 32)
        c = srv.matrix(S.F32, 64, 64)
        d = srv.matrix(S.F64, 64, 64)
        assert srv.submit_status(S.SGEMM, m=64, n=64, k=48, a=a, b=a, c=c)[0] == M.EINVAL
        assert srv.submit_status(S.SGEMM, m=64, n=64, k=32, a=a, b=c, c=c)[0] == M.EINVAL
        assert srv.submit_status(S.DGEMM, m=64, n=64, k=32, a=a, b=a, c=d)[0] == M.EINVAL
        assert srv.submit_status(S.DTRSM, m=64, n=64, a=d, b=d)[0] == M.EINVAL
        assert srv.submit_status(S.SSYRK, n=64, k=32, a=a, c=c, uplo=0)[0] == M.EINVAL
        st, tk = srv.submit_status(S.SSYRK, n=64, k=32, a=a, c=c, uplo=S.LOWER)
        assert st == 0 and srv.wait(tk)[0] == 0
        assert srv.wait(tk)[0] == M.EINVAL  # each ticket is waited once
        for x in held + [a, c, d]:
            x.destroy()
        assert srv.stats()["hbm_used"] == 0


def test_destroy_with_requests_in_flight(lib):
    """Matrices destroyed while their requests run are released afterwards,
    and closing the server completes what was submitted."""
    rng = np.random.default_rng(3)
    srv = S.Server(hbm_budget=512 << 20)
    try:
        mats = []
        for _ in range(12):
            n = int(rng.integers(256, 768))
            x = [srv.matrix(S.F32, n, n, data=rng.uniform(-1, 1, (n, n))) for _ in range(3)]
            srv.submit(S.SGEMM, m=n, n=n, k=n, a=x[0], b=x[1], c=x[2])
            for h in x:
                h.destroy()
            mats.append(x)
    finally:
        srv.close()
