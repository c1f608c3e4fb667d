"""Python mirror of the serving loop (include/mublis_gfx950.h, mublis_server_*).

Thin ctypes wrappers: a ``Server`` owns the C server handle, ``Matrix``
objects are its device-resident operands, and ``submit`` takes the same
arguments as the cblas_* call a request stands for (column-major).  Every
call goes to libmublis.so; there is no Python-side execution path.
"""
import ctypes

import numpy as np

from . import _VP, _I, _D, check, load

F32, F64 = 0, 1
SGEMM, SSYRK, SSYR2K, DGEMM, STRSM, DTRSM = range(6)
NO_TRANS, TRANS = 0, 1
LOWER, UPPER = 1, 2
LEFT, RIGHT = 0, 1
NON_UNIT, UNIT = 0, 1
OP_NAMES = {SGEMM: "sgemm", SSYRK: "ssyrk", SSYR2K: "ssyr2k", DGEMM: "dgemm",
            STRSM: "strsm", DTRSM: "dtrsm"}


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("streams", _I), ("slots_per_stream", _I),
                ("max_group", _I), ("split_flop", _D)]


class Request(ctypes.Structure):
    _fields_ = [("op", _I), ("trans_a", _I), ("trans_b", _I), ("side", _I), ("uplo", _I),
                ("diag", _I), ("m", _I), ("n", _I), ("k", _I), ("alpha", _D), ("beta", _D),
                ("a", _VP), ("b", _VP), ("c", _VP)]


class Stats(ctypes.Structure):
    _U = ctypes.c_uint64
    _fields_ = [("submitted", _U), ("completed", _U), ("failed", _U), ("launches", _U),
                ("grouped_launches", _U), ("grouped_requests", _U), ("max_pending", _U),
                ("max_inflight", _U), ("p50_ms", _D), ("p99_ms", _D), ("mean_ms", _D),
                ("max_ms", _D), ("flop", _D), ("elapsed_s", _D), ("tflops", _D),
                ("hbm_budget", ctypes.c_size_t), ("hbm_used", ctypes.c_size_t),
                ("hbm_peak", ctypes.c_size_t)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Matrix:
    """A rows x cols column-major matrix held by the server."""

    def __init__(self, server, dtype, rows, cols, wait=True):
        self.server, self.dtype, self.rows, self.cols = server, dtype, rows, cols
        self.handle = ctypes.c_void_p()
        check(server.lib.mublis_matrix_create(server.handle, dtype, rows, cols, int(wait),
                                              ctypes.byref(self.handle)),
              "mublis_matrix_create")
        self.ld = server.lib.mublis_matrix_ld(self.handle)
        self.ptr = server.lib.mublis_matrix_device_ptr(self.handle)

    @property
    def np_dtype(self):
        return np.float64 if self.dtype == F64 else np.float32

    def write(self, host):
        host = np.asfortranarray(host, dtype=self.np_dtype)
        assert host.shape == (self.rows, self.cols)
        check(self.server.lib.mublis_matrix_write(self.server.handle, self.handle,
                                                  host.ctypes.data, max(1, self.rows)),
              "mublis_matrix_write")

    def read(self):
        out = np.empty((self.rows, self.cols), dtype=self.np_dtype, order="F")
        check(self.server.lib.mublis_matrix_read(self.server.handle, self.handle,
                                                 out.ctypes.data, max(1, self.rows)),
              "mublis_matrix_read")
        return out

    def destroy(self):
        if self.handle:
            check(self.server.lib.mublis_matrix_destroy(self.server.handle, self.handle),
                  "mublis_matrix_destroy")
            self.handle = ctypes.c_void_p()


class Server:
    def __init__(self, hbm_budget=1 << 30, streams=0, slots_per_stream=0, max_group=0,
                 split_flop=0.0):
        self.lib = load()
        self.cfg = ServerConfig(hbm_budget, streams, slots_per_stream, max_group, split_flop)
        self.handle = ctypes.c_void_p()
        check(self.lib.mublis_server_create(ctypes.byref(self.cfg), ctypes.byref(self.handle)),
              "mublis_server_create")

    def matrix(self, dtype, rows, cols, data=None, wait=True):
        m = Matrix(self, dtype, rows, cols, wait)
        if data is not None:
            m.write(data)
        return m

    @staticmethod
    def request(op, m=0, n=0, k=0, alpha=1.0, beta=0.0, a=None, b=None, c=None,
                trans_a=NO_TRANS, trans_b=NO_TRANS, side=LEFT, uplo=LOWER, diag=NON_UNIT):
        """The mublis_request_t of a request, for submit_request."""
        return Request(op, trans_a, trans_b, side, uplo, diag, m, n, k, alpha, beta,
                       a.handle if a is not None else None,
                       b.handle if b is not None else None,
                       c.handle if c is not None else None)

    def submit(self, op, **kw):
        """Queue a request; returns its ticket.  Raises on a rejected request."""
        return self.submit_request(self.request(op, **kw))

    def submit_request(self, req):
        """Queue a request built by request(); returns its ticket.  Raises on a
        rejected request."""
        t = ctypes.c_uint64()
        check(self.lib.mublis_submit(self.handle, ctypes.byref(req), ctypes.byref(t)),
              "mublis_submit")
        return t.value

    def submit_status(self, op, **kw):
        """Like submit, but returns (status, ticket) instead of raising."""
        req = Request(op, kw.get("trans_a", 0), kw.get("trans_b", 0), kw.get("side", 0),
                      kw.get("uplo", LOWER), kw.get("diag", 0), kw.get("m", 0),
                      kw.get("n", 0), kw.get("k", 0), kw.get("alpha", 1.0),
                      kw.get("beta", 0.0),
                      *[kw[x].handle if kw.get(x) is not None else None for x in "abc"])
        t = ctypes.c_uint64()
        st = self.lib.mublis_submit(self.handle, ctypes.byref(req), ctypes.byref(t))
        return st, t.value

    def replay(self, requests, arrivals):
        """Submit requests (built by request()) at arrival offsets in seconds
        from a native loop (mublis_server_replay); returns (tickets, largest
        submission delay in ms).  Raises on a rejected request, after waiting
        for the ones already queued."""
        n = len(requests)
        arr = (Request * n)(*requests)
        at = (ctypes.c_double * n)(*[float(x) for x in arrivals])
        tk = (ctypes.c_uint64 * n)()
        done, lag = ctypes.c_int(), ctypes.c_double()
        st = self.lib.mublis_server_replay(self.handle, ctypes.cast(arr, ctypes.c_void_p),
                                           ctypes.cast(at, ctypes.c_void_p), n,
                                           ctypes.cast(tk, ctypes.c_void_p), ctypes.byref(done),
                                           ctypes.byref(lag))
        if st:
            for i in range(done.value):
                self.wait(tk[i])
            check(st, "mublis_server_replay")
        return list(tk), lag.value

    def wait(self, ticket):
        """Block until the request completes; returns (status, latency_ms)."""
        lat = ctypes.c_double()
        st = self.lib.mublis_wait(self.handle, ticket, ctypes.byref(lat))
        return st, lat.value

    def stats(self):
        s = Stats()
        check(self.lib.mublis_server_stats(self.handle, ctypes.byref(s)), "mublis_server_stats")
        return s.as_dict()

    def reset_stats(self):
        check(self.lib.mublis_server_reset_stats(self.handle), "mublis_server_reset_stats")

    def close(self):
        if self.handle:
            check(self.lib.mublis_server_destroy(self.handle), "mublis_server_destroy")
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def plain_call(lib, r):
    """Run one serving request (a dict of Server.submit keyword arguments
    plus "op") as the plain synchronous cblas_* call on the same device
    memory -- the serial path the serving loop is measured and checked
    against."""
    a, b, c = (ctypes.c_void_p(r[x].ptr) if r.get(x) is not None else None for x in "abc")
    la, lb, lc = (int(r[x].ld) if r.get(x) is not None else 0 for x in "abc")
    op = r["op"]
    col, notr, tr, lo, up = 102, 111, 112, 122, 121
    uplo = lo if r.get("uplo", LOWER) == LOWER else up
    ta = tr if r.get("trans_a", NO_TRANS) else notr
    tb = tr if r.get("trans_b", NO_TRANS) else notr
    if op == SGEMM:
        lib.cblas_sgemm(col, ta, tb, r["m"], r["n"], r["k"], r["alpha"], a, la, b, lb,
                        r["beta"], c, lc)
    elif op == DGEMM:
        lib.cblas_dgemm(col, ta, tb, r["m"], r["n"], r["k"], r["alpha"], a, la, b, lb,
                        r["beta"], c, lc)
    elif op == SSYRK:
        lib.cblas_ssyrk(col, uplo, ta, r["n"], r["k"], r["alpha"], a, la, r["beta"], c, lc)
    elif op == SSYR2K:
        lib.cblas_ssyr2k(col, uplo, ta, r["n"], r["k"], r["alpha"], a, la, b, lb, r["beta"],
                         c, lc)
    else:
        side = 142 if r.get("side", LEFT) == RIGHT else 141
        diag = 132 if r.get("diag", NON_UNIT) == UNIT else 131
        f = lib.cblas_strsm if op == STRSM else lib.cblas_dtrsm
        f(col, side, uplo, ta, diag, r["m"], r["n"], r["alpha"], a, la, b, lb)
