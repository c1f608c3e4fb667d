"""MuBLIS L3 engine for MI355X (gfx950): Python host mirror.

The product is the C library ``mublis_amd/lib/libmublis.so`` (CBLAS and
mublis_* entry points, see include/cblas.h and include/mublis.h).  This
module builds it, loads it with ctypes and declares every entry point's
argument types, so callers can pass numpy arrays (host memory, staged by the
library) or torch tensors (device memory, used in place).

Loading fails loudly when the library has not been built: there is no CPU
fallback.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "lib", "libmublis.so")

_VP, _I, _I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
_F, _D = ctypes.c_float, ctypes.c_double

# include/mublis_gfx950.h status codes
OK, ECONTEXT, ENOMEM, EINVAL, EHIP = 0, 1, 2, 3, 4

_lib = None

# include/mublis_gfx950.h batched extensions
BATCH_GEMM, BATCH_SYRK, BATCH_SYR2K, BATCH_TRSM = 0, 1, 2, 3


class SbatchOp(ctypes.Structure):
    """mublis_sbatch_op_t: one FP32 GEMM, SYRK, SYR2K or TRSM problem on device
    memory (TRSM: B at c is solved in place; side and diag are its own)."""
    _fields_ = [("op", _I), ("trans_a", _I), ("trans_b", _I), ("uplo", _I),
                ("m", _I), ("n", _I), ("k", _I), ("alpha", _F), ("beta", _F),
                ("a", _VP), ("lda", _I), ("b", _VP), ("ldb", _I), ("c", _VP), ("ldc", _I),
                ("side", _I), ("diag", _I)]


def build(jobs=8):
    """Compile libmublis.so for gfx950 (always recompiles what changed)."""
    subprocess.check_call(["make", "-C", HERE, f"-j{jobs}"])


def _declare(lib):
    for p, ct in (("s", _F), ("d", _D)):
        cb = {
            "gemm": [_I, _I, _I, _I, _I, _I, ct, _VP, _I, _VP, _I, ct, _VP, _I],
            "symm": [_I, _I, _I, _I, _I, ct, _VP, _I, _VP, _I, ct, _VP, _I],
            "syrk": [_I, _I, _I, _I, _I, ct, _VP, _I, ct, _VP, _I],
            "syr2k": [_I, _I, _I, _I, _I, ct, _VP, _I, _VP, _I, ct, _VP, _I],
            "trmm": [_I, _I, _I, _I, _I, _I, _I, ct, _VP, _I, _VP, _I],
            "trsm": [_I, _I, _I, _I, _I, _I, _I, ct, _VP, _I, _VP, _I],
        }
        for r, args in cb.items():
            f = getattr(lib, f"cblas_{p}{r}")
            f.argtypes, f.restype = args, None
        mb = {
            "gemm": [_I, _I, _I, _I, _I, ct, _VP, _I, _I, _VP, _I, _I, ct, _VP, _I, _I],
            "symm": [_I, _I, _I, _I, ct, _VP, _I, _I, _VP, _I, _I, ct, _VP, _I, _I],
            "syrk": [_I, _I, _I, _I, ct, _VP, _I, _I, ct, _VP, _I, _I],
            "syr2k": [_I, _I, _I, _I, ct, _VP, _I, _I, _VP, _I, _I, ct, _VP, _I, _I],
            "trmm": [_I, _I, _I, _I, _I, _I, ct, _VP, _I, _I, _VP, _I, _I],
            "trsm": [_I, _I, _I, _I, _I, _I, ct, _VP, _I, _I, _VP, _I, _I],
        }
        for r, args in mb.items():
            f = getattr(lib, f"mublis_{p}{r}")
            f.argtypes, f.restype = args, _I
        f = getattr(lib, f"mublis_l3_{p}driver")
        f.argtypes = [_I, _I, _I, ct, _VP, _I, _I, _VP, _I, _I, ct, _VP, _I, _I, _VP]
        f.restype = _I
        f = getattr(lib, f"mublis_gfx950_{p}gemm_dev")
        f.argtypes = [_I, _I, _I, ct] + [_VP, _I64, _I64] * 4 + [ct, _VP, _I64, _I64,
                                                                  _I, _VP]
        f.restype = _I
        f = getattr(lib, f"mublis_gfx950_{p}trsm_dev")
        f.argtypes = [_I, _I, _I, _I, ct, _VP, _I64, _I64, _VP, _I64, _I64, _VP]
        f.restype = _I
        f = getattr(lib, f"mublis_gfx950_{p}pack_dev")
        f.argtypes = [_I, _I, _VP, _I64, _I64, _I, _I, _I, _I, _I, _VP, _I64, _VP]
        f.restype = _I
        f = getattr(lib, f"mublis_gfx950_{p}trsm_inverted_dev")
        f.argtypes = [_I, _I, _I, _VP, _I64, _I64, _VP, _I64, _I64, _VP]
        f.restype = _I
        f = getattr(lib, f"mublis_gfx950_{p}copy_dev")
        f.argtypes = [_I, _I, _VP, _I64, _I64, _VP, _I64, _I64, _VP]
        f.restype = _I
        f = getattr(lib, f"mublis_gfx950_{p}fill_dev")
        f.argtypes = [_I, _I, ct, _VP, _I64, _I64, _VP]
        f.restype = _I
    lib.mublis_sgrouped_plan_create.argtypes = [_VP, _I, _I, _VP]
    lib.mublis_sgrouped_plan_create.restype = _I
    lib.mublis_grouped_plan_execute.argtypes = [_VP, _VP]
    lib.mublis_grouped_plan_execute.restype = _I
    lib.mublis_grouped_plan_info.argtypes = [_VP, _VP, _VP, _VP, _VP]
    lib.mublis_grouped_plan_info.restype = _I
    lib.mublis_grouped_plan_destroy.argtypes = [_VP]
    lib.mublis_grouped_plan_destroy.restype = _I
    lib.mublis_sgemm_grouped.argtypes = [_VP, _I]
    lib.mublis_sgemm_grouped.restype = _I
    lib.mublis_grouped_autotune.argtypes = [ctypes.c_char_p, _I, _I, _I, _I]
    lib.mublis_grouped_autotune.restype = _I
    lib.mublis_grouped_load_table.argtypes = [ctypes.c_char_p]
    lib.mublis_grouped_load_table.restype = _I
    lib.mublis_sgemm_bf16x3.argtypes = [_I] * 5 + [_F, _VP, _I, _I, _VP, _I, _I, _F, _VP, _I,
                                                  _I]
    lib.mublis_sgemm_bf16x3.restype = _I
    lib.mublis_strsm_refined.argtypes = [_I] * 6 + [_F, _VP, _I, _I, _VP, _I, _I, _I]
    lib.mublis_strsm_refined.restype = _I
    lib.mublis_set_sgemm_mode.argtypes = [_I]
    lib.mublis_set_sgemm_mode.restype = _I
    lib.mublis_get_sgemm_mode.restype = _I
    lib.mublis_rt_init.restype = _I
    lib.mublis_rt_arch.restype = ctypes.c_char_p
    lib.mublis_rt_last_error.restype = ctypes.c_char_p
    lib.mublis_rt_stream.restype = _VP
    lib.mublis_rt_sync.argtypes = [_VP]
    lib.mublis_rt_sync.restype = _I
    lib.mublis_rt_is_device_ptr.argtypes = [_VP]
    lib.mublis_rt_is_device_ptr.restype = _I
    lib.mublis_get_context.argtypes = [_VP]
    lib.mublis_get_context.restype = _I
    _SZ = ctypes.c_size_t
    lib.mublis_sgrouped_ws_bytes.argtypes = [_VP, _I]
    lib.mublis_sgrouped_ws_bytes.restype = _SZ
    lib.mublis_sgrouped_launch_ws.argtypes = [_VP, _I, _VP, _VP, _SZ, _VP]
    lib.mublis_sgrouped_launch_ws.restype = _I
    lib.mublis_gfx950_trsm_ws_bytes.argtypes = [_I]
    lib.mublis_gfx950_trsm_ws_bytes.restype = _SZ
    lib.mublis_gfx950_trsm_pipe_max_n.argtypes = [_I]
    lib.mublis_gfx950_trsm_pipe_max_n.restype = _I
    for name, ct in (("mublis_gfx950_dtrsm_ws_dev", _D), ("mublis_gfx950_strsm_ws_dev", _F)):
        f = getattr(lib, name)
        f.argtypes = [_I, _I, _I, _I, ct, _VP, _I64, _I64, _VP, _I64, _I64, _VP, _SZ, _VP]
        f.restype = _I
    f = lib.mublis_gfx950_dtrsm_noalloc_dev
    f.argtypes = [_I, _I, _I, _I, _D, _VP, _I64, _I64, _VP, _I64, _I64, _VP]
    f.restype = _I
    # serving loop
    for name, args in (("mublis_server_create", [_VP, _VP]),
                       ("mublis_server_destroy", [_VP]),
                       ("mublis_matrix_create", [_VP, _I, _I, _I, _I, _VP]),
                       ("mublis_matrix_write", [_VP, _VP, _VP, _I64]),
                       ("mublis_matrix_read", [_VP, _VP, _VP, _I64]),
                       ("mublis_matrix_destroy", [_VP, _VP]),
                       ("mublis_submit", [_VP, _VP, _VP]),
                       ("mublis_server_replay", [_VP, _VP, _VP, _I, _VP, _VP, _VP]),
                       ("mublis_wait", [_VP, ctypes.c_uint64, _VP]),
                       ("mublis_server_stats", [_VP, _VP]),
                       ("mublis_server_reset_stats", [_VP])):
        f = getattr(lib, name)
        f.argtypes, f.restype = args, _I
    lib.mublis_matrix_device_ptr.argtypes = [_VP]
    lib.mublis_matrix_device_ptr.restype = _VP
    lib.mublis_matrix_ld.argtypes = [_VP]
    lib.mublis_matrix_ld.restype = _I64
    return lib


def load():
    """Load libmublis.so (built in-tree).  Raises if it is missing.

    When torch is importable it is imported first: torch ships its own
    libamdhip64 (SONAME libamdhip64.so.7) and requests it under another
    name, so loading this library first would put two HIP runtimes in one
    process and torch would then see no GPU.  With torch loaded first, this
    library's libamdhip64.so.7 dependency binds to the runtime already
    present and device pointers are shared.
    """
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run mublis_amd.build() (no CPU fallback)")
        _lib = _declare(ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL))
    return _lib


def ptr(x):
    """Address of a numpy array or torch tensor (or None)."""
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return ctypes.c_void_p(x.data_ptr())
    return ctypes.c_void_p(x.ctypes.data)


def check(status, what="mublis"):
    if status != OK:
        msg = load().mublis_rt_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed with status {status}: {msg}")
    return status


class GroupedPlan:
    """A captured grouped launch (mublis_sgrouped_plan_create / execute)."""

    def __init__(self, ops, flags=0):
        self.lib = load()
        self.ops = (SbatchOp * len(ops))(*ops)
        self.handle = ctypes.c_void_p()
        check(self.lib.mublis_sgrouped_plan_create(ctypes.cast(self.ops, _VP), len(ops),
                                                   flags, ctypes.byref(self.handle)),
              "mublis_sgrouped_plan_create")

    def execute(self, stream=None):
        check(self.lib.mublis_grouped_plan_execute(self.handle, stream),
              "mublis_grouped_plan_execute")

    def info(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int()
        h = (ctypes.c_int * 3)()
        check(self.lib.mublis_grouped_plan_info(self.handle, ctypes.byref(f), ctypes.byref(b),
                                                ctypes.byref(n), h))
        return {"flops": f.value, "bytes": b.value, "items": n.value, "tiles": list(h)}

    def close(self):
        if self.handle:
            self.lib.mublis_grouped_plan_destroy(self.handle)
            self.handle = ctypes.c_void_p()
