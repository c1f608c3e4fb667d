"""The Python ctypes mirrors of the public structs (mublis_amd.SbatchOp,
mublis_amd.serving.ServerConfig / Request / Stats) against the C layout of
include/mublis_gfx950.h: tests/native/abi_layout.c, built by the host C
compiler, prints every struct's size and each field's offset (no GPU)."""
import ctypes
import os
import shutil
import subprocess

import pytest

import l3cases as L
import mublis_amd as M
from mublis_amd import serving as S

HERE = os.path.dirname(os.path.abspath(__file__))
MIRRORS = {"mublis_sbatch_op_t": M.SbatchOp, "mublis_server_config_t": S.ServerConfig,
           "mublis_request_t": S.Request, "mublis_server_stats_t": S.Stats}


def c_layout(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no host C compiler")
    exe = tmp_path / "abi_layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(L.ROOT, "include"),
                    os.path.join(HERE, "native", "abi_layout.c"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    layout = {}
    for line in out.splitlines():
        key, value = line.rsplit(" ", 1)
        layout[key] = int(value)
    return layout


def test_ctypes_mirrors_match_header(tmp_path):
    layout = c_layout(tmp_path)
    for name, cls in MIRRORS.items():
        assert layout[f"{name} sizeof"] == ctypes.sizeof(cls), name
        fields = [f for f, _ in cls._fields_]
        c_fields = sorted((k.split(".", 1)[1] for k in layout if k.startswith(name + ".")),
                          key=lambda f: layout[f"{name}.{f}"])
        assert fields == c_fields, (name, fields, c_fields)
        for f in fields:
            assert getattr(cls, f).offset == layout[f"{name}.{f}"], (name, f)
