"""Host-only checks of the serving loop's building blocks (no GPU).

* The budget allocator (mublis_amd/runtime/budget_arena.h) against a
  brute-force model of every 256-byte unit, built with AddressSanitizer and
  UndefinedBehaviorSanitizer: no overlap, best fit never misses a free run
  large enough, coalescing leaves no two adjacent free blocks, double release
  is refused.
* The serving structs of include/mublis_gfx950.h have the layout the Python
  mirror (mublis_amd/serving.py) declares.
"""
import ctypes
import os
import shutil
import subprocess

import pytest

import l3cases as L

HERE = os.path.dirname(os.path.abspath(__file__))


def _gxx():
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    return gxx


def test_budget_arena_under_sanitizers(tmp_path):
    exe = tmp_path / "arena_check"
    subprocess.run([_gxx(), "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all",
                    "-I", os.path.join(L.ROOT, "mublis_amd", "runtime"),
                    os.path.join(HERE, "native", "arena_check.cpp"), "-o", str(exe)],
                   check=True)
    for seed in (1, 2, 3):
        out = subprocess.run([str(exe), str(seed)], capture_output=True, text=True,
                             timeout=300)
        assert out.returncode == 0, out.stderr
        assert f"ok seed {seed}" in out.stdout


def test_serving_struct_layout(tmp_path):
    import mublis_amd.serving as S
    src = tmp_path / "layout.c"
    fields = {"mublis_server_config_t": S.ServerConfig, "mublis_request_t": S.Request,
              "mublis_server_stats_t": S.Stats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mublis_gfx950.h"',
             'int main(void) {']
    for cname, py in fields.items():
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ['  return 0;', '}']
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(L.ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = dict(line.rsplit(" ", 1) for line in
               subprocess.run([str(exe)], capture_output=True, text=True,
                              check=True).stdout.splitlines())
    for cname, py in fields.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)
