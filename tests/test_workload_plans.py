"""The benchmark workloads' sizes are drawn on the host (workloads.py:
grouped_shapes, serving_plan), so that the CPU baseline runs the same
problems as the GPU.  These pin the draws: counts, ranges, the operation
mix, a few leading values and determinism."""
from collections import Counter

import numpy as np

from mublis_amd import serving as S
from mublis_amd.workloads import _SERVE_POOL, grouped_shapes, serving_plan


def test_grouped_shapes():
    ms, ns, ks, gf, sn, sk, sf = grouped_shapes()
    assert len(ms) == len(ns) == len(ks) == len(gf) == 4096
    assert len(sn) == len(sk) == len(sf) == 1024
    for x in (ms, ns, ks, sk):
        assert x.min() >= 16 and x.max() <= 512
    assert sn.min() >= 32 and sn.max() <= 1024
    assert list(ms[:5]) == [376, 47, 152, 267, 22]
    assert list(sn[:5]) == [150, 68, 90, 247, 194]
    assert set(gf) == {(0, 0)} and set(sf) == {(0, 0)}
    again = grouped_shapes()
    assert all(np.array_equal(a, b) for a, b in zip((ms, ns, ks, sn, sk),
                                                    (again[0], again[1], again[2],
                                                     again[4], again[5])))
    # flags are drawn after the GEMM sizes: those stay the same draws
    fl = grouped_shapes(vary_flags=True)
    assert np.array_equal(fl[0], ms) and np.array_equal(fl[2], ks)
    assert {t for t, _ in fl[3]} == {0, 1}


def test_serving_plan():
    plan = serving_plan()
    assert len(plan) == 2000
    mix = Counter(r["op"] for r in plan)
    assert mix == {S.SGEMM: 1176, S.SSYR2K: 235, S.SSYRK: 198, S.DTRSM: 391}
    total = sum(r["flop"] for r in plan)
    assert abs(total - 3.311837496197e12) < 1e3
    for r in plan:
        for role, dt, rows, cols, off in r["mats"]:
            assert 0 <= off <= _SERVE_POOL - rows * cols
            assert dt in (S.F32, S.F64)
        if r["op"] == S.SGEMM:
            assert [m[0] for m in r["mats"]] == ["a", "b", "c"]
            assert all(64 <= r[x] <= 4096 for x in "mnk")
        elif r["op"] == S.DTRSM:
            assert 256 <= r["m"] <= 4096 and 1 <= r["n"] <= 512
        else:
            assert r["uplo"] in (S.LOWER, S.UPPER)
    assert serving_plan() == plan
