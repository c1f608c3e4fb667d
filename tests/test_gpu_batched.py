"""GPU checks of the grouped persistent batched kernel (SURVEY-named batched
workload: uneven FP32 GEMM and SYRK problems in one hipGraph-captured
launch), through the C ABI.

Bars: every problem's result is bit-identical to the corresponding single
cblas_sgemm / cblas_ssyrk call on the same device operands (the grouped
kernel runs the same per-tile MFMA routine and the same K order); forced
tile shapes give identical bits; graph replays accumulate exactly like
repeated calls; a sample is within the normwise FP32 bound of the oracle.
"""
import ctypes
import os

import numpy as np
import pytest

import l3cases as L
import mublis_amd as M
from mublis_amd.workloads import grouped_problems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    lib = M.load()
    assert lib.mublis_rt_init() == 0
    return lib


def single_calls(lib, torch, specs, initial, alpha, beta, times=1):
    outs = []
    for s, c0 in zip(specs, initial):
        C = c0.clone()
        for _ in range(times):
            if s["op"] == "gemm":
                lib.cblas_sgemm(102, 112 if s["ta"] else 111, 112 if s["tb"] else 111, s["m"],
                                s["n"], s["k"], alpha, M.ptr(s["A"]), s["lda"], M.ptr(s["B"]),
                                s["ldb"], beta, M.ptr(C), s["m"])
            else:
                lib.cblas_ssyrk(102, 121 if s["upper"] else 122, 112 if s["trans"] else 111,
                                s["n"], s["k"], alpha, M.ptr(s["A"]), s["lda"], beta, M.ptr(C),
                                s["n"])
        outs.append(C)
    torch.cuda.synchronize()
    return outs


@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (0.75, -0.5), (1.0, 0.0), (0.0, 2.0)])
def test_grouped_equals_single_calls(lib, alpha, beta):
    torch = pytest.importorskip("torch")
    ops, keep, specs = grouped_problems(torch, seed=7, n_gemm=300, n_syrk=80,
                                        syrk_n=(32, 512), alpha=alpha, beta=beta,
                                        vary_flags=True)
    initial = [s["C"].clone() for s in specs]
    want = single_calls(lib, torch, specs, initial, alpha, beta)
    for s, c0 in zip(specs, initial):
        s["C"].copy_(c0)
    plan = M.GroupedPlan(ops)
    try:
        plan.execute()
        torch.cuda.synchronize()
        info = plan.info()
    finally:
        plan.close()
    assert info["items"] > 0 and sum(info["tiles"]) == len(ops)
    bad = [i for i, (s, w) in enumerate(zip(specs, want)) if not torch.equal(s["C"], w)]
    assert not bad, f"{len(bad)} problems differ, first {specs[bad[0]]['op']}"


def test_forced_tiles_and_replay(lib):
    torch = pytest.importorskip("torch")
    ops, keep, specs = grouped_problems(torch, seed=8, n_gemm=120, n_syrk=40,
                                        syrk_n=(32, 300), vary_flags=True)
    initial = [s["C"].clone() for s in specs]
    want = single_calls(lib, torch, specs, initial, 1.0, 1.0, times=2)
    for t in range(3):
        for s, c0 in zip(specs, initial):
            s["C"].copy_(c0)
        plan = M.GroupedPlan(ops, flags=(t + 1) << 4)
        try:
            assert plan.info()["tiles"][t] == len(ops)
            plan.execute()
            torch.cuda.synchronize()
            plan.execute()  # beta = 1: the second replay accumulates again
            torch.cuda.synchronize()
        finally:
            plan.close()
        bad = [i for i, (s, w) in enumerate(zip(specs, want)) if not torch.equal(s["C"], w)]
        assert not bad, f"tile {t}: {len(bad)} problems differ"


def test_grouped_sample_vs_oracle(lib):
    torch = pytest.importorskip("torch")
    orc = L.Cblas(L.ORACLE_LIB, prefix="orc_", returns_status=True)
    ops, keep, specs = grouped_problems(torch, seed=9, n_gemm=40, n_syrk=12,
                                        syrk_n=(32, 256), alpha=0.5, beta=-1.0,
                                        vary_flags=True)
    initial = [s["C"].cpu().numpy().T.copy(order="F") for s in specs]
    assert lib.mublis_sgemm_grouped(ctypes.cast((M.SbatchOp * len(ops))(*ops),
                                                ctypes.c_void_p), len(ops)) == 0
    eps = np.finfo(np.float32).eps
    for s, c0 in zip(specs, initial):
        got = s["C"].cpu().numpy().T
        want = c0.copy(order="F")
        a = np.asfortranarray(s["A"].cpu().numpy().T)
        if s["op"] == "gemm":
            b = np.asfortranarray(s["B"].cpu().numpy().T)
            assert orc.call(np.float32, "gemm", 102, 112 if s["ta"] else 111,
                            112 if s["tb"] else 111, s["m"], s["n"], s["k"], 0.5, a, s["lda"],
                            b, s["ldb"], -1.0, want, s["m"]) == 0
            scale = 0.5 * np.linalg.norm(a) * np.linalg.norm(b) + np.linalg.norm(c0)
        else:
            assert orc.call(np.float32, "syrk", 102, 121 if s["upper"] else 122,
                            112 if s["trans"] else 111, s["n"], s["k"], 0.5, a, s["lda"], -1.0,
                            want, s["n"]) == 0
            scale = 0.5 * np.linalg.norm(a) ** 2 + np.linalg.norm(c0)
        assert np.linalg.norm(got - want) <= 2 * s["k"] * eps * scale


def test_autotuner_writes_and_uses_table(lib, tmp_path):
    torch = pytest.importorskip("torch")
    path = str(tmp_path / "tune.txt").encode()
    assert lib.mublis_grouped_autotune(path, 4, 5, 4, 1) == 0
    lines = [ln for ln in open(path.decode()) if not ln.startswith("#")]
    assert len(lines) == 8
    assert lib.mublis_grouped_load_table(path) == 8
    ops, keep, specs = grouped_problems(torch, seed=10, n_gemm=50, n_syrk=0,
                                        gemm_range=(16, 63))
    plan = M.GroupedPlan(ops)
    try:
        hist = plan.info()["tiles"]
    finally:
        plan.close()
    table = {tuple(map(int, ln.split()[:3])): int(ln.split()[3]) for ln in lines}
    lg = lambda v: min(max(int(np.floor(np.log2(v))), 4), 13)  # noqa: E731
    expect = [0, 0, 0]
    for s in specs:
        expect[table[(lg(s["m"]), lg(s["n"]), lg(s["k"]))]] += 1
    assert hist == expect
    lib.mublis_grouped_load_table(None)
