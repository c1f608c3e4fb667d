"""Write the golden L3 vectors from the reference's own C build.

Run in the development container after `make -C oracle` has compiled
oracle/_ref/ from the reference sources.  For each case of the reference's
correctness harness (192 cases, seed 314159, a fresh generator per case as in
test/correctness/check_cblas.py:157) and for the larger ragged cases, the
inputs and the outputs of both reference configs (linux_avx2 and reference)
are stored, so the tests never need the reference at run time.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import l3cases as L  # noqa: E402


def run(lib, c, prob):
    out = np.array(prob["c0"], copy=True, order=c["order"])
    lib.call(c["dtype"], c["routine"], *prob["args"](out))
    return out


def main():
    avx2 = L.Cblas(L.REF_AVX2_LIB)
    refc = L.Cblas(L.REF_REF_LIB)
    for name, cases, seeder in (
            ("l3_harness", L.all_cases(), lambda i: L.SEED),
            ("l3_ragged", L.ragged_cases(), lambda i: L.SEED + 1 + i)):
        arrays, meta = {}, []
        for i, c in enumerate(cases):
            prob = L.build_problem(c, np.random.default_rng(seeder(i)))
            arrays[f"{i}_a"] = prob["inputs"]["a"]
            if prob["inputs"].get("b") is not None:
                arrays[f"{i}_b"] = prob["inputs"]["b"]
            arrays[f"{i}_c0"] = prob["c0"]
            arrays[f"{i}_avx2"] = run(avx2, c, prob)
            arrays[f"{i}_ref"] = run(refc, c, prob)
            m = {k: (v.name if isinstance(v, np.dtype) else v) for k, v in c.items()}
            meta.append(m)
        arrays["meta"] = np.array(json.dumps(meta))
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
        print(name, len(cases), "cases")


if __name__ == "__main__":
    main()
