"""Run the grouped-kernel autotuner on the current GPU and write its table.

    python -m mublis_amd.autotune [--out PATH] [--lo 4] [--hi 10]

The table holds the launch schedule of each tile class ("sched class
variant"), the tile shape of each GEMM bucket ("gemm lm ln lk tile") and of
each SYRK / SYR2K bucket ("syrk ln lk tile"), buckets by floor(log2) of the
sizes; a trailing "~" marks a near-tie that took the size heuristic's
choice.  The library loads it from mublis_amd/tune/grouped_gfx950.txt (or
from $MUBLIS_TUNE_TABLE) when it plans a grouped launch.
"""
import argparse
import os
import time

from . import HERE, check, load


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--out", default=os.path.join(HERE, "tune", "grouped_gfx950.txt"))
    p.add_argument("--lo", type=int, default=4)
    p.add_argument("--hi", type=int, default=10)
    p.add_argument("--per-bucket", type=int, default=16)
    p.add_argument("--reps", type=int, default=3)
    a = p.parse_args()
    lib = load()
    check(lib.mublis_rt_init(), "mublis_rt_init")
    t0 = time.time()
    check(lib.mublis_grouped_autotune(a.out.encode(), a.lo, a.hi, a.per_bucket, a.reps),
          "mublis_grouped_autotune")
    print(f"wrote {a.out} in {time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    main()
