"""Autotune the compaction for this GPU and write the tuning table.

    python -m gpu_compressstore2_amd.tune [--out FILE] [--reps R]

For every element width (1, 2, 4, 8 bytes), size (2^16, 2^20, 2^22, 2^24,
2^26, 2^28 elements) and Bernoulli selectivity (0.1, 1, 3, 10, 25, 50, 75, 90,
99 %)
csx_autotune times
every candidate configuration on device-generated inputs and records the
fastest; the table is written as text (csx_tune_save).  The package loads
gpu_compressstore2_amd/tuned_mi355x.txt, when present, on first use, so
compact_auto and csx_compact_auto pick from it.
"""
import argparse
import json
import os
import sys

import gpu_compressstore2_amd as csx

WIDTHS = (1, 2, 4, 8)
LOG2N = (16, 20, 22, 24, 26, 28)
SELS = (0.001, 0.01, 0.03, 0.1, 0.25, 0.5, 0.75, 0.9, 0.99)
PIPELINES = {0: "fused", 1: "three_phase", 2: "two_pass", 3: "sparse", 4: "single"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_mi355x.txt"))
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()
    csx.tune_clear()
    rows = []
    for w in WIDTHS:
        for l in LOG2N:
            for p in SELS:
                cfg, ms = csx.autotune(w, 1 << l, p, reps=args.reps)
                # the default configuration on the same inputs, for the record
                dflt = csx.tune_lookup(w, 1 << l, p)
                row = {"w": w, "log2n": l, "sel": p, "best": {**cfg.as_dict(), "pipeline": PIPELINES[cfg.pipeline]}, "ms": round(ms, 5),
                       "GB/s": round(((1 << l) * (w + 0.125 + p * w)) / (ms * 1e-3) / 1e9, 1)}
                assert dflt.as_dict() == cfg.as_dict()
                rows.append(row)
                print(json.dumps(row), flush=True)
    csx.tune_save(args.out)
    print(f"wrote {args.out}", file=sys.stderr)


if __name__ == "__main__":
    main()
