"""Write tests/golden/ref_vectors.json from the reference's own CPU code.

The vectors come from oracle/_ref/libcsx_ref.so, which `make -C oracle ref`
compiles from RememberOfLife/gpu_compressstore2's src/fast_prng.cuh and
src/data_generator.cuh where they lie.  Only inputs and outputs are stored
(hex strings); no reference source enters this repository.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.json")


def hx(a):
    return np.ascontiguousarray(a).tobytes().hex()


def main():
    R = oracle.Reference()
    v = {"source": "RememberOfLife/gpu_compressstore2 src/fast_prng.cuh, src/data_generator.cuh (compiled, oracle/_ref)", "cases": []}
    cases = v["cases"]
    for seed in (42, 0, 7, 123456789):
        cases.append({"fn": "prng_seq", "seed": seed, "n": 64, "out": hx(R.prng_seq(seed, 64))})
    for nb, p in ((64, 0.5), (64, 0.1), (37, 0.9), (16, 0.0), (16, 1.0)):
        m, c = R.mask_uniform(nb, p)
        cases.append({"fn": "mask_uniform", "nbytes": nb, "p": p, "out": hx(m), "count": c})
    m, c = R.mask_zipf(256)
    cases.append({"fn": "mask_zipf", "nbytes": 256, "out": hx(m), "count": c})
    for ss in (0.05, 0.2):
        cases.append({"fn": "mask_burst", "nbytes": 128, "segment_sizer": ss, "out": hx(R.mask_burst(128, ss))})
    for sp in (3, -5, 1, 0):
        cases.append({"fn": "mask_offset", "nbytes": 64, "spacing": sp, "out": hx(R.mask_offset(64, sp))})
    for pat, L in ((0xB5000000, 8), (0xDEADBEEF, 32), (0x80000000, 1), (0x6C000000, 7)):
        m, c = R.mask_pattern(64, pat, L)
        cases.append({"fn": "mask_pattern", "nbytes": 64, "pattern": pat, "length": L, "out": hx(m), "count": c})
    for s, cc, t in ((0.5, 1, 1024), (0.3, 4, 2048), (0.75, 3, 800)):
        m, c = R.mask_clustering(s, cc, t)
        cases.append({"fn": "mask_clustering", "selectivity": s, "clusters": cc, "total": t, "out": hx(m), "count": c})
    # generate_validation on reference-style inputs: data bytes = uniform mask at p=0.5 (src/main.cu:76)
    for dt, n, p in ((np.uint32, 1000, 0.5), (np.uint64, 333, 0.25), (np.uint8, 509, 0.7), (np.uint16, 64, 0.05)):
        itemsize = np.dtype(dt).itemsize
        data_bytes, _ = R.mask_uniform(n * itemsize, 0.5)
        data = data_bytes.view(dt)
        mask, _ = R.mask_uniform((n + 7) // 8, p)
        out, cnt = R.generate_validation(data, mask)
        cases.append({"fn": "generate_validation", "dtype": np.dtype(dt).name, "n": n, "data": hx(data), "mask": hx(mask),
                      "out": hx(out), "count": cnt})
    with open(OUT, "w") as f:
        json.dump(v, f, indent=1)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
