"""Randomized parity over the configuration space: element width, size
(ragged tails included), selectivity, bit order, pipeline with its chunk
length and tile size, and input/mask/output pointers off 16-byte alignment,
each case checked against the oracle with a canary past the count.  Seeded,
so a failure names a reproducible case."""
import numpy as np
import pytest
import torch

import gpu_compressstore2_amd as csx

pytestmark = pytest.mark.gpu

NP = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
TV = {1: torch.uint8, 2: torch.int16, 4: torch.int32, 8: torch.int64}
VIEW = {1: np.uint8, 2: np.int16, 4: np.int32, 8: np.int64}


def random_config(rng, w):
    pipeline = int(rng.choice([csx.PIPELINE_FUSED, csx.PIPELINE_THREE_PHASE, csx.PIPELINE_TWO_PASS, csx.PIPELINE_SPARSE, csx.PIPELINE_SINGLE]))
    chunk = int(rng.choice([32, 64, 128, 256, 1024, 4096, 8192, 1 << 15]))
    vpt, block = 8, 0
    if pipeline == csx.PIPELINE_THREE_PHASE:
        vpt = int(rng.choice([2, 4, 8, 16] if w >= 2 else [2, 4, 8]))
        block = int(rng.choice([0, 128, 256, 512]))
    return csx.Config(pipeline, chunk, vpt, 0, block)


@pytest.mark.parametrize("block", range(8))
def test_random_configs(O, block):
    rng = np.random.default_rng(1000 + block)
    for case in range(100):
        w = int(rng.choice([1, 2, 4, 8]))
        n = int(rng.integers(0, 1 << int(rng.integers(1, 21))))
        p = float(rng.choice([0.0, 0.001, 0.02, 0.3, 0.5, 0.8, 0.999, 1.0]))
        order = int(rng.integers(0, 2))
        cfg = random_config(rng, w)
        ioff, moff, ooff = (int(x) for x in rng.integers(0, 4, 3))
        print(f"random block={block} case={case} w={w} n={n} p={p} order={order} cfg={cfg.as_dict()} offsets={ioff},{moff},{ooff}", flush=True)
        data = O.fill_values(n + ioff, NP[w], seed=case * 7 + block)
        mask_full = O.mask_bernoulli(n + 8 * moff + 64, p, seed=case * 11 + block, bit_order=order)
        mask = mask_full[moff:]
        expect, ke = O.compact(data[ioff:], mask, n, order)
        d_in = torch.from_numpy(data.view(VIEW[w])).cuda()[ioff:]
        d_mask = torch.from_numpy(mask_full).cuda()[moff:]
        canary_val = 0x5A
        d_outbuf = torch.full(((n + ooff + 16) * w,), canary_val, dtype=torch.uint8, device="cuda").view(TV[w])
        d_out = d_outbuf[ooff:]
        _, k = csx.compact_config(d_in, d_mask, cfg, n=n, bit_order=order, out=d_out)
        assert k == ke, (case, k, ke)
        got = d_out.cpu().numpy().view(NP[w])
        assert (got[:k] == expect).all(), case
        rest = d_outbuf.view(torch.uint8).cpu().numpy()[(ooff + k) * w:]
        assert (rest == canary_val).all(), case
