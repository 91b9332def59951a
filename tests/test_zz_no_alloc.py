"""csx_compact_noalloc (the library's static device scratch, no workspace
argument) against csx_compact_config (a caller-owned workspace) through the C
ABI: bit-exact output and the same count, for every pipeline, on two calls in
a row whose second input is smaller -- the scratch keeps the first call's
contents, and nothing of them may leak into the second."""
import numpy as np
import pytest
import torch

import gpu_compressstore2_amd as csx

pytestmark = pytest.mark.gpu

# (pipeline, chunk, vpt) of csx_compact_config; every form the entry point runs
CONFIGS = [
    (csx.PIPELINE_THREE_PHASE, 1024, 2),
    (csx.PIPELINE_THREE_PHASE, 256, 8),
    (csx.PIPELINE_TWO_PASS, 1024, 8),
    (csx.PIPELINE_FUSED, 1024, 8),
    (csx.PIPELINE_SPARSE, 4096, 8),
]


def inputs(n, w, p, seed):
    data = csx.gen_values(n, w, seed=seed)
    mask = csx.gen_mask_bernoulli(n, p, seed=seed + 1)
    return data, mask


@pytest.mark.parametrize("w", [1, 4, 8])
@pytest.mark.parametrize("pipeline,chunk,vpt", CONFIGS)
def test_noalloc_equals_workspace_call(w, pipeline, chunk, vpt):
    cfg = csx.Config(pipeline, chunk, vpt, 0)
    # first call large, second smaller, with a different selectivity
    for n, p, seed in ((3_000_017, 0.5, 11), (100_003, 0.03, 23)):
        print(f"noalloc w={w} pipeline={pipeline} chunk={chunk} vpt={vpt} n={n} p={p}", flush=True)
        data, mask = inputs(n, w, p, seed)
        ref, k_ref = csx.compact_config(data, mask, cfg, n=n)
        out = torch.full_like(data, -1)
        count = torch.full((1,), -1, dtype=torch.int64, device="cuda")
        got, k = csx.compact_noalloc(data, mask, cfg, n=n, out=out, count=count)
        torch.cuda.synchronize()
        assert k == k_ref, (n, k, k_ref)
        assert int(count.item()) == k_ref
        assert torch.equal(got[:k], ref[:k_ref])
        # nothing written at or past the count
        assert bool((got[k:] == -1).all())


def test_noalloc_one_tile_and_empty():
    cfg = csx.Config(csx.PIPELINE_THREE_PHASE, 1024, 8, 0)
    for n in (0, 1, 8192, 77):
        data, mask = inputs(max(n, 1), 4, 0.4, 5 + n)
        ref, k_ref = csx.compact_config(data, mask, cfg, n=n)
        got, k = csx.compact_noalloc(data, mask, cfg, n=n)
        assert k == k_ref
        assert torch.equal(got[:k], ref[:k_ref])


def test_noalloc_rejects_oversized_request():
    cfg = csx.Config(csx.PIPELINE_THREE_PHASE, 32, 8, 0)
    need = csx.workspace_size(1 << 30, 4, 32)
    assert need > csx.NOALLOC_SCRATCH_BYTES
    data = torch.empty(1, dtype=torch.int32, device="cuda")
    mask = torch.empty(1, dtype=torch.uint8, device="cuda")
    with pytest.raises(csx.CsxError) as e:
        csx.compact_noalloc(data, mask, cfg, n=1 << 30)
    assert e.value.status == 4
