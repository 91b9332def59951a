"""The configurations' own shapes at full size, through the C ABI.

* 4096 u32 columns with lengths log-uniform in [2^10, 2^24] (about 7.1e9
  elements) and per-column Bernoulli selectivity uniform in [0.01, 0.99],
  compacted in one batched call: every column's count against its mask's
  popcount and its contents against torch.masked_select.
* 2^28 four-byte elements with Bernoulli masks at 10 % and 90 % and a mask of
  alternating runs with geometric lengths of mean 256, at chunk lengths 256,
  1024 and 4096, through every writeout variant (popcount, scan and writeout
  called phase by phase) and every pipeline: the data are the indices
  0..n-1, so the output must be exactly the selected positions
  (torch.nonzero of the unpacked mask).

Inputs come from the device generators, bit-exact with the oracle's
(tests/test_gpu_generators.py)."""
import numpy as np
import pytest
import torch

import gpu_compressstore2_amd as csx

pytestmark = pytest.mark.gpu


def unpack_msb_first(mask, n):
    sh = torch.arange(7, -1, -1, dtype=torch.uint8, device=mask.device)
    return ((mask.view(-1, 1) >> sh) & 1).view(-1)[:n].bool()


def test_config_batch_4096_log_uniform_columns():
    seed, nseg = 2024, 4096
    rng = np.random.default_rng(seed)
    lengths = np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 24), nseg)).astype(np.int64)
    sels = rng.uniform(0.01, 0.99, nseg)
    starts, mstarts = np.zeros(nseg, np.int64), np.zeros(nseg, np.int64)
    pos = mpos = 0
    for i, n in enumerate(lengths):
        starts[i], mstarts[i] = pos, mpos
        pos += (int(n) + 63) // 64 * 64
        mpos += ((int(n) + 7) // 8 + 255) // 256 * 256
    total = int(lengths.sum())
    print(f"batch config: {nseg} columns, {total} elements", flush=True)
    data = csx.gen_values(pos, 4, csx.VALUES_INTEGER, seed=seed)
    out = torch.full((pos,), -1, dtype=torch.int32, device="cuda")
    masks = torch.empty(mpos, dtype=torch.uint8, device="cuda")
    ins, ms, outs = [], [], []
    for i in range(nseg):
        n = int(lengths[i])
        mv = masks[mstarts[i]: mstarts[i] + (n + 7) // 8]
        csx.gen_mask_bernoulli(n, float(sels[i]), seed=seed + i, out=mv)
        ins.append(data[starts[i]: starts[i] + n])
        ms.append(mv)
        outs.append(out[starts[i]: starts[i] + n])
    table = csx.segment_table(ins, ms, outs, lengths)
    counts = csx.compact_batch(table, 4, total)
    torch.cuda.synchronize()
    got_counts = counts.cpu().numpy()
    for i in range(nseg):
        n = int(lengths[i])
        bits = unpack_msb_first(ms[i], n)
        expect = torch.masked_select(ins[i], bits)
        k = int(got_counts[i])
        assert k == int(bits.sum().item()) == expect.numel(), (i, n, k)
        assert torch.equal(outs[i][:k], expect), (i, n)
        if k < n:
            assert int(outs[i][k:].ne(-1).sum().item()) == 0, (i, "written past the count")
    del data, out, masks, table, counts
    torch.cuda.empty_cache()


VARIANTS = {"staged": csx.WRITEOUT_STAGED, "direct": csx.WRITEOUT_DIRECT, "thread_per_chunk": csx.WRITEOUT_THREAD_PER_CHUNK,
            "sparse": csx.WRITEOUT_SPARSE}
PIPELINES = {"three_phase": csx.PIPELINE_THREE_PHASE, "two_pass": csx.PIPELINE_TWO_PASS, "fused": csx.PIPELINE_FUSED,
             "sparse": csx.PIPELINE_SPARSE}


@pytest.mark.parametrize("shape", ["bernoulli_0.1", "bernoulli_0.9", "runs_256"])
def test_config_full_size_2p28_every_variant(shape):
    n = 1 << 28
    mask = csx.gen_mask_bernoulli(n, float(shape.split("_")[1]), seed=7) if shape.startswith("bernoulli") else csx.gen_mask_runs(n, 256, seed=7)
    expect = torch.nonzero(unpack_msb_first(mask, n)).flatten().to(torch.int32)
    k_expect = expect.numel()
    data = torch.arange(n, dtype=torch.int32, device="cuda")
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    for chunk in (256, 1024, 4096):
        counts = csx.popc(mask, n, chunk)
        offsets, total = csx.scan(counts)
        torch.cuda.synchronize()
        assert int(total.item()) == k_expect, (shape, chunk)
        for vname, v in VARIANTS.items():
            print(f"2^28 {shape} chunk={chunk} writeout={vname}", flush=True)
            out.fill_(-1)
            csx.writeout(data, mask, offsets, n, chunk, variant=v, counts=counts, out=out)
            torch.cuda.synchronize()
            assert torch.equal(out[:k_expect], expect), (shape, chunk, vname)
            assert int(out[k_expect:].ne(-1).sum().item()) == 0, (shape, chunk, vname)
        ws = csx.alloc_workspace(csx.workspace_size(n, 4, chunk))
        for pname, pl in PIPELINES.items():
            print(f"2^28 {shape} chunk={chunk} pipeline={pname}", flush=True)
            out.fill_(-1)
            _, k = csx.compact(data, mask, n=n, chunk_len=chunk, pipeline=pl, out=out, workspace=ws)
            assert k == k_expect, (shape, chunk, pname)
            assert torch.equal(out[:k], expect), (shape, chunk, pname)
            assert int(out[k:].ne(-1).sum().item()) == 0, (shape, chunk, pname)
        del ws, counts, offsets
    # the tuned entry point at the mask's own selectivity
    _, k = csx.compact_auto(data, mask, k_expect / n, n=n, out=out)
    assert k == k_expect and torch.equal(out[:k], expect)
    del data, out, mask, expect
    torch.cuda.empty_cache()
