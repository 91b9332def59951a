"""Full-size parity of the 2^29-element cases: every element width with a
Bernoulli(0.25) mask, a seeded periodic 32-bit pattern and Zipf(s=1.2)
per-4096-block selectivity over {0, 1/64, ..., 1}.  Inputs come from the
device generators (bit-exact with the oracle's sequential generators, which
tests/test_gpu_generators.py checks); results are checked against
torch.masked_select, an independent implementation, bit for bit, and the
counts against the mask's popcount."""
import numpy as np
import pytest
import torch

import gpu_compressstore2_amd as csx

pytestmark = pytest.mark.gpu

N = 1 << 29
TYPES = [  # (elem_size, value kind, torch view for the comparison)
    (1, csx.VALUES_INTEGER, torch.uint8),
    (2, csx.VALUES_INTEGER, torch.int16),
    (4, csx.VALUES_INTEGER, torch.int32),
    (8, csx.VALUES_INTEGER, torch.int64),
    (8, csx.VALUES_FLOAT, torch.int64),  # float64, compared as bits
]
SHIFTS = None


def unpack(mask, n):
    global SHIFTS
    if SHIFTS is None:
        SHIFTS = torch.arange(7, -1, -1, dtype=torch.uint8, device="cuda")
    return ((mask[: (n + 7) // 8].unsqueeze(1) >> SHIFTS) & 1).reshape(-1)[:n].bool()


def masks(O):
    pattern = int(O.prng_seq(29, 1)[0])  # a seeded 32-bit pattern
    yield "bernoulli_0.25", csx.gen_mask_bernoulli(N, 0.25, seed=3), None
    yield "periodic", csx.gen_mask_pattern(N, pattern, 32), pattern
    yield "zipf_blocks", csx.gen_mask_zipf_blocks(N, 4096, 1.2, 65, seed=5), None


@pytest.mark.parametrize("w,kind,view", TYPES, ids=["u8", "u16", "u32", "u64", "f64"])
def test_full_size(O, w, kind, view):
    data = csx.gen_values(N, w, kind, seed=100 + w + kind)
    flat = data.view(view)
    out = torch.empty_like(flat)
    ws = csx.alloc_workspace(csx.auto_workspace_size(N, w))
    for name, mask, pattern in masks(O):
        sel = unpack(mask, N)
        expect = torch.masked_select(flat, sel)
        k_exp = expect.numel()
        print(f"[full-size] w={w} kind={kind} mask={name} k={k_exp}", flush=True)
        _, k = csx.compact(flat, mask, n=N, out=out, workspace=ws)
        assert k == k_exp, name
        assert torch.equal(out[:k], expect), name
        out.fill_(0)
        _, k2 = csx.compact_auto(flat, mask, k_exp / N, n=N, out=out, workspace=ws)
        assert k2 == k_exp and torch.equal(out[:k2], expect), name
        if pattern is not None:
            k3 = csx.pattern_count(N, pattern, 32)
            out.fill_(0)
            _, k4 = csx.pattern_compact(flat, pattern, 32, n=N, out=out)
            assert k3 == k4 == k_exp and torch.equal(out[:k4], expect), name
        del sel, expect
    if w == 4:
        # the same data through the fused predicate path: value > 2^31 as uint32
        mask, cnt = csx.predicate_mask(flat, ">", 2 ** 31, n=N, col_type=2)
        expect = torch.masked_select(flat, unpack(mask, N))
        assert int(cnt.item()) == expect.numel()
        assert bool(((expect.to(torch.int64) & 0xFFFFFFFF) > 2 ** 31).all())
        _, k = csx.compact_predicate(flat, ">", 2 ** 31, n=N, col_type=2, out=out)
        assert k == expect.numel() and torch.equal(out[:k], expect)


def test_beyond_32_bit_index():
    """A u8 column longer than 2^32 elements (a ragged tail byte included):
    offsets, counts and tile indices are 64-bit on every pipeline."""
    n = (1 << 32) + 4099
    data = csx.gen_values(n, 1, csx.VALUES_INTEGER, seed=77)
    mask = csx.gen_mask_bernoulli(n, 0.5, seed=78)
    sel = unpack(mask, n)
    # torch.masked_select overflows past 2^32 elements: select slice by slice
    step = 1 << 30
    expect = torch.cat([torch.masked_select(data[a:a + step], sel[a:a + step]) for a in range(0, n, step)])
    del sel
    k_exp = expect.numel()
    assert k_exp > (1 << 31)
    out = torch.empty_like(data)
    ws = csx.alloc_workspace(csx.auto_workspace_size(n, 1))
    for pipeline in (csx.PIPELINE_THREE_PHASE, csx.PIPELINE_TWO_PASS):
        print(f"[beyond-2^32] n={n} pipeline={pipeline} k={k_exp}", flush=True)
        out[k_exp:k_exp + 4096].fill_(0xA5)  # nothing past the count is written
        _, k = csx.compact(data, mask, n=n, pipeline=pipeline, out=out, workspace=ws)
        assert k == k_exp, pipeline
        assert torch.equal(out[:k], expect), pipeline
        assert bool((out[k:k + 4096] == 0xA5).all()), pipeline
    out.fill_(0)
    _, k = csx.compact_auto(data, mask, 0.5, n=n, out=out, workspace=ws)
    assert k == k_exp and torch.equal(out[:k], expect)
