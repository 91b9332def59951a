"""Batched compaction of many columns (csx_compact_batch): every column
against the oracle, and graph capture with replays on new inputs."""
import numpy as np
import pytest
import torch

import gpu_compressstore2_amd as csx
import oracle

pytestmark = pytest.mark.gpu

NP = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
VIEW = {1: np.uint8, 2: np.int16, 4: np.int32, 8: np.int64}
TILE = {1: 32768, 2: 16384, 4: 8192, 8: 4096}


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(VIEW[a.dtype.itemsize])).cuda()


def build(O, w, lengths, seed, order=oracle.MSB_FIRST):
    rng = np.random.default_rng(seed)
    cols = []
    for i, n in enumerate(lengths):
        data = O.fill_values(n, NP[w], seed=seed * 1000 + i)
        p = float(rng.uniform(0.01, 0.99))
        mask = O.mask_bernoulli(n, p, seed=seed * 1000 + i + 7, bit_order=order)
        expect, k = O.compact(data, mask, n, order)
        cols.append((data, mask, expect, k))
    ins = [dev(c[0]) if c[0].size else torch.zeros(1, dtype=dev(np.zeros(1, NP[w])).dtype, device="cuda") for c in cols]
    masks = [dev(c[1]) if c[1].size else torch.zeros(1, dtype=torch.uint8, device="cuda") for c in cols]
    outs = [torch.full((max(n, 1) + 4,), -1, dtype=ins[0].dtype, device="cuda") for n in lengths]
    return cols, ins, masks, outs


def check(cols, outs, counts, w):
    counts = counts.cpu().numpy()
    for i, (data, mask, expect, k) in enumerate(cols):
        assert counts[i] == k, (i, counts[i], k)
        got = outs[i].cpu().numpy().view(NP[w])
        assert (got[:k] == expect).all(), i
        assert (got[k:].view(np.uint8) == 0xFF).all(), i  # nothing written past the count


@pytest.mark.parametrize("order", [oracle.MSB_FIRST, oracle.LSB_FIRST])
@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_batch_parity(O, w, order):
    T = TILE[w]
    lengths = [0, 1, 7, 8, 100, T - 1, T, T + 1, 3 * T + 5, 1000, 0, 65536, 250003, 31, 2 * T]
    print(f"batch w={w} order={order} nseg={len(lengths)}", flush=True)
    cols, ins, masks, outs = build(O, w, lengths, seed=w + 10 * order, order=order)
    table = csx.segment_table(ins, masks, outs, lengths)
    counts = csx.compact_batch(table, w, sum(lengths), bit_order=order)
    torch.cuda.synchronize()
    check(cols, outs, counts, w)


@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_batch_every_tile_size_and_count_grid(O, w):
    """Every tile size of the batched form (2, 4 and 8 vectors per thread: for
    2- to 8-byte elements 2 to 8 tiles share one wave load in the count
    kernel) and count grids from one workgroup up, on columns shorter and
    longer than a tile, with empty ones between."""
    T = TILE[w] // 4  # the 2-vector tile
    lengths = [0, 1, T - 1, T, T + 1, 0, 0, 5, 2 * T + 3, 17, 9 * T + 1, 0, 3, 4 * T, 100003]
    for order in (oracle.MSB_FIRST, oracle.LSB_FIRST):
        cols, ins, masks, outs = build(O, w, lengths, seed=40 + w + order, order=order)
        table = csx.segment_table(ins, masks, outs, lengths)
        total = sum(lengths)
        ws = csx.alloc_workspace(csx.batch_workspace_size(len(lengths), total, w))
        counts = torch.zeros(len(lengths), dtype=torch.int64, device="cuda")
        for vpt in (2, 4, 8):
            for cb, block in ((1, 256), (7, 128), (64, 512), (4096, 256), (4096, 128), (4096, 512)):
                print(f"batch w={w} order={order} vpt={vpt} count_blocks={cb} block={block}", flush=True)
                for o in outs:
                    o.fill_(-1)
                csx.compact_batch_config(table, w, total, csx.Config(csx.PIPELINE_THREE_PHASE, 0, vpt, cb, block), counts=counts, workspace=ws,
                                         bit_order=order)
                torch.cuda.synchronize()
                check(cols, outs, counts, w)


@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_batch_many_short_columns(O, w):
    """Hundreds of columns shorter than a tile, a third of them empty, then a
    long one: a wave load of the count kernel holds tiles of several columns
    (up to 16 with 2-vector tiles) and its cursor walks them and the empty
    ones between within one load; each tile's record is checked through the
    write kernels' output."""
    rng = np.random.default_rng(300 + w)
    lengths = [0 if rng.random() < 0.33 else int(rng.integers(1, 300)) for _ in range(400)] + [3 * TILE[w] + 11]
    for order in (oracle.MSB_FIRST, oracle.LSB_FIRST):
        cols, ins, masks, outs = build(O, w, lengths, seed=60 + w + order, order=order)
        table = csx.segment_table(ins, masks, outs, lengths)
        total = sum(lengths)
        ws = csx.alloc_workspace(csx.batch_workspace_size(len(lengths), total, w))
        counts = torch.zeros(len(lengths), dtype=torch.int64, device="cuda")
        for vpt, block in ((2, 128), (2, 256), (4, 512), (8, 512)):
            for cb in (1, 3, 4096):
                print(f"batch short columns w={w} order={order} vpt={vpt} block={block} count_blocks={cb}", flush=True)
                for o in outs:
                    o.fill_(-1)
                csx.compact_batch_config(table, w, total, csx.Config(csx.PIPELINE_THREE_PHASE, 0, vpt, cb, block), counts=counts, workspace=ws,
                                         bit_order=order)
                torch.cuda.synchronize()
                check(cols, outs, counts, w)


@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_batch_unaligned_segments(O, w):
    """Columns sliced at offsets that break 16-byte alignment: input and
    output off by whole elements, masks off by 1..15 bytes, every third
    column aligned, so aligned and unaligned columns share the packed wave
    loads of the count kernel (2- and 4-vector tiles) and the write kernel
    runs its unaligned tile body on tiles after a column's first."""
    T2 = TILE[w] // 4  # elements of the 2-vector tile
    rng = np.random.default_rng(70 + w)
    lengths = [int(x) for x in rng.integers(1, 3 * T2, 40)] + [T2, 2 * T2 + 1, 5, 9 * T2 + 3]
    per16 = 16 // w
    for order in (oracle.MSB_FIRST, oracle.LSB_FIRST):
        cols, ins, masks, outs = [], [], [], []
        for i, n in enumerate(lengths):
            aligned = i % 3 == 0
            eo = 0 if aligned or per16 == 1 else int(rng.integers(1, per16))
            mo = 0 if aligned else int(rng.integers(1, 16))
            oo = 0 if aligned or per16 == 1 else int(rng.integers(1, per16))
            data = O.fill_values(n, NP[w], seed=900 + 31 * i + w)
            mask = O.mask_bernoulli(n, float(rng.uniform(0.05, 0.95)), seed=901 + 31 * i + w, bit_order=order)
            expect, k = O.compact(data, mask, n, order)
            cols.append((data, mask, expect, k))
            ibuf = torch.zeros(n + per16 + 1, dtype=dev(data[:1]).dtype, device="cuda")
            ibuf[eo:eo + n] = dev(data)
            mbuf = torch.zeros(mask.size + 16, dtype=torch.uint8, device="cuda")
            mbuf[mo:mo + mask.size] = dev(mask)
            obuf = torch.full((n + per16 + 4,), -1, dtype=ibuf.dtype, device="cuda")
            ins.append(ibuf[eo:eo + n])
            masks.append(mbuf[mo:mo + mask.size])
            outs.append(obuf[oo:])
        table = csx.segment_table(ins, masks, outs, lengths)
        total = sum(lengths)
        ws = csx.alloc_workspace(csx.batch_workspace_size(len(lengths), total, w))
        counts = torch.zeros(len(lengths), dtype=torch.int64, device="cuda")
        for vpt in (2, 4, 8):
            for cb in (7, 4096):
                print(f"batch unaligned w={w} order={order} vpt={vpt} count_blocks={cb}", flush=True)
                for o in outs:
                    o.fill_(-1)
                csx.compact_batch_config(table, w, total, csx.Config(csx.PIPELINE_THREE_PHASE, 0, vpt, cb), counts=counts, workspace=ws,
                                         bit_order=order)
                torch.cuda.synchronize()
                check(cols, outs, counts, w)


@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_batch_sparse_tiles(O, w):
    """Columns whose tiles hold at most 1/64 selected elements go to the
    sparse kernel: half the columns at 0.1-1.5 % selectivity, some with tiles
    at exactly that threshold (a periodic mask, one element in 64), among
    dense columns, at every tile size; every column against the oracle."""
    rng = np.random.default_rng(80 + w)
    T2 = TILE[w] // 4
    lengths = [int(x) for x in rng.integers(1, 12 * T2, 36)]
    for order in (oracle.MSB_FIRST, oracle.LSB_FIRST):
        cols, ins, masks, outs = [], [], [], []
        for i, n in enumerate(lengths):
            data = O.fill_values(n, NP[w], seed=300 + 7 * i + w)
            if i % 6 == 5:
                mask = O.mask_periodic(n, 0x80000000, order)[: (n + 7) // 8]  # 1 in 32 ...
                mask = mask & np.tile(np.array([0xFF, 0xFF, 0xFF, 0xFF, 0, 0, 0, 0], np.uint8), mask.size // 8 + 1)[: mask.size]  # ... in every other word: 1 in 64
            elif i % 2:
                mask = O.mask_bernoulli(n, float(rng.uniform(0.001, 0.015)), seed=301 + 7 * i + w, bit_order=order)
            else:
                mask = O.mask_bernoulli(n, float(rng.uniform(0.05, 0.95)), seed=301 + 7 * i + w, bit_order=order)
            expect, k = O.compact(data, mask, n, order)
            cols.append((data, mask, expect, k))
            ins.append(dev(data))
            masks.append(dev(mask))
            outs.append(torch.full((n + 4,), -1, dtype=ins[-1].dtype, device="cuda"))
        table = csx.segment_table(ins, masks, outs, lengths)
        total = sum(lengths)
        ws = csx.alloc_workspace(csx.batch_workspace_size(len(lengths), total, w))
        counts = torch.zeros(len(lengths), dtype=torch.int64, device="cuda")
        for vpt in (2, 4, 8):
            for cb, block in ((3, 128), (4096, 256), (4096, 512)):
                print(f"batch sparse tiles w={w} order={order} vpt={vpt} count_blocks={cb} block={block}", flush=True)
                for o in outs:
                    o.fill_(-1)
                csx.compact_batch_config(table, w, total, csx.Config(csx.PIPELINE_THREE_PHASE, 0, vpt, cb, block), counts=counts, workspace=ws,
                                         bit_order=order)
                torch.cuda.synchronize()
                check(cols, outs, counts, w)


def test_batch_many_log_uniform_columns(O):
    rng = np.random.default_rng(4)
    lengths = [int(x) for x in np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 16), 300))]
    cols, ins, masks, outs = build(O, 4, lengths, seed=3)
    table = csx.segment_table(ins, masks, outs, lengths)
    counts = csx.compact_batch(table, 4, sum(lengths))
    torch.cuda.synchronize()
    check(cols, outs, counts, 4)


def test_batch_bound_too_small_sets_error(O):
    lengths = [50000, 50000]
    cols, ins, masks, outs = build(O, 4, lengths, seed=5)
    table = csx.segment_table(ins, masks, outs, lengths)
    ws = csx.alloc_workspace(csx.batch_workspace_size(2, 10, 4))
    csx.compact_batch(table, 4, 10, workspace=ws)
    assert csx.workspace_error(ws) == 2


@pytest.mark.parametrize("w", [1, 4, 8])
def test_batch_loose_bound_dirty_workspace(O, w):
    """A bound far above the columns' total and a workspace full of 0xFF: the
    surplus tiles of the bound find records and counts nobody wrote and must
    leave before using them.  A second, smaller batch in the same workspace
    then meets the first batch's stale records past its own total."""
    rng = np.random.default_rng(90 + w)
    T = TILE[w]
    lengths = [int(x) for x in rng.integers(0, 3 * T, 30)] + [5 * T + 3]
    small = lengths[::3]
    bound = 4 * sum(lengths)
    ws = csx.alloc_workspace(csx.batch_workspace_size(len(lengths), bound, w))
    ws.fill_(0xFF)
    for lens, seed in ((lengths, 21), (small, 22)):
        cols, ins, masks, outs = build(O, w, lens, seed=seed + w)
        table = csx.segment_table(ins, masks, outs, lens)
        counts = torch.zeros(len(lens), dtype=torch.int64, device="cuda")
        for vpt, block in ((2, 128), (8, 512)):
            print(f"batch loose bound w={w} nseg={len(lens)} vpt={vpt} block={block}", flush=True)
            for o in outs:
                o.fill_(-1)
            csx.compact_batch_config(table, w, bound, csx.Config(csx.PIPELINE_THREE_PHASE, 0, vpt, 64, block), counts=counts, workspace=ws)
            torch.cuda.synchronize()
            check(cols, outs, counts, w)
            assert csx.workspace_error(ws) == 0


def test_batch_graph_capture_replays(O):
    lengths = [5000, 70000, 123, 9000, 0, 40000]
    cols, ins, masks, outs = build(O, 4, lengths, seed=8)
    table = csx.segment_table(ins, masks, outs, lengths)
    counts = torch.zeros(len(lengths), dtype=torch.int64, device="cuda")
    ws = csx.alloc_workspace(csx.batch_workspace_size(len(lengths), sum(lengths), 4))
    csx.compact_batch(table, 4, sum(lengths), counts=counts, workspace=ws)  # eager, warms up
    torch.cuda.synchronize()
    check(cols, outs, counts, 4)
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            csx.compact_batch(table, 4, sum(lengths), counts=counts, workspace=ws, stream=s)
    torch.cuda.synchronize()
    for rep in range(4):
        # new inputs in the same buffers, outputs reset, then replay
        new = []
        for i, n in enumerate(lengths):
            data = O.fill_values(n, np.uint32, seed=500 + 10 * rep + i)
            mask = O.mask_bernoulli(n, 0.2 + 0.15 * rep, seed=600 + 10 * rep + i)
            expect, k = O.compact(data, mask, n)
            if n:
                ins[i].copy_(dev(data))
                masks[i].copy_(dev(mask))
            outs[i].fill_(-1)
            new.append((data, mask, expect, k))
        counts.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        check(new, outs, counts, 4)
