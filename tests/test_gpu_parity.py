"""GPU parity: every phase and pipeline of libcsx.so against the oracle,
bit-exact, on seeded inputs (oracle sizes), plus full-size property checks."""
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
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a.view(VIEW[a.dtype.itemsize])).cuda()


def host(t, dtype):
    return t.cpu().numpy().view(dtype)


def make_mask(O, kind, n, seed, order):
    if kind == "bern50":
        return O.mask_bernoulli(n, 0.5, seed, order)
    if kind == "bern01":
        return O.mask_bernoulli(n, 0.01, seed, order)
    if kind == "bern99":
        return O.mask_bernoulli(n, 0.99, seed, order)
    if kind == "zeros":
        return np.zeros((n + 7) // 8, np.uint8)
    if kind == "ones":
        m = np.full((n + 7) // 8, 0xFF, np.uint8)
        return m
    if kind == "runs":
        return O.mask_geometric_runs(n, 256, seed, order)
    if kind == "zipf":
        return O.mask_zipf_blocks(n, 4096, 1.2, 65, seed, order)
    if kind == "periodic":
        return O.mask_periodic(n, 0x9E3779B9, order)
    raise ValueError(kind)


def dirty_tail(mask, n, order):
    """Set every bit past n in the last byte: the engine must ignore them."""
    m = mask.copy()
    if n % 8:
        r = n % 8
        m[-1] |= np.uint8(0xFF >> r) if order == oracle.MSB_FIRST else np.uint8((0xFF << r) & 0xFF)
    return m


def run_compact(data, mask, n, order, pipeline, chunk=1024, canary=True):
    w = data.dtype.itemsize
    d_in = dev(data)
    d_mask = dev(mask)
    # output with a canary pattern: nothing at or past out[count] may change
    out_h = np.frombuffer(np.full((max(n, 1) + 64) * w, 0xA5, np.uint8).tobytes(), dtype=NP[w]).copy()
    d_out = dev(out_h)
    out, k = csx.compact(d_in, d_mask, n=n, chunk_len=chunk, bit_order=order, pipeline=pipeline, out=d_out)
    torch.cuda.synchronize()
    res = host(out, NP[w])
    if canary:
        assert (res[k:] == out_h[k:]).all(), "write past the selected count"
    return res[:k], k


CASES = [
    (1, "bern50"), (7, "bern50"), (8, "ones"), (9, "bern50"), (100, "bern01"), (1000, "runs"),
    ("T-1", "bern50"), ("T", "ones"), ("T+1", "bern99"), ("3T+17", "bern50"), ("3T+17", "zeros"),
    (200003, "zipf"), (131072, "periodic"), ((1 << 20) + 5, "bern50"), ((1 << 20) + 5, "bern01"),
]


def resolve_n(n, w):
    if isinstance(n, int):
        return n
    T = TILE[w]
    return {"T-1": T - 1, "T": T, "T+1": T + 1, "3T+17": 3 * T + 17}[n]


PIPELINES = [csx.PIPELINE_FUSED, csx.PIPELINE_THREE_PHASE, csx.PIPELINE_TWO_PASS, csx.PIPELINE_SPARSE, csx.PIPELINE_SINGLE]
# the pipelines without a size cap (the single launch stops at 2^24 elements)
PIPELINES_ANY_N = [csx.PIPELINE_FUSED, csx.PIPELINE_THREE_PHASE, csx.PIPELINE_TWO_PASS, csx.PIPELINE_SPARSE]


@pytest.mark.parametrize("pipeline", PIPELINES)
@pytest.mark.parametrize("order", [oracle.MSB_FIRST, oracle.LSB_FIRST])
@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_compact_parity(O, w, order, pipeline):
    for i, (n, kind) in enumerate(CASES):
        n = resolve_n(n, w)
        print(f"compact w={w} order={order} pipeline={pipeline} n={n} mask={kind}", flush=True)
        data = O.fill_values(n, NP[w], seed=1000 + i)
        mask = dirty_tail(make_mask(O, kind, n, 2000 + i, order), n, order)
        expect, ke = O.compact(data, mask, n, order)
        got, k = run_compact(data, mask, n, order, pipeline)
        assert k == ke, (n, kind)
        assert (got == expect).all(), (n, kind)


@pytest.mark.parametrize("chunk", [32, 64, 256, 1024, 4096, 16384, 65536])
def test_three_phase_chunk_lengths(O, chunk):
    for w in (1, 4, 8):
        for n in (chunk - 1, chunk * 5 + 3, 300007):
            print(f"three-phase chunk={chunk} w={w} n={n}", flush=True)
            data = O.fill_values(n, NP[w], seed=n)
            mask = O.mask_bernoulli(n, 0.3, seed=n + 1)
            expect, ke = O.compact(data, mask, n)
            got, k = run_compact(data, mask, n, oracle.MSB_FIRST, csx.PIPELINE_THREE_PHASE, chunk=chunk)
            assert k == ke and (got == expect).all()


@pytest.mark.parametrize("order", [oracle.MSB_FIRST, oracle.LSB_FIRST])
@pytest.mark.parametrize("chunk", [32, 64, 128, 256, 1024, 4096, 8192, 16384, 1 << 17])
def test_popc_parity(O, chunk, order):
    for n in (1, 7, 31, 32, 33, 127, 128, 129, chunk - 1, chunk, chunk + 1, 1000003):
        print(f"popc chunk={chunk} order={order} n={n}", flush=True)
        mask = dirty_tail(O.mask_bernoulli(n, 0.45, seed=n, bit_order=order), n, order)
        expect = O.popc(mask, n, chunk, order)
        got = host(csx.popc(dev(mask), n, chunk, order), np.uint32)
        assert (got == expect).all()


def test_popc_unaligned_mask(O):
    n = 100000
    base = O.mask_bernoulli(n + 64, 0.5, seed=5)
    for off in (1, 3, 8):
        mask = base[off:]
        expect = O.popc(mask, n, 256)
        d = dev(base)[off:]
        got = host(csx.popc(d, n, 256), np.uint32)
        assert (got == expect).all()


@pytest.mark.parametrize("m", [1, 2, 4095, 4096, 4097, 100000, (1 << 20) + 3, 5 << 20])
def test_scan_parity(O, m):
    rng = np.random.default_rng(m)
    counts = rng.integers(0, 4096, m).astype(np.uint32)
    expect, total = O.exclusive_scan(counts)
    offs, tot = csx.scan(dev(counts))
    assert int(tot.item()) == total
    assert (host(offs, np.uint64) == expect).all()


def test_scan_large_values(O):
    m = 300000
    counts = np.full(m, 0xFFFFFFFF, np.uint32)
    expect, total = O.exclusive_scan(counts)
    offs, tot = csx.scan(dev(counts))
    assert int(tot.item()) == total == m * 0xFFFFFFFF
    assert (host(offs, np.uint64) == expect).all()


@pytest.mark.parametrize("m", [1, 2, 3, 7, 64, 1000, 4097, 262144, 300001])
def test_scan_tree_and_walk(O, m):
    rng = np.random.default_rng(m + 1)
    counts = rng.integers(0, 1024, m).astype(np.uint32)
    tree_expect = O.upsweep_tree(counts)
    d = dev(counts)
    tree, tot = csx.scan_tree(d)
    assert int(tot.item()) == int(counts.astype(np.int64).sum())
    got_tree = host(tree, np.uint32)
    assert (got_tree == tree_expect).all()
    offs = host(csx.tree_offsets(tree), np.uint32)
    sample = np.unique(np.concatenate([np.arange(min(m, 50)), rng.integers(0, m, 200)]))
    for idx in sample:
        assert offs[idx] == O.tree_prefix(tree_expect, int(idx))
    csum = np.concatenate([[0], np.cumsum(counts.astype(np.int64))[:-1]]).astype(np.uint32)
    assert (offs == csum).all()


@pytest.mark.parametrize("variant", [csx.WRITEOUT_STAGED, csx.WRITEOUT_DIRECT, csx.WRITEOUT_THREAD_PER_CHUNK, csx.WRITEOUT_SPARSE])
@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_writeout_variants(O, w, variant):
    for chunk, n, kind, skip in ((32, 70001, "bern50", False), (1024, 250007, "zipf", True), (4096, 123457, "runs", True),
                                 (65536, 300000, "bern01", True), (256, 9999, "zeros", True)):
        for order in (oracle.MSB_FIRST, oracle.LSB_FIRST):
            print(f"writeout w={w} variant={variant} chunk={chunk} n={n} mask={kind} skip={skip} order={order}", flush=True)
            data = O.fill_values(n, NP[w], seed=n + w)
            mask = dirty_tail(make_mask(O, kind, n, n + 9, order), n, order)
            counts = O.popc(mask, n, chunk, order)
            offs, total = O.exclusive_scan(counts)
            expect, ke = O.compact(data, mask, n, order)
            assert total == ke
            canary = np.frombuffer(np.full((n + 16) * w, 0x5A, np.uint8).tobytes(), dtype=NP[w]).copy()
            # the staged tile at every workgroup size, the other variants at 256
            for block in ((0, 128, 512) if variant == csx.WRITEOUT_STAGED else (0,)):
                d_out = dev(canary)
                csx.writeout(dev(data), dev(mask), dev(offs), n, chunk, order, variant, counts=dev(counts) if skip else None, out=d_out,
                             block_size=block)
                got = host(d_out, NP[w])
                assert (got[:ke] == expect).all(), block
                assert (got[ke:] == canary[ke:]).all(), block


def test_unaligned_pointers(O):
    n = 100003
    for w in (1, 2, 4, 8):
        for pipeline in PIPELINES:
            data = O.fill_values(n + 8, NP[w], seed=w)
            mask_full = O.mask_bernoulli((n + 64) * 1, 0.5, seed=w + 1)
            for ioff, moff, ooff in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (3, 5, 7)):
                print(f"unaligned w={w} pipeline={pipeline} in+{ioff} mask+{moff} out+{ooff}", flush=True)
                mask = mask_full[moff:]
                expect, ke = O.compact(data[ioff:], mask, n)
                d_in = dev(data)[ioff:]
                d_mask = dev(mask_full)[moff:]
                d_outbuf = torch.zeros(n + 16, dtype=dev(data).dtype, device="cuda")
                d_out = d_outbuf[ooff:]
                _, k = csx.compact(d_in, d_mask, n=n, pipeline=pipeline, out=d_out)
                got = host(d_out, NP[w])[:k]
                assert k == ke and (got == expect).all()


def test_zero_elements(O):
    d = torch.zeros(4, dtype=torch.int32, device="cuda")
    m = torch.zeros(1, dtype=torch.uint8, device="cuda")
    for pipeline in PIPELINES:
        _, k = csx.compact(d, m, n=0, pipeline=pipeline)
        assert k == 0


def test_reference_generators_end_to_end(O):
    """The reference's own inputs (src/main.cu:76, src/data_generator.cuh:14-188):
    data bytes = uniform mask at p=0.5, seed 42; masks from each generator;
    output checked against the restated generate_validation."""
    n = 1 << 18
    data = O.mask_uniform(n * 4, 0.5)[0].view(np.uint32)
    masks = {
        "uniform": O.mask_uniform(n // 8, 0.3)[0],
        "zipf": O.mask_zipf(n // 8)[0],
        "burst": O.mask_burst(n // 8, 0.01),
        "offset": O.mask_offset(n // 8, 7),
        "pattern": O.mask_pattern(n // 8, O.pattern_from_selectivity(8, 0.5), 8)[0],
        "clustering": O.mask_clustering(0.5, 1, n)[0],
    }
    for name, mask in masks.items():
        for pipeline in PIPELINES:
            print(f"reference inputs mask={name} pipeline={pipeline}", flush=True)
            expect, ke = O.generate_validation(data, mask)
            got, k = run_compact(data, mask, n, oracle.MSB_FIRST, pipeline)
            assert k == ke and (got == expect[:ke]).all()


def test_validate_counts_mismatches(O):
    a = np.arange(10000, dtype=np.uint32)
    b = a.copy()
    b[[0, 17, 9999]] += 1
    f = csx.validate(dev(a), dev(b), 10000)
    assert int(f.item()) == 3
    f = csx.validate(dev(a), dev(a), 10000)
    assert int(f.item()) == 0


def test_repeat_is_deterministic(O):
    n = 1 << 20
    data = O.fill_values(n, np.uint32, seed=77)
    mask = O.mask_bernoulli(n, 0.5, seed=78)
    for pipeline in PIPELINES:
        a, ka = run_compact(data, mask, n, oracle.MSB_FIRST, pipeline)
        b, kb = run_compact(data, mask, n, oracle.MSB_FIRST, pipeline)
        assert ka == kb and (a == b).all()


def test_two_pass_many_groups(O):
    # more tiles than 2048 groups of 16: exercises larger groups and
    # multi-load group sums in the writeout
    n = (1 << 26) + 12345
    data = O.fill_values(n, np.uint8, seed=3)
    mask = O.mask_bernoulli(n, 0.2, seed=4)
    expect, ke = O.compact(data, mask, n)
    got, k = run_compact(data, mask, n, oracle.MSB_FIRST, csx.PIPELINE_TWO_PASS, canary=False)
    assert k == ke and (got == expect).all()


@pytest.mark.parametrize("block", [128, 512])
@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_three_phase_workgroup_sizes(O, w, block):
    """The three-phase pipeline with 128- and 512-thread tiles at every tile
    size, both bit orders, ragged sizes and chunk lengths below and above the
    tile (units of min(chunk, tile))."""
    for vpt in ((2, 4, 8) if w == 1 else (2, 4, 8, 16)):
        te = block * vpt * (16 // w)
        for n, chunk, kind in ((te * 3 + 17, 1024, "bern50"), (te - 1, 64, "bern99"), (5 * te + 1, 1 << 17, "runs"), (777, 256, "bern01")):
            for order in (oracle.MSB_FIRST, oracle.LSB_FIRST):
                print(f"three-phase block={block} vpt={vpt} w={w} n={n} chunk={chunk} mask={kind} order={order}", flush=True)
                data = O.fill_values(n, NP[w], seed=n + vpt)
                mask = dirty_tail(make_mask(O, kind, n, n + block, order), n, order)
                expect, ke = O.compact(data, mask, n, order)
                canary = np.frombuffer(np.full((n + 16) * w, 0xC3, np.uint8).tobytes(), dtype=NP[w]).copy()
                d_out = dev(canary)
                _, k = csx.compact_config(dev(data), dev(mask), csx.Config(csx.PIPELINE_THREE_PHASE, chunk, vpt, 0, block), n=n, bit_order=order,
                                          out=d_out)
                # the writeout ran block-thread tiles; an input of at most one 256-thread tile of
                # vpt vectors runs whole in one workgroup
                one_tile = 256 * vpt * (16 // w)
                assert csx.last_launch() == ((256, 1) if n <= one_tile else (block, (n + te - 1) // te))
                got = host(d_out, NP[w])
                assert k == ke and (got[:k] == expect).all()
                assert (got[k:] == canary[k:]).all()


@pytest.mark.parametrize("w", [1, 4, 8])
def test_few_blocks_edges(O, w):
    """Columns of at most 128 popcount blocks skip the block prefixes: every
    writeout wave (staged tile or sparse step) sums the totals before its
    block and the last tile or step writes the count.  Sizes where the blocks
    grow to 2^18 and 2^19 elements, the largest such column (2^26: 128
    blocks) and one element more (the prefix kernel again); the last 300001
    elements unselected, so the last tiles are empty and the count comes
    from an empty tile."""
    for n in ((1 << 24) + 12345, (1 << 25) + 99, 1 << 26, (1 << 26) + 1):
        data = O.fill_values(n, NP[w], seed=n % 1000 + w)
        d_in = dev(data)
        for order in (oracle.MSB_FIRST, oracle.LSB_FIRST):
            mask = O.mask_bernoulli(n, 0.3, seed=n % 997 + order, bit_order=order)
            mask[(n - 300001) // 8 + 1:] = 0
            mask = dirty_tail(mask, n, order)
            expect, ke = O.compact(data, mask, n, order)
            d_mask, d_exp = dev(mask), dev(expect)
            for pipeline, chunk, vpt, block in ((csx.PIPELINE_THREE_PHASE, 1024, 4, 256), (csx.PIPELINE_THREE_PHASE, 1024, 8, 512),
                                                (csx.PIPELINE_SPARSE, 4096, 8, 256)):
                print(f"few blocks w={w} n={n} order={order} pipeline={pipeline} vpt={vpt} block={block}", flush=True)
                d_out = torch.full(((n + 16) * w,), 0xC3, dtype=torch.uint8, device="cuda").view(d_in.dtype)
                _, k = csx.compact_config(d_in, d_mask, csx.Config(pipeline, chunk, vpt, 0, block), n=n, bit_order=order, out=d_out)
                assert k == ke and torch.equal(d_out[:k], d_exp)
                assert bool((d_out[k:].view(torch.uint8) == 0xC3).all())


@pytest.mark.parametrize("w", [1, 4, 8])
def test_block_prefix_path_unaligned(O, w):
    """Columns past 2^26 elements take the block prefixes (popcount, prefix
    kernel, writeout).  Input, mask and output off their 16-byte alignment
    (the unaligned popcount and tile bodies), both bit orders, a ragged
    length, the staged tile at two shapes and the sparse writeout."""
    n = (1 << 26) + 4097
    ioff, moff, ooff = 3, 5, 1
    data = O.fill_values(n + ioff, NP[w], seed=w + 11)
    d_data = dev(data)
    for order in (oracle.MSB_FIRST, oracle.LSB_FIRST):
        mask_full = dirty_tail(O.mask_bernoulli(n + 8 * moff, 0.4, seed=order + w, bit_order=order)[: (n + 7) // 8 + moff], n + 8 * moff, order)
        mask = mask_full[moff:]
        expect, ke = O.compact(data[ioff:], mask, n, order)
        d_mask_full, d_exp = dev(mask_full), dev(expect)
        for pipeline, chunk, vpt, block in ((csx.PIPELINE_THREE_PHASE, 4096, 2, 256), (csx.PIPELINE_THREE_PHASE, 1024, 8, 512),
                                            (csx.PIPELINE_SPARSE, 4096, 8, 256)):
            print(f"block prefixes unaligned w={w} n={n} order={order} pipeline={pipeline} vpt={vpt} block={block}", flush=True)
            d_outbuf = torch.full(((n + 16) * w,), 0xC3, dtype=torch.uint8, device="cuda").view(d_data.dtype)
            d_out = d_outbuf[ooff:]
            _, k = csx.compact_config(d_data[ioff:], d_mask_full[moff:], csx.Config(pipeline, chunk, vpt, 0, block), n=n, bit_order=order,
                                      out=d_out)
            assert k == ke and torch.equal(d_out[:k], d_exp)
            assert bool((d_out[k:].view(torch.uint8) == 0xC3).all())


@pytest.mark.parametrize("vpt", [2, 4, 8])
@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_single_launch_tile_sizes(O, w, vpt):
    """The single-launch pipeline with 2-, 4- and 8-vector tiles: at most one
    tile (one workgroup), just past it, a few tiles, and columns whose tiles
    count more than one round (64 KiB) of mask before them, in both bit
    orders, with the launch shape it reports."""
    te = 256 * vpt * (16 // w)
    for n, kind in ((te - 5, "bern50"), (te, "runs"), (te + 1, "bern50"), (3 * te + 17, "runs"), (600_000 + 333, "bern01"),
                    ((1 << 20) + 77, "bern99")):
        for order in (oracle.MSB_FIRST, oracle.LSB_FIRST):
            print(f"single launch vpt={vpt} w={w} n={n} mask={kind} order={order}", flush=True)
            data = O.fill_values(n, NP[w], seed=n + vpt)
            mask = dirty_tail(make_mask(O, kind, n, n + vpt, order), n, order)
            expect, ke = O.compact(data, mask, n, order)
            canary = np.frombuffer(np.full((n + 16) * w, 0xC3, np.uint8).tobytes(), dtype=NP[w]).copy()
            d_out = dev(canary)
            _, k = csx.compact_config(dev(data), dev(mask), csx.Config(csx.PIPELINE_SINGLE, 1024, vpt, 0), n=n, bit_order=order, out=d_out)
            assert csx.last_launch() == (256, (n + te - 1) // te)  # up to one tile: the one-workgroup form
            got = host(d_out, NP[w])
            assert k == ke and (got[:k] == expect).all()
            assert (got[k:] == canary[k:]).all()


def test_single_launch_at_its_cap():
    """The single-launch pipeline at CSX_SINGLE_MAX_ELEMENTS (the last tile
    counts 2 MiB of mask before it) against torch.nonzero, a ragged size just
    below, and the first size past the cap refused before any launch."""
    for n, p in ((csx.SINGLE_MAX_ELEMENTS, 0.5), (csx.SINGLE_MAX_ELEMENTS - 77, 0.02)):
        print(f"single launch n={n} p={p}", flush=True)
        g = torch.Generator(device="cuda")
        g.manual_seed(n)
        bits = torch.rand(n, device="cuda", generator=g) < p
        pad = (-n) % 8
        padded = torch.cat([bits, torch.zeros(pad, dtype=torch.bool, device="cuda")])
        weights = torch.tensor([128, 64, 32, 16, 8, 4, 2, 1], dtype=torch.uint8, device="cuda")
        mask = (padded.view(-1, 8).to(torch.uint8) * weights).sum(1, dtype=torch.uint8)
        data = torch.arange(n, dtype=torch.int32, device="cuda")
        out, k = csx.compact(data, mask, n=n, pipeline=csx.PIPELINE_SINGLE)
        expect = torch.nonzero(bits).flatten().to(torch.int32)
        assert k == expect.numel()
        assert torch.equal(out[:k], expect)
    big = csx.SINGLE_MAX_ELEMENTS + 1
    with pytest.raises(csx.CsxError) as e:
        csx.compact(torch.empty(big, dtype=torch.uint8, device="cuda"), torch.empty((big + 7) // 8, dtype=torch.uint8, device="cuda"), n=big,
                    pipeline=csx.PIPELINE_SINGLE)
    assert e.value.status == 4


@pytest.mark.parametrize("p", [0.01, 0.5, 0.99])
@pytest.mark.parametrize("pipeline", PIPELINES_ANY_N)
def test_full_size_2p28_index_identity(p, pipeline):
    """2^28 u32 with data[i] = i: the output must be exactly the selected
    indices, checked against torch.nonzero of the unpacked mask."""
    n = 1 << 28
    print(f"full size p={p} pipeline={pipeline}", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(int(p * 1000) + pipeline)
    bits = torch.rand(n, device="cuda", generator=g) < p
    weights = torch.tensor([128, 64, 32, 16, 8, 4, 2, 1], dtype=torch.uint8, device="cuda")
    mask = (bits.view(-1, 8).to(torch.uint8) * weights).sum(1, dtype=torch.uint8)
    data = torch.arange(n, dtype=torch.int32, device="cuda")
    out, k = csx.compact(data, mask, pipeline=pipeline)
    expect = torch.nonzero(bits).flatten().to(torch.int32)
    assert k == expect.numel()
    assert torch.equal(out[:k], expect)
    del bits, mask, data, out, expect
    torch.cuda.empty_cache()


def test_scan_u32_in_place_and_separate(O):
    rng = np.random.default_rng(9)
    for m in (1, 4096, 4097, 262144 + 5):
        counts = rng.integers(0, 1025, m).astype(np.uint32)
        expect, total = O.exclusive_scan(counts)
        d = dev(counts)
        offs, tot = csx.scan_u32(d)
        assert int(tot.item()) == total and (host(offs, np.uint32) == expect.astype(np.uint32)).all()
        offs2, tot2 = csx.scan_u32(d, offsets=d)  # in place
        assert int(tot2.item()) == total and (host(d, np.uint32) == expect.astype(np.uint32)).all()


@pytest.mark.parametrize("variant", [csx.WRITEOUT_STAGED, csx.WRITEOUT_DIRECT, csx.WRITEOUT_THREAD_PER_CHUNK, csx.WRITEOUT_SPARSE])
def test_writeout_u32_offsets_and_out_base(O, variant):
    n, chunk = 200003, 512
    for w in (1, 4, 8):
        data = O.fill_values(n, NP[w], seed=w + 40)
        mask = O.mask_bernoulli(n, 0.4, seed=w + 41)
        counts = O.popc(mask, n, chunk)
        offs, total = O.exclusive_scan(counts)
        expect, ke = O.compact(data, mask, n)
        base = 777
        d_out = torch.zeros(n + base + 16, dtype=dev(data).dtype, device="cuda")
        d_base = torch.tensor([base], dtype=torch.int64, device="cuda")
        csx.writeout(dev(data), dev(mask), dev(offs.astype(np.uint32)), n, chunk, variant=variant, out=d_out, out_base=d_base)
        got = host(d_out, NP[w])
        assert (got[base:base + ke] == expect).all()
        assert (got[:base] == 0).all() and (got[base + ke:] == 0).all()


@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_pattern_compact(O, w):
    rng = np.random.default_rng(w)
    cases = [(0xB5000000, 8), (0xDEADBEEF, 32), (0x80000000, 1), (0x6C000000, 7), (0xFFFFFFFF, 32), (0x00000001, 32),
             (0x12345678, 27), (0xF0000000, 5), (0x0, 8)]
    for pat, L in cases:
        for n in (0, 1, 7, 33, 1000, 65537, 1 << 20):
            print(f"pattern w={w} pattern={pat:#x} L={L} n={n}", flush=True)
            data = O.fill_values(n, NP[w], seed=n + L)
            mask, _ = O.mask_pattern((n + 7) // 8, pat & (0xFFFFFFFF << (32 - L)) & 0xFFFFFFFF, L)
            expect, ke = O.compact(data, mask, n)
            assert csx.pattern_count(n, pat, L) == ke
            canary = np.frombuffer(np.full((n + 8) * w, 0x3C, np.uint8).tobytes(), dtype=NP[w]).copy()
            d_out = dev(canary)
            d_cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
            _, k = csx.pattern_compact(dev(data) if n else torch.zeros(1, dtype=dev(canary).dtype, device="cuda"), pat, L, n=n, out=d_out,
                                       count=d_cnt)
            got = host(d_out, NP[w])
            assert k == ke and int(d_cnt.item()) == ke
            assert (got[:k] == expect).all() and (got[k:] == canary[k:]).all()
    # an output of exactly the k selected elements right before the input in
    # one allocation (an arena's neighbouring ranges): accepted, and nothing
    # of the input beyond is touched
    pat, L, n = 0xB5000000, 8, 100003
    data = O.fill_values(n, NP[w], seed=17)
    mask, _ = O.mask_pattern((n + 7) // 8, pat, L)
    expect, ke = O.compact(data, mask, n)
    buf = dev(np.concatenate([np.zeros(ke, NP[w]), data]))
    out, k = csx.pattern_compact(buf[ke:], pat, L, n=n, out=buf[:ke])
    got = host(buf, NP[w])
    assert k == ke and (got[:k] == expect).all() and (got[ke:] == data).all()
    # the reference driver's own pattern (src/main.cu:54-64) at p = 0.5, L = 8
    pat = O.pattern_from_selectivity(8, 0.5)
    n = 300000
    data = O.fill_values(n, NP[w], seed=5)
    mask, _ = O.mask_pattern((n + 7) // 8, pat, 8)
    expect, ke = O.compact(data, mask, n)
    out, k = csx.pattern_compact(dev(data), pat, 8)
    assert k == ke and (host(out, NP[w])[:k] == expect).all()
    # input not 16-byte aligned (element loads), output offset too
    if w < 8:
        d = dev(data)
        for off in (1, 3):
            mask, _ = O.mask_pattern((n - off + 7) // 8, pat, 8)
            expect, ke = O.compact(data[off:], mask, n - off)
            d_out = torch.zeros(n + 8, dtype=d.dtype, device="cuda")[off:]
            _, k = csx.pattern_compact(d[off:], pat, 8, out=d_out)
            assert k == ke and (host(d_out, NP[w])[:k] == expect).all(), off


@pytest.mark.parametrize("nstreams", [1, 2, 3, 8])
def test_compact_streams(O, nstreams):
    streams = [torch.cuda.Stream() for _ in range(nstreams)]
    for w, n, chunk in ((4, (1 << 20) + 77, 1024), (1, 3 * 32768 + 5, 4096), (8, 100, 32), (2, 0, 256), (4, 5 * 8192, 65536)):
        print(f"streams={nstreams} w={w} n={n} chunk={chunk}", flush=True)
        data = O.fill_values(n, NP[w], seed=n + 3)
        mask = O.mask_bernoulli(n, 0.37, seed=n + 4)
        expect, ke = O.compact(data, mask, n)
        d_in = dev(data) if n else torch.zeros(1, dtype=dev(np.zeros(1, NP[w])).dtype, device="cuda")
        d_mask = dev(mask) if n else torch.zeros(1, dtype=torch.uint8, device="cuda")
        canary = np.frombuffer(np.full((n + 8) * w, 0x77, np.uint8).tobytes(), dtype=NP[w]).copy()
        d_out = dev(canary)
        torch.cuda.synchronize()
        _, k = csx.compact_streams(d_in, d_mask, streams, n=n, chunk_len=chunk, out=d_out)
        torch.cuda.synchronize()
        got = host(d_out, NP[w])
        assert k == ke and (got[:k] == expect).all() and (got[k:] == canary[k:]).all()


def test_rocprim_baseline_agrees(O):
    from gpu_compressstore2_amd import baseline

    for w, n in ((4, (1 << 20) + 3), (1, 77777), (8, 4096)):
        data = O.fill_values(n, NP[w], seed=w)
        mask = O.mask_bernoulli(n, 0.5, seed=w + 1)
        expect, ke = O.compact(data, mask, n)
        out, cnt = baseline.rocprim_select(dev(data), dev(mask), n=n)
        assert int(cnt.item()) == ke
        assert (host(out, NP[w])[:ke] == expect).all()
