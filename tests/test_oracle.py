"""CPU checks that pin the oracle (oracle/csx_oracle.c).

Pinned by: the reference's own CPU code compiled from its sources
(oracle/_ref), golden vectors written from it (tests/golden), published
known answers (splitmix64, PCG32), numpy's independent bit unpacking, and
brute force on tiny inputs.
"""
import itertools

import numpy as np
import pytest

import oracle
from conftest import unhex

# ------------------------------------------------------------------ PRNG


def test_splitmix64_known_answer(O):
    # splitmix64 (Steele, Lea, Flood 2014) first output for state 0
    st = oracle.ctypes_prng()
    O.lib.csxo_prng_seed(oracle.ctypes.byref(st), 0)
    assert st.state == 0xE220A8397B1DCDAF
    assert st.inc == 1


def test_pcg32_known_answer(O):
    # pcg32_srandom_r(42, 54) then pcg32_random_r, the PCG basic C demo's first outputs
    st = oracle.ctypes_prng()
    st.state = 0
    st.inc = (54 << 1) | 1
    O.lib.csxo_prng_next(oracle.ctypes.byref(st))
    st.state = (st.state + 42) & (2**64 - 1)
    O.lib.csxo_prng_next(oracle.ctypes.byref(st))
    got = [O.lib.csxo_prng_next(oracle.ctypes.byref(st)) for _ in range(6)]
    assert got == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_prng_advance_matches_stepping(O):
    seq = O.prng_seq(99, 5000)
    for delta in (0, 1, 2, 63, 64, 1000, 4999):
        st = oracle.ctypes_prng()
        O.lib.csxo_prng_seed(oracle.ctypes.byref(st), 99)
        O.lib.csxo_prng_advance(oracle.ctypes.byref(st), delta)
        assert O.lib.csxo_prng_next(oracle.ctypes.byref(st)) == seq[delta]


# ------------------------------------------------------- golden vectors


def test_golden_vectors(O, golden):
    for c in golden:
        fn = c["fn"]
        if fn == "prng_seq":
            assert (O.prng_seq(c["seed"], c["n"]) == unhex(c["out"], np.uint32)).all()
        elif fn == "mask_uniform":
            m, cnt = O.mask_uniform(c["nbytes"], c["p"])
            assert (m == unhex(c["out"])).all() and cnt == c["count"]
        elif fn == "mask_zipf":
            m, cnt = O.mask_zipf(c["nbytes"])
            assert (m == unhex(c["out"])).all() and cnt == c["count"]
        elif fn == "mask_burst":
            assert (O.mask_burst(c["nbytes"], c["segment_sizer"]) == unhex(c["out"])).all()
        elif fn == "mask_offset":
            assert (O.mask_offset(c["nbytes"], c["spacing"]) == unhex(c["out"])).all()
        elif fn == "mask_pattern":
            m, cnt = O.mask_pattern(c["nbytes"], c["pattern"], c["length"])
            assert (m == unhex(c["out"])).all() and cnt == c["count"]
        elif fn == "mask_clustering":
            m, cnt = O.mask_clustering(c["selectivity"], c["clusters"], c["total"])
            assert (m == unhex(c["out"])).all() and cnt == c["count"]
        elif fn == "generate_validation":
            dt = np.dtype(c["dtype"])
            data, mask = unhex(c["data"], dt), unhex(c["mask"])
            out, cnt = O.generate_validation(data, mask)
            assert cnt == c["count"] and (out == unhex(c["out"], dt)).all()
            # the GPU semantics agree with the reference oracle for whole bytes
            n8 = c["n"] // 8 * 8
            comp, k = O.compact(data, mask, n8)
            assert k == cnt and (comp == out[:k]).all()
        else:
            raise AssertionError(fn)


# ---------------------------------------------- against the reference build


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.uint64])
@pytest.mark.parametrize("n", [0, 1, 7, 8, 9, 255, 1024, 4099])
def test_generate_validation_vs_reference(O, R, dtype, n):
    data = O.fill_values(n, dtype, seed=11)
    mask = O.mask_bernoulli(n, 0.37, seed=12)
    a, ca = O.generate_validation(data, mask)
    b, cb = R.generate_validation(data, mask)
    assert ca == cb and (a == b).all()


def test_generators_vs_reference(O, R):
    for nb, p in ((4000, 0.5), (999, 0.013), (512, 0.99)):
        assert (O.mask_uniform(nb, p)[0] == R.mask_uniform(nb, p)[0]).all()
    assert (O.mask_zipf(8192)[0] == R.mask_zipf(8192)[0]).all()
    assert (O.mask_burst(4096, 0.01) == R.mask_burst(4096, 0.01)).all()
    for pat, L in ((0x12345678, 29), (0xF0000000, 5)):
        assert (O.mask_pattern(1000, pat, L)[0] == R.mask_pattern(1000, pat, L)[0]).all()
    a = O.mask_clustering(0.4, 5, 100000)
    b = R.mask_clustering(0.4, 5, 100000)
    assert (a[0] == b[0]).all() and a[1] == b[1]


# -------------------------------------------------------- independent checks


@pytest.mark.parametrize("order", [oracle.MSB_FIRST, oracle.LSB_FIRST])
@pytest.mark.parametrize("n", [0, 1, 5, 8, 13, 64, 100, 1000, 77777])
def test_compact_vs_numpy(O, order, n):
    data = O.fill_values(n, np.uint32, seed=n + 1)
    mask = O.mask_bernoulli(n, 0.3, seed=n + 2, bit_order=order)
    # dirty tail bits past n must be ignored
    if n % 8:
        mask[-1] |= np.uint8(0xFF >> (n % 8)) if order == oracle.MSB_FIRST else np.uint8((0xFF << (n % 8)) & 0xFF)
    out, k = O.compact(data, mask, n, order)
    ref = oracle.numpy_compact(data, mask, n, order)
    assert k == ref.size and (out == ref).all()


def test_compact_brute_force_tiny(O):
    # every mask of up to 10 elements, both orders
    for n in range(0, 11):
        data = np.arange(100, 100 + n, dtype=np.uint16)
        for bits in itertools.product((0, 1), repeat=n):
            for order in (oracle.MSB_FIRST, oracle.LSB_FIRST):
                mask = np.zeros((n + 7) // 8 + 1, np.uint8)
                for i, b in enumerate(bits):
                    if b:
                        mask[i // 8] |= (1 << (7 - i % 8)) if order == oracle.MSB_FIRST else (1 << (i % 8))
                out, k = O.compact(data, mask, n, order)
                expect = [100 + i for i, b in enumerate(bits) if b]
                assert k == len(expect) and list(out) == expect


def test_msb_first_byte_0x80_selects_element_0(O):
    data = np.arange(8, dtype=np.uint32)
    out, k = O.compact(data, np.array([0x80], np.uint8), 8, oracle.MSB_FIRST)
    assert k == 1 and out[0] == 0
    out, k = O.compact(data, np.array([0x01], np.uint8), 8, oracle.LSB_FIRST)
    assert k == 1 and out[0] == 0


def test_trailing_partial_byte_deviation(O):
    # generate_validation walks n/8 whole bytes (src/data_generator.cuh:175):
    # elements of a trailing partial byte are dropped, the GPU semantics keep them
    data = np.arange(12, dtype=np.uint32)
    mask = np.array([0x00, 0xF0], np.uint8)  # elements 8..11 selected
    gv, c = O.generate_validation(data, mask)
    assert c == 0
    out, k = O.compact(data, mask, 12)
    assert k == 4 and list(out) == [8, 9, 10, 11]


@pytest.mark.parametrize("chunk", [32, 64, 128, 1024, 4096])
@pytest.mark.parametrize("n", [1, 31, 32, 33, 1000, 65537])
def test_popc_and_scan_closed_forms(O, chunk, n):
    mask = O.mask_bernoulli(n, 0.6, seed=n)
    mask_dirty = mask.copy()
    if n % 8:
        mask_dirty[-1] |= np.uint8(0xFF >> (n % 8))
    counts = O.popc(mask_dirty, n, chunk)
    bits = np.unpackbits(mask, bitorder="big")[:n].astype(np.int64)
    pad = (-n) % chunk
    expect = np.concatenate([bits, np.zeros(pad, np.int64)]).reshape(-1, chunk).sum(1)
    assert (counts == expect).all()
    offs, total = O.exclusive_scan(counts)
    assert total == bits.sum()
    assert (offs == np.concatenate([[0], np.cumsum(expect)[:-1]])).all()


@pytest.mark.parametrize("m", [1, 2, 3, 5, 8, 13, 64, 100, 1000, 4097])
def test_upsweep_tree_is_fenwick_and_walk_is_prefix(O, m):
    rng = np.random.default_rng(m)
    counts = rng.integers(0, 1000, m).astype(np.uint32)
    tree = O.upsweep_tree(counts)
    csum = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    for j in range(m):
        low = (j + 1) & -(j + 1)
        assert tree[j] == csum[j + 1] - csum[j + 1 - low]
    for idx in range(m):
        assert O.tree_prefix(tree, idx) == csum[idx]


def test_seeded_generators_properties(O):
    n = 1 << 16
    m = O.mask_bernoulli(n, 0.25, seed=3)
    frac = O.mask_count(m, n) / n
    assert abs(frac - 0.25) < 0.01
    # draw k decides element k: the same stream as the reference's uniform rule
    ref, _ = O.mask_uniform(n // 8, 0.25, seed=3)
    assert (m == ref).all()
    g = O.mask_geometric_runs(n, 256, seed=4)
    bits = np.unpackbits(g, bitorder="big")[:n]
    runs = np.diff(np.flatnonzero(np.diff(np.concatenate([[2], bits, [2]])) != 0))
    assert 150 < runs.mean() < 400
    z = O.mask_zipf_blocks(1 << 18, 4096, 1.2, 65, seed=5)
    per_block = np.unpackbits(z, bitorder="big").reshape(-1, 4096).mean(1)
    assert (per_block < 0.05).sum() > 10  # skewed towards empty blocks
    pat = 0xDEADBEEF
    pm = O.mask_periodic(96, pat)
    assert (np.unpackbits(pm, bitorder="big")[:32] == np.array([(pat >> (31 - i)) & 1 for i in range(32)])).all()


def test_fill_values_f32_range(O):
    v = O.fill_values(10000, np.float32, seed=8)
    assert v.min() >= 0.0 and v.max() < 1.0
    d = O.fill_values(10000, np.float64, seed=8)
    assert d.min() >= 0.0 and d.max() < 1.0


def test_pattern_from_selectivity(O):
    for L, s in ((8, 0.5), (32, 0.25), (7, 0.99), (1, 1.0)):
        p = O.pattern_from_selectivity(L, s)
        assert bin(p).count("1") == int(L * np.float32(s))
        assert p & ((1 << (32 - L)) - 1) == 0


def test_pattern_from_selectivity_from_reference_stream(O, golden):
    """src/main.cu:54-64 draws r from fast_prng(42) and sets bit 31 - (r mod L)
    unless it is set, until int(L * s) bits are set: the set positions are the
    first int(L * s) distinct values of r mod L in the stream.  The stream is
    the reference's own (the seed-42 golden vector written from oracle/_ref),
    so this pins the loop, its modulus, bit placement and count rounding
    independently of the oracle's PRNG."""
    stream = next(unhex(c["out"], np.uint32) for c in golden if c["fn"] == "prng_seq" and c["seed"] == 42)
    for L, s in ((8, 0.5), (32, 0.25), (32, 0.5), (7, 0.99), (16, 0.75), (5, 0.2), (3, 0.7), (1, 1.0), (12, 0.0)):
        ones = int(np.float32(L) * np.float32(s))  # the reference's int * float, truncated
        distinct = list(dict.fromkeys(int(r) % L for r in stream))[:ones]
        assert len(distinct) == ones  # 64 reference draws cover these cases
        expect = 0
        for i in distinct:
            expect |= 1 << (31 - i)
        assert O.pattern_from_selectivity(L, s) == expect, (L, s)
