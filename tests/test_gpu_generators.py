"""Device generators of libcsx.so against the oracle's sequential generators,
bit-exact (and, through the oracle, against the reference's own generators)."""
import numpy as np
import pytest
import torch

import gpu_compressstore2_amd as csx
import oracle

pytestmark = pytest.mark.gpu

NS = [1, 7, 8, 31, 32, 33, 1000, 65536 + 3, (1 << 20) + 17]


@pytest.mark.parametrize("order", [oracle.MSB_FIRST, oracle.LSB_FIRST])
def test_bernoulli(O, order):
    for n in NS:
        for p, seed in ((0.5, 42), (0.01, 7), (0.99, 1), (0.0, 3), (1.0, 4)):
            got = csx.gen_mask_bernoulli(n, p, seed, order).cpu().numpy()[: (n + 7) // 8]
            assert (got == O.mask_bernoulli(n, p, seed, order)).all(), (n, p, seed)


def test_bernoulli_is_reference_uniform_mask(O):
    # seed 42, MSB-first: generate_mask_uniform of src/data_generator.cuh:14-30
    for nbytes, p in ((4096, 0.5), (1001, 0.3), (77, 0.99)):
        ref, _ = O.mask_uniform(nbytes, p)
        got = csx.gen_mask_bernoulli(nbytes * 8, p, 42).cpu().numpy()
        assert (got == ref).all()


def test_reference_data_fill(O):
    # the reference driver fills the column with uniform-mask bytes (src/main.cu:76)
    n = 100003
    ref, _ = O.mask_uniform(n * 4, 0.5)
    got = csx.gen_mask_bernoulli(n * 32, 0.5, 42).cpu().numpy()
    assert (got == ref).all()


@pytest.mark.parametrize("order", [oracle.MSB_FIRST, oracle.LSB_FIRST])
def test_pattern(O, order):
    for n in NS:
        for pat, L in ((0xB5000000, 8), (0xDEADBEEF, 32), (0x6C000000, 7), (0x80000000, 1)):
            got = csx.gen_mask_pattern(n, pat, L, order).cpu().numpy()[: (n + 7) // 8]
            bits = np.array([(pat >> (31 - (k % L))) & 1 for k in range(n)], np.uint8)
            expect = np.packbits(bits, bitorder="big" if order == oracle.MSB_FIRST else "little")
            assert (got == expect).all(), (n, pat, L)
            if order == oracle.MSB_FIRST and n % 8 == 0:
                ref, _ = O.mask_pattern(n // 8, pat, L)
                assert (got == ref).all()
    for n in (1000, 4099):
        got = csx.gen_mask_pattern(n, 0x9E3779B9, 32, order).cpu().numpy()[: (n + 7) // 8]
        assert (got == O.mask_periodic(n, 0x9E3779B9, order)).all()


@pytest.mark.parametrize("order", [oracle.MSB_FIRST, oracle.LSB_FIRST])
def test_zipf_blocks(O, order):
    for n in (1, 4096, 4097, 1 << 18, (1 << 20) + 100):
        for block, s, levels, seed in ((4096, 1.2, 65, 5), (1024, 2.0, 17, 9)):
            got = csx.gen_mask_zipf_blocks(n, block, s, levels, seed, order).cpu().numpy()[: (n + 7) // 8]
            assert (got == O.mask_zipf_blocks(n, block, s, levels, seed, order)).all(), (n, block)


@pytest.mark.parametrize("order", [oracle.MSB_FIRST, oracle.LSB_FIRST])
def test_runs(O, order):
    for n in (1, 31, 32, 33, 5000, 1 << 18, (1 << 20) + 5):
        for mean, seed in ((256, 3), (16, 4), (4096, 5)):
            got = csx.gen_mask_runs(n, mean, seed, order).cpu().numpy()[: (n + 7) // 8]
            assert (got == O.mask_geometric_runs(n, mean, seed, order)).all(), (n, mean)


@pytest.mark.parametrize("w,kind", [(1, 0), (2, 0), (4, 0), (8, 0), (4, 1), (8, 1)])
def test_values(O, w, kind):
    dt = {(1, 0): np.uint8, (2, 0): np.uint16, (4, 0): np.uint32, (8, 0): np.uint64, (4, 1): np.float32, (8, 1): np.float64}[(w, kind)]
    for n in (1, 5, 1000, 100003):
        got = csx.gen_values(n, w, kind, seed=n).cpu().numpy().view(dt)[:n]
        expect = O.fill_values(n, dt, seed=n, kind=kind)
        assert (got.view(np.uint8) == expect.view(np.uint8)).all()
