"""Predicate -> mask and fused predicate compaction on the device, bit-exact
against the oracle (csxo_predicate_mask_typed + compact)."""
import numpy as np
import pytest
import torch

import gpu_compressstore2_amd as csx
from test_predicate_oracle import TYPES, column

pytestmark = pytest.mark.gpu


def raw(a, offset_elems=0):
    """Device bytes of `a`, starting offset_elems elements into a fresh buffer."""
    b = np.ascontiguousarray(a).view(np.uint8)
    w = a.dtype.itemsize
    buf = torch.zeros(b.size + 64 + offset_elems * w, dtype=torch.uint8, device="cuda")
    buf[offset_elems * w:offset_elems * w + b.size] = torch.from_numpy(b).cuda()
    return buf[offset_elems * w:]


def bounds(op):
    return (0, 1) if op == 6 else (1, 0)


@pytest.mark.parametrize("dtype", TYPES)
def test_predicate_mask_and_compaction_match_oracle(O, dtype):
    t = oracle_type = TYPES.index(dtype)
    w = np.dtype(dtype).itemsize
    for n in (1, 37, 1024, 5000, 131083):
        c = column(dtype, n, 100 + n)
        for op in range(7):
            a, b = bounds(op)
            for order in (0, 1):
                expect_mask, k = O.predicate_mask_typed(c, op, a, b, bit_order=order)
                d = raw(c, offset_elems=(op + order) % 3)  # 16-byte aligned and not
                mask, count = csx.predicate_mask(d, op, a, b, n=n, col_type=t, bit_order=order)
                nb = (n + 7) // 8
                assert int(count.item()) == k, (dtype, n, op, order)
                assert (mask[:nb].cpu().numpy() == expect_mask).all(), (dtype, n, op, order)
                # fused compaction, mask written too; canary past the output
                out = torch.full((n * w + 64,), 0xA5, dtype=torch.uint8, device="cuda")
                m2 = torch.full((nb + 16,), 0x5A, dtype=torch.uint8, device="cuda")
                _, kk = csx.compact_predicate(d, op, a, b, n=n, col_type=oracle_type, bit_order=order, out=out, mask=m2)
                exp_out, k2 = O.compact(c, expect_mask, n, bit_order=order)
                assert kk == k == k2, (dtype, n, op, order)
                got = out.cpu().numpy()
                assert (got[: k * w].view(c.dtype).view(np.uint8) == exp_out[:k].view(np.uint8)).all(), (dtype, n, op, order)
                assert (got[k * w:] == 0xA5).all()
                assert (m2[:nb].cpu().numpy() == expect_mask).all() and (m2[nb:].cpu().numpy() == 0x5A).all()


def test_empty_column():
    d = torch.zeros(16, dtype=torch.float32, device="cuda")
    mask, count = csx.predicate_mask(d, ">", 0.0, n=0)
    assert int(count.item()) == 0
    _, k = csx.compact_predicate(d, ">", 0.0, n=0)
    assert k == 0


def test_reference_predicate_on_device(O):
    """`value > 200` over float32 values in [0, 400): the reference driver's predicate."""
    rng = np.random.default_rng(9)
    c = (rng.random(1 << 20, dtype=np.float32) * 400).astype(np.float32)
    m_ref, k_ref = O.predicate_mask(c, 200.0)
    d = torch.from_numpy(c).cuda()
    mask, count = csx.predicate_mask(d, ">", 200.0)
    assert int(count.item()) == k_ref and (mask.cpu().numpy() == m_ref).all()
    out, k = csx.compact_predicate(d, ">", 200.0)
    assert k == k_ref and torch.equal(out[:k].cpu(), torch.from_numpy(c[c > 200.0]))


def test_full_size_fused_equals_two_steps():
    """2^28 float32: fused predicate compaction equals predicate_mask then
    csx_compact, every output passes the predicate, and the count is the
    mask's popcount."""
    n = 1 << 28
    g = torch.Generator(device="cuda").manual_seed(4)
    col = torch.rand(n, device="cuda", generator=g)
    mask, count = csx.predicate_mask(col, ">", 0.5)
    out1, k1 = csx.compact(col, mask, n=n)
    out2, k2 = csx.compact_predicate(col, ">", 0.5)
    assert k1 == k2 == int(count.item())
    assert abs(k1 - n / 2) < 6 * (n ** 0.5)
    assert torch.equal(out1[:k1], out2[:k2])
    assert bool((out2[:k2] > 0.5).all())
    del out1
    ref = torch.masked_select(col, col > 0.5)
    assert torch.equal(ref, out2[:k2])
