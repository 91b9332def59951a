"""CPU checks of the predicate -> mask step: the oracle's comparison against
numpy's, the reference's own predicate case, and argument checks of the C ABI
that return before any device work."""
import ctypes

import numpy as np
import pytest

import gpu_compressstore2_amd as csx
import oracle

TYPES = [np.uint8, np.uint16, np.uint32, np.uint64, np.int8, np.int16, np.int32, np.int64, np.float32, np.float64]
NP_OPS = [np.greater, np.greater_equal, np.less, np.less_equal, np.equal, np.not_equal]


def column(dtype, n, seed):
    rng = np.random.default_rng(seed)
    dt = np.dtype(dtype)
    if n == 0:
        return np.zeros(0, dt)
    if dt.kind == "f":
        c = rng.uniform(-3, 3, n).astype(dt)
        c[rng.integers(0, n, max(n // 10, 1))] = 1.0  # ties with the threshold
        specials = np.array([np.nan, -np.nan, np.inf, -np.inf, -0.0, 0.0, 1.0, np.finfo(dt).max, np.finfo(dt).tiny], dt)
        c[: specials.size] = specials[: min(n, specials.size)]
    else:
        info = np.iinfo(dt)
        c = rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)
        c[rng.integers(0, n, max(n // 10, 1))] = 1
        specials = np.array([info.min, info.max, 0, 1], dt)
        c[: specials.size] = specials[: min(n, specials.size)]
    return c


def np_mask(sel, bit_order):
    return np.packbits(sel, bitorder="big" if bit_order == 0 else "little")


@pytest.mark.parametrize("dtype", TYPES)
def test_typed_predicate_matches_numpy(O, dtype):
    for n in (0, 1, 9, 1000, 4099):
        c = column(dtype, n, 5 + n)
        for op in range(7):
            for order in (0, 1):
                a, b = (1, 2) if op < 6 else (0, 1)
                got, ones = O.predicate_mask_typed(c, op, a, b, bit_order=order)
                sel = NP_OPS[op](c, np.array(a, dtype)) if op < 6 else (c >= np.array(a, dtype)) & (c <= np.array(b, dtype))
                assert ones == int(sel.sum()), (dtype, n, op)
                assert (got == np_mask(sel, order)).all(), (dtype, n, op, order)


def test_reference_predicate_case(O):
    """The driver's predicate `value > 200` (src/main.cu:24-28) packed
    MSB-first, 8 elements per byte, by gen_predicate (src/utils.cuh:102-122)."""
    c = np.array([199.0, 200.0, 200.5, 1e9, np.nan, -0.0, 201.0, 0.0, 300.0], np.float32)
    m, ones = O.predicate_mask(c, 200.0)
    assert ones == 4 and list(m) == [0b00110010, 0b10000000]
    t, tones = O.predicate_mask_typed(c, 0, 200.0)
    assert tones == ones and (t == m).all()
    u = np.arange(190, 230, dtype=np.uint32)
    m2, k2 = O.predicate_mask(u, 200)
    t2, tk2 = O.predicate_mask_typed(u, 0, 200)
    assert k2 == tk2 == 29 and (m2 == t2).all()


def test_brute_force_tiny(O):
    """Every 8-element u8 column over {0, 1, 2} against 1, all operators."""
    rng = np.random.default_rng(3)
    for _ in range(300):
        c = rng.integers(0, 3, 8).astype(np.uint8)
        for op in range(6):
            m, k = O.predicate_mask_typed(c, op, 1)
            expect = 0
            for i, v in enumerate(c):
                if [v > 1, v >= 1, v < 1, v <= 1, v == 1, v != 1][op]:
                    expect |= 0x80 >> i
            assert m[0] == expect and k == bin(expect).count("1")


def test_predicate_struct_layout():
    assert ctypes.sizeof(csx.Predicate) == 24
    q = csx.predicate(8, ">", 200.0)
    assert q.type == 8 and q.op == 0 and q.a == 0x43480000
    q = csx.predicate(7, "between", -1, 5)
    assert q.a == 2**64 - 1 and q.b == 5


def test_predicate_abi_rejects_bad_arguments():
    L = csx.lib()
    q = csx.predicate(2, ">", 1)
    sz = ctypes.c_size_t(0)
    assert L.csx_predicate_workspace_size(1000, 2, ctypes.byref(sz)) == 0 and sz.value > 125
    assert L.csx_predicate_workspace_size(1000, 11, ctypes.byref(sz)) == 1
    assert L.csx_predicate_mask(None, 10, None, None, 0, None, None) == 1  # null predicate
    bad = csx.Predicate(2, 9, 0, 0)
    assert L.csx_predicate_mask(None, 10, ctypes.byref(bad), None, 0, None, None) == 1  # unknown op
    assert L.csx_predicate_mask(None, 10, ctypes.byref(q), None, 0, None, None) == 1  # null buffers
    assert L.csx_predicate_mask(None, 10, ctypes.byref(q), None, 7, None, None) == 1  # bad order
    assert L.csx_predicate_mask(ctypes.c_void_p(2), 10, ctypes.byref(q), ctypes.c_void_p(16), 0, None, None) == 1  # misaligned
    ws = ctypes.c_void_p(256)
    assert L.csx_compact_predicate(None, 10, ctypes.byref(q), None, None, 0, None, None, None, 0, None) == 2  # no workspace
    assert L.csx_compact_predicate(None, 10, ctypes.byref(q), None, None, 0, None, None, ws, 1 << 20, None) == 1  # null buffers
