"""CPU oracle for bitmask stream compaction -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker (or, for the baseline,
as the CPU path timed beside the GPU).  The product library never uses it.

``Oracle`` wraps ``oracle/_build/libcsx_oracle.so`` (the C restatement in
``csx_oracle.c``); ``Reference`` wraps ``oracle/_ref/libcsx_ref.so`` (the
reference's own CPU code, built by ``make -C oracle ref`` where
RememberOfLife/gpu_compressstore2 is present).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CSX_ORACLE_LIBRARY names another build of the restatement (the sanitized one)
ORACLE_SO = os.environ.get("CSX_ORACLE_LIBRARY") or os.path.join(HERE, "_build", "libcsx_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcsx_ref.so")

MSB_FIRST = 0
LSB_FIRST = 1

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32


def build(ref_root="/root/reference"):
    """Compile the C restatement, and the reference CPU code when present."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir(os.path.join(ref_root, "src")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF_ROOT={ref_root}"], check=True)


class ctypes_prng(ctypes.Structure):
    """Layout of csxo_prng."""

    _fields_ = [("state", ctypes.c_uint64), ("inc", ctypes.c_uint64)]


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class Oracle:
    """ctypes view of the C restatement."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        self.lib = L
        sig = {
            "csxo_prng_seq": (None, [_u64, _vp, ctypes.c_size_t]),
            "csxo_prng_seed": (None, [_vp, _u64]),
            "csxo_prng_next": (_u32, [_vp]),
            "csxo_prng_advance": (None, [_vp, _u64]),
            "csxo_clear_tail": (None, [_vp, _u64, ctypes.c_int]),
            "csxo_mask_uniform": (None, [_vp, _u64, _u64, ctypes.c_double, _u64, _vp]),
            "csxo_mask_zipf": (None, [_vp, _u64, _u64, _u64, _vp]),
            "csxo_mask_burst": (None, [_vp, _u64, _u64, _u64, ctypes.c_double]),
            "csxo_mask_offset": (None, [_vp, _u64, _u64, ctypes.c_int64]),
            "csxo_mask_pattern": (None, [_vp, _u64, _u64, _u32, _u32, _vp]),
            "csxo_mask_clustering": (None, [_vp, ctypes.c_float, _u64, _u64, _vp]),
            "csxo_pattern_from_selectivity": (_u32, [ctypes.c_int, ctypes.c_float, _u64]),
            "csxo_mask_bernoulli": (None, [_vp, _u64, ctypes.c_double, _u64, ctypes.c_int]),
            "csxo_mask_geometric_runs": (None, [_vp, _u64, _u32, _u64, ctypes.c_int]),
            "csxo_zipf_thresholds": (None, [ctypes.c_double, _u32, _vp]),
            "csxo_mask_zipf_blocks": (None, [_vp, _u64, _u32, ctypes.c_double, _u32, _u64, ctypes.c_int]),
            "csxo_mask_periodic": (None, [_vp, _u64, _u32, ctypes.c_int]),
            "csxo_fill_values": (None, [_vp, _u64, _u32, ctypes.c_int, _u64]),
            "csxo_predicate_mask_f32": (_u64, [_vp, _u64, ctypes.c_float, _vp, ctypes.c_int]),
            "csxo_predicate_mask_u32": (_u64, [_vp, _u64, _u32, _vp, ctypes.c_int]),
            "csxo_predicate_mask_typed": (_u64, [_vp, _u64, ctypes.c_int, ctypes.c_int, _u64, _u64, _vp, ctypes.c_int]),
            "csxo_generate_validation": (_u64, [_vp, _vp, _vp, _u64, _u32]),
            "csxo_compact": (_u64, [_vp, _vp, _vp, _u64, _u32, ctypes.c_int]),
            "csxo_popc": (None, [_vp, _vp, _u64, _u32, ctypes.c_int]),
            "csxo_exclusive_scan": (_u64, [_vp, _vp, _u64]),
            "csxo_upsweep_tree": (None, [_vp, _u64]),
            "csxo_tree_prefix": (_u32, [_vp, _u64, _u64]),
            "csxo_mask_count": (_u64, [_vp, _u64, ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args

    # ---- generators -------------------------------------------------------
    def prng_seq(self, seed, n):
        out = np.zeros(n, np.uint32)
        self.lib.csxo_prng_seq(seed, _ptr(out), n)
        return out

    def mask_uniform(self, nbytes, selectivity, seed=42, offset=0):
        buf = np.zeros(offset + nbytes, np.uint8)
        cnt = np.zeros(1, np.uint64)
        self.lib.csxo_mask_uniform(_ptr(buf), offset, nbytes, selectivity, seed, _ptr(cnt))
        return buf, int(cnt[0])

    def mask_zipf(self, nbytes):
        buf = np.zeros(nbytes, np.uint8)
        cnt = np.zeros(1, np.uint64)
        self.lib.csxo_mask_zipf(_ptr(buf), nbytes, 0, nbytes, _ptr(cnt))
        return buf, int(cnt[0])

    def mask_burst(self, nbytes, segment_sizer):
        buf = np.zeros(nbytes, np.uint8)
        self.lib.csxo_mask_burst(_ptr(buf), nbytes, 0, nbytes, segment_sizer)
        return buf

    def mask_offset(self, nbytes, spacing):
        buf = np.zeros(nbytes, np.uint8)
        self.lib.csxo_mask_offset(_ptr(buf), 0, nbytes, spacing)
        return buf

    def mask_pattern(self, nbytes, pattern, pattern_length):
        buf = np.zeros(nbytes, np.uint8)
        cnt = np.zeros(1, np.uint64)
        self.lib.csxo_mask_pattern(_ptr(buf), 0, nbytes, pattern, pattern_length, _ptr(cnt))
        return buf, int(cnt[0])

    def mask_clustering(self, selectivity, cluster_count, total_elements):
        buf = np.zeros(total_elements // 8, np.uint8)
        cnt = np.zeros(1, np.uint64)
        self.lib.csxo_mask_clustering(_ptr(buf), selectivity, cluster_count, total_elements, _ptr(cnt))
        return buf, int(cnt[0])

    def pattern_from_selectivity(self, pattern_length, selectivity, seed=42):
        return int(self.lib.csxo_pattern_from_selectivity(pattern_length, selectivity, seed))

    def mask_bernoulli(self, n, p, seed, bit_order=MSB_FIRST):
        buf = np.zeros((n + 7) // 8, np.uint8)
        self.lib.csxo_mask_bernoulli(_ptr(buf), n, p, seed, bit_order)
        return buf

    def mask_geometric_runs(self, n, mean_run, seed, bit_order=MSB_FIRST):
        buf = np.zeros((n + 7) // 8, np.uint8)
        self.lib.csxo_mask_geometric_runs(_ptr(buf), n, mean_run, seed, bit_order)
        return buf

    def zipf_thresholds(self, s, levels):
        out = np.zeros(levels, np.uint32)
        self.lib.csxo_zipf_thresholds(s, levels, _ptr(out))
        return out

    def mask_zipf_blocks(self, n, block_len, s, levels, seed, bit_order=MSB_FIRST):
        buf = np.zeros((n + 7) // 8, np.uint8)
        self.lib.csxo_mask_zipf_blocks(_ptr(buf), n, block_len, s, levels, seed, bit_order)
        return buf

    def mask_periodic(self, n, pattern, bit_order=MSB_FIRST):
        buf = np.zeros((n + 7) // 8, np.uint8)
        self.lib.csxo_mask_periodic(_ptr(buf), n, pattern, bit_order)
        return buf

    def fill_values(self, n, dtype, seed, kind=None):
        dt = np.dtype(dtype)
        if kind is None:
            kind = 1 if dt.kind == "f" else 0
        out = np.zeros(n, dt)
        self.lib.csxo_fill_values(_ptr(out), n, dt.itemsize, kind, seed)
        return out

    def predicate_mask(self, col, threshold, bit_order=MSB_FIRST):
        col = np.ascontiguousarray(col)
        n = col.size
        mask = np.zeros((n + 7) // 8, np.uint8)
        if col.dtype == np.float32:
            ones = self.lib.csxo_predicate_mask_f32(_ptr(col), n, float(threshold), _ptr(mask), bit_order)
        elif col.dtype == np.uint32:
            ones = self.lib.csxo_predicate_mask_u32(_ptr(col), n, int(threshold), _ptr(mask), bit_order)
        else:
            raise TypeError(col.dtype)
        return mask, int(ones)

    def predicate_mask_typed(self, col, op, a, b=0, bit_order=MSB_FIRST):
        """Mask of `col OP a` (op index as csx_pred_op; 6 = a <= v <= b)."""
        col = np.ascontiguousarray(col)
        n = col.size
        mask = np.zeros((n + 7) // 8, np.uint8)
        ones = self.lib.csxo_predicate_mask_typed(_ptr(col), n, PRED_TYPES[col.dtype], op, scalar_bits(a, col.dtype), scalar_bits(b, col.dtype),
                                                  _ptr(mask), bit_order)
        if ones == 2**64 - 1:
            raise ValueError(f"bad predicate op {op}")
        return mask, int(ones)

    # ---- compaction path --------------------------------------------------
    def generate_validation(self, data, mask):
        n = data.size
        out = np.zeros(n, data.dtype)
        cnt = self.lib.csxo_generate_validation(_ptr(data), _ptr(mask), _ptr(out), n, data.dtype.itemsize)
        return out, int(cnt)

    def compact(self, data, mask, n=None, bit_order=MSB_FIRST):
        data = np.ascontiguousarray(data)
        n = data.size if n is None else n
        out = np.zeros(max(n, 1), data.dtype)
        cnt = self.lib.csxo_compact(_ptr(data), _ptr(mask), _ptr(out), n, data.dtype.itemsize, bit_order)
        return out[:cnt].copy(), int(cnt)

    def popc(self, mask, n, chunk_len, bit_order=MSB_FIRST):
        chunks = (n + chunk_len - 1) // chunk_len
        out = np.zeros(max(chunks, 1), np.uint32)
        self.lib.csxo_popc(_ptr(mask), _ptr(out), n, chunk_len, bit_order)
        return out[:chunks]

    def exclusive_scan(self, counts):
        counts = np.ascontiguousarray(counts, np.uint32)
        out = np.zeros(max(counts.size, 1), np.uint64)
        total = self.lib.csxo_exclusive_scan(_ptr(counts), _ptr(out), counts.size)
        return out[: counts.size], int(total)

    def upsweep_tree(self, counts):
        tree = np.array(counts, np.uint32, copy=True)
        self.lib.csxo_upsweep_tree(_ptr(tree), tree.size)
        return tree

    def tree_prefix(self, tree, idx):
        tree = np.ascontiguousarray(tree, np.uint32)
        return int(self.lib.csxo_tree_prefix(_ptr(tree), idx, tree.size))

    def mask_count(self, mask, n, bit_order=MSB_FIRST):
        return int(self.lib.csxo_mask_count(_ptr(mask), n, bit_order))


class Reference:
    """ctypes view of the reference's own CPU code (oracle/_ref)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref`")
        L = ctypes.CDLL(path)
        self.lib = L
        sig = {
            "ref_prng_seq": (None, [_u64, _vp, ctypes.c_size_t]),
            "ref_mask_uniform": (None, [_vp, _u64, _u64, ctypes.c_double, _vp]),
            "ref_mask_zipf": (None, [_vp, _u64, _u64, _u64, _vp]),
            "ref_mask_burst": (None, [_vp, _u64, _u64, _u64, ctypes.c_double]),
            "ref_mask_offset": (None, [_vp, _u64, _u64, ctypes.c_int64]),
            "ref_mask_pattern": (None, [_vp, _u64, _u64, _u32, _u32, _vp]),
            "ref_mask_clustering": (_u64, [_vp, ctypes.c_float, _u64, _u64, _vp]),
            "ref_generate_validation": (_u64, [_vp, _vp, _vp, _u64, _u32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args

    def prng_seq(self, seed, n):
        out = np.zeros(n, np.uint32)
        self.lib.ref_prng_seq(seed, _ptr(out), n)
        return out

    def mask_uniform(self, nbytes, selectivity, offset=0):
        buf = np.zeros(offset + nbytes, np.uint8)
        cnt = np.zeros(1, np.uint64)
        self.lib.ref_mask_uniform(_ptr(buf), offset, nbytes, selectivity, _ptr(cnt))
        return buf, int(cnt[0])

    def mask_zipf(self, nbytes):
        buf = np.zeros(nbytes, np.uint8)
        cnt = np.zeros(1, np.uint64)
        self.lib.ref_mask_zipf(_ptr(buf), nbytes, 0, nbytes, _ptr(cnt))
        return buf, int(cnt[0])

    def mask_burst(self, nbytes, segment_sizer):
        buf = np.zeros(nbytes, np.uint8)
        self.lib.ref_mask_burst(_ptr(buf), nbytes, 0, nbytes, segment_sizer)
        return buf

    def mask_offset(self, nbytes, spacing):
        buf = np.zeros(nbytes, np.uint8)
        self.lib.ref_mask_offset(_ptr(buf), 0, nbytes, spacing)
        return buf

    def mask_pattern(self, nbytes, pattern, pattern_length):
        buf = np.zeros(nbytes, np.uint8)
        cnt = np.zeros(1, np.uint64)
        self.lib.ref_mask_pattern(_ptr(buf), 0, nbytes, pattern, pattern_length, _ptr(cnt))
        return buf, int(cnt[0])

    def mask_clustering(self, selectivity, cluster_count, total_elements):
        buf = np.zeros(max(total_elements // 8, 1), np.uint8)
        cnt = np.zeros(1, np.uint64)
        nb = self.lib.ref_mask_clustering(_ptr(buf), selectivity, cluster_count, total_elements, _ptr(cnt))
        return buf[:nb], int(cnt[0])

    def generate_validation(self, data, mask):
        data = np.ascontiguousarray(data)
        n = data.size
        out = np.zeros(n, data.dtype)
        m = np.array(mask, np.uint8, copy=True)
        cnt = self.lib.ref_generate_validation(_ptr(data), _ptr(m), _ptr(out), n, data.dtype.itemsize)
        return out, int(cnt)


# numpy dtype -> csx_type index
PRED_TYPES = {np.dtype(t): i for i, t in enumerate([np.uint8, np.uint16, np.uint32, np.uint64, np.int8, np.int16, np.int32, np.int64,
                                                    np.float32, np.float64])}


def scalar_bits(v, dtype):
    """A value of `dtype` as the little-endian low bytes of a uint64."""
    raw = np.array([v], dtype=dtype).view(np.uint8)
    return int.from_bytes(bytes(raw) + bytes(8 - raw.size), "little")


def reference_available():
    return os.path.exists(REF_SO)


def numpy_compact(data, mask, n=None, bit_order=MSB_FIRST):
    """Independent check through numpy's own bit unpacking and boolean indexing."""
    n = data.size if n is None else n
    bits = np.unpackbits(np.asarray(mask, np.uint8), bitorder="big" if bit_order == MSB_FIRST else "little")[:n]
    sel = bits.astype(bool)
    return np.ascontiguousarray(data[:n][sel])
