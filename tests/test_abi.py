"""CPU checks of the C ABI: the library loads, exports every symbol that
include/csx.h declares, and rejects bad arguments before touching the GPU."""
import ctypes
import os
import re

import pytest

import gpu_compressstore2_amd as csx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if hdr.endswith(".h"):
            text = open(os.path.join(ROOT, "include", hdr)).read()
            names |= set(re.findall(r"^\s*(?:csx_status|const char\*|int|uint64_t)\s+(csx_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("csx_popc", "csx_scan", "csx_writeout", "csx_compact", "csx_workspace_size", "csx_validate"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = csx.lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert set(declared_symbols()) <= set(csx.SIGNATURES)


def test_status_strings_and_version():
    L = csx.lib()
    assert L.csx_status_string(0) == b"ok"
    assert L.csx_status_string(2).startswith(b"workspace")
    assert L.csx_version() >= 100
    assert L.csx_chunk_count(1000, 32) == 32
    assert L.csx_chunk_count(0, 1024) == 0


@pytest.mark.parametrize("chunk", [0, 16, 31, 33, 100, 1000])
def test_bad_chunk_length_rejected(chunk):
    L = csx.lib()
    sz = ctypes.c_size_t(0)
    assert L.csx_workspace_size(1000, 4, chunk, ctypes.byref(sz)) == 1
    assert b"power of two" in L.csx_last_error()
    assert L.csx_popc(None, None, 1000, chunk, 0, None) == 1


@pytest.mark.parametrize("w", [0, 3, 5, 16])
def test_bad_width_rejected(w):
    sz = ctypes.c_size_t(0)
    assert csx.lib().csx_workspace_size(1000, w, 1024, ctypes.byref(sz)) == 1


def test_bad_bit_order_and_variant_rejected():
    L = csx.lib()
    assert L.csx_popc(None, None, 10, 32, 2, None) == 1
    assert L.csx_writeout(None, None, None, None, None, 10, 4, 32, 0, 7, None) == 1


def test_too_many_elements_rejected():
    sz = ctypes.c_size_t(0)
    assert csx.lib().csx_workspace_size(csx.MAX_ELEMENTS + 1, 4, 1024, ctypes.byref(sz)) == 1


def test_workspace_checks():
    L = csx.lib()
    need = csx.workspace_size(1 << 20, 4, 1024)
    assert need > 0 and need % 256 == 0
    # missing workspace
    assert L.csx_compact(None, None, None, 1 << 20, 4, 1024, 0, 0, None, None, None, 0, None) == 2
    # too small (a fake, never dereferenced pointer: the size check comes first)
    assert L.csx_compact(None, None, None, 1 << 20, 4, 1024, 0, 0, None, None, ctypes.c_void_p(4096), need - 1, None) == 2
    assert b"smaller" in L.csx_last_error()
    # misaligned
    assert L.csx_compact(None, None, None, 1 << 20, 4, 1024, 0, 0, None, None, ctypes.c_void_p(4097), need, None) == 2


def test_noalloc_rejects_oversized_request_before_device_work():
    """csx_compact_noalloc checks the request against its static scratch
    before it looks the scratch up on the device (fake pointers, never
    dereferenced)."""
    L = csx.lib()
    assert csx.workspace_size(1 << 30, 4, 32) > csx.NOALLOC_SCRATCH_BYTES
    cfg = csx.Config(csx.PIPELINE_THREE_PHASE, 32, 8, 0)
    fake = ctypes.c_void_p(4096)
    assert L.csx_compact_noalloc(fake, fake, ctypes.c_void_p(1 << 40), 1 << 30, 4, 0, ctypes.byref(cfg), None, None, None) == 4
    assert b"static scratch" in L.csx_last_error()
    assert L.csx_compact_noalloc(fake, fake, fake, 100, 4, 0, None, None, None, None) == 1
    bad = csx.Config(csx.PIPELINE_THREE_PHASE, 100, 8, 0)
    assert L.csx_compact_noalloc(fake, fake, fake, 100, 4, 0, ctypes.byref(bad), None, None, None) == 1


def test_null_data_pointers_rejected():
    L = csx.lib()
    need = csx.workspace_size(100, 4, 1024)
    assert L.csx_compact(None, None, None, 100, 4, 1024, 0, 0, None, None, ctypes.c_void_p(4096), need, None) == 1
    assert L.csx_writeout(None, None, None, None, None, 100, 4, 1024, 0, 0, None) == 1


def test_unknown_pipeline_rejected():
    need = csx.workspace_size(100, 4, 1024)
    assert csx.lib().csx_compact(None, None, None, 100, 4, 1024, 0, 9, None, None, ctypes.c_void_p(4096), need, None) == 1
    # the highest known pipeline (the single launch) passes the pipeline check and stops at the null pointers;
    # one past it is unknown
    L = csx.lib()
    assert L.csx_compact(None, None, None, 100, 4, 1024, 0, csx.PIPELINE_SINGLE, None, None, ctypes.c_void_p(4096), need, None) == 1
    assert b"null" in L.csx_last_error()
    assert L.csx_compact(None, None, None, 100, 4, 1024, 0, csx.PIPELINE_SINGLE + 1, None, None, ctypes.c_void_p(4096), need, None) == 1
    assert b"unknown pipeline" in L.csx_last_error()


def test_single_launch_cap_rejected_before_device_work():
    L = csx.lib()
    n = csx.SINGLE_MAX_ELEMENTS + 1
    need = csx.workspace_size(n, 4, 1024)
    fake = ctypes.c_void_p(4096)
    assert L.csx_compact(fake, fake, ctypes.c_void_p(1 << 44), n, 4, 1024, 0, csx.PIPELINE_SINGLE, None, None, fake, need, None) == 4
    assert b"single-launch" in L.csx_last_error()


def test_grid_limit_counts_the_workgroup_size():
    """A writeout grid must hold fewer than 2^32 work-items.  2^35 u32
    elements in 2-vector tiles of 512 threads (4096 elements) make 2^23
    tiles, 2^32 work-items: refused before any launch, although 2^23
    workgroups of 256 threads would fit (fake addresses, never
    dereferenced)."""
    L = csx.lib()
    n, w = 1 << 35, 4
    V = ctypes.c_void_p
    data, mask, out, ws = 1 << 44, 1 << 46, 1 << 47, 1 << 43
    need = csx.workspace_size(n, w, 1024)
    cfg = csx.Config(csx.PIPELINE_THREE_PHASE, 1024, 2, 0, 512)
    st = L.csx_compact_config(V(data), V(mask), V(out), n, w, 0, ctypes.byref(cfg), None, None, V(ws), need, None)
    assert st == 4 and b"exceed one launch" in L.csx_last_error()


def test_in_place_calls_rejected_before_device_work():
    """Compaction is not in place: an output range that shares a byte with
    the input or the mask is refused with CSX_ERR_INVALID before any launch
    (addresses only, nothing is dereferenced)."""
    L = csx.lib()
    n, w = 1000, 4
    base = 1 << 30
    data, mask, ws = base, base + (1 << 24), base + (1 << 26)
    V = ctypes.c_void_p
    for out in (data, data + 4 * (n - 1), data - 4 * (n - 1), mask - 4, mask + 100):
        st = L.csx_compact(V(data), V(mask), V(out), n, w, 1024, 0, csx.PIPELINE_THREE_PHASE, None, None, V(ws), 1 << 30, None)
        assert st == 1 and b"overlaps" in L.csx_last_error(), hex(out - base)
        wa = csx.WriteoutArgs(V(data), V(mask), V(out), V(ws), 0, None, None, n, w, 1024, 0, csx.WRITEOUT_STAGED)
        assert L.csx_writeout_ex(ctypes.byref(wa), None) == 1 and b"overlaps" in L.csx_last_error()
    st = L.csx_pattern_compact(V(data), V(data + 8), n, w, 0xAAAAAAAA, 32, None, None, None)
    assert st == 1 and b"overlaps" in L.csx_last_error()
    u32 = csx.TYPE_NAMES.index("uint32")
    q = csx.predicate(u32, csx.PRED_GT, 5)
    need = csx.predicate_workspace_size(n, u32)
    for out, m in ((data + 64, None), (data + (1 << 20), V(data + 100)), (data + (1 << 20), V(data + (1 << 20) + 8))):
        st = L.csx_compact_predicate(V(data), n, ctypes.byref(q), V(out), m, 0, None, None, V(ws), need, None)
        assert st == 1 and b"must not overlap" in L.csx_last_error()
    assert L.csx_predicate_mask(V(data), n, ctypes.byref(q), V(data + 12), 0, None, None) == 1
    assert b"overlaps" in L.csx_last_error()


def test_writeout_variant_range():
    L = csx.lib()
    # the sparse variant is known: rejected for the null pointers, not the variant
    assert L.csx_writeout(None, None, None, None, None, 10, 4, 32, 0, csx.WRITEOUT_SPARSE, None) == 1
    assert b"null" in L.csx_last_error()
    assert L.csx_writeout(None, None, None, None, None, 10, 4, 32, 0, csx.WRITEOUT_SPARSE + 1, None) == 1
    assert b"unknown writeout variant" in L.csx_last_error()


def test_tuning_table_accepts_sparse_and_rejects_bad_rows(tmp_path):
    good = tmp_path / "t.txt"
    good.write_text("# csx tuning table 1\ncompact 4 20 0.010000 3 4096 8 0.01\n")
    csx.tune_clear()
    try:
        csx.tune_load(str(good))
        c = csx.tune_lookup(4, 1 << 20, 0.01)
        assert (c.pipeline, c.chunk_len, c.vpt) == (csx.PIPELINE_SPARSE, 4096, 8)
        for bad in ("compact 4 20 0.01 3 4096 2 0.01\n",  # sparse takes no tile size but 8
                    "compact 4 20 0.01 5 1024 8 0.01\n",  # unknown pipeline
                    "compact 4 25 0.01 4 1024 8 0.01\n",  # the single launch past its 2^24 cap
                    "compact 4 20 1.5 1 1024 8 0.01\n"):  # selectivity out of range
            f = tmp_path / "bad.txt"
            f.write_text(bad)
            with pytest.raises(csx.CsxError):
                csx.tune_load(str(f))
    finally:
        csx.tune_clear()
        csx.tune_load(csx.TUNED_TABLE)


def test_single_launch_cells_never_serve_columns_past_the_cap(tmp_path):
    f = tmp_path / "t.txt"
    f.write_text("compact 4 24 0.500000 4 1024 8 0.005\ncompact 4 28 0.500000 1 1024 2 0.27\n")
    csx.tune_clear()
    try:
        csx.tune_load(str(f))
        assert csx.tune_lookup(4, 1 << 24, 0.5).pipeline == csx.PIPELINE_SINGLE
        # 2^25 is nearest to the 2^24 cell, which it may not take
        c = csx.tune_lookup(4, 1 << 25, 0.5)
        assert (c.pipeline, c.vpt) == (csx.PIPELINE_THREE_PHASE, 2)
    finally:
        csx.tune_clear()
        csx.tune_load(csx.TUNED_TABLE)


def test_tuning_table_parse_is_strict_and_all_or_nothing(tmp_path):
    f = tmp_path / "t.txt"
    csx.tune_clear()
    try:
        # indented comments, CRLF endings, blank lines and a last row without a newline load
        f.write_bytes(b"  # indented comment\r\n\r\n\tcompact 4 20 0.500000 1 256 4 0.02  \r\nbatch 4 4 2048")
        csx.tune_load(str(f))
        c = csx.tune_lookup(4, 1 << 20, 0.5)
        assert (c.pipeline, c.chunk_len, c.vpt) == (csx.PIPELINE_THREE_PHASE, 256, 4)
        ok_row = "compact 4 20 0.500000 0 4096 8 0.02\n"
        for bad in ("compact 4 20 0.5 1 256 4 0.02 junk\n",  # trailing text after the row
                    "compact 4 20 0.5 1 256 4 nan\n",  # time not a finite non-negative number
                    "compact 4 20 0.5 1 256 4 -0.5\n",
                    "compact 4 20 0.5 1 256 4 inf\n",
                    "batch 4 4 2048 9\n",  # one field too many
                    "compact 4 20 0.5 1 256 4 0.02" + " " * 600 + "compact 4 20 0.5 1 256 4 0.02\n"):  # over-long line
            f.write_text(ok_row + bad)
            with pytest.raises(csx.CsxError):
                csx.tune_load(str(f))
            # the good first row of a rejected file is not taken either
            c = csx.tune_lookup(4, 1 << 20, 0.5)
            assert (c.pipeline, c.chunk_len, c.vpt) == (csx.PIPELINE_THREE_PHASE, 256, 4)
    finally:
        csx.tune_clear()
        csx.tune_load(csx.TUNED_TABLE)


def test_workspace_size_grows_with_n_and_shrinking_chunks():
    a = csx.workspace_size(1 << 24, 4, 1024)
    b = csx.workspace_size(1 << 26, 4, 1024)
    c = csx.workspace_size(1 << 24, 4, 32)
    assert b > a and c > a
    assert csx.scan_workspace_size(0) >= 256


def test_error_raises_python_exception():
    with pytest.raises(csx.CsxError) as e:
        csx.workspace_size(10, 3, 1024)
    assert e.value.status == 1


def test_zipf_table_matches_oracle(O):
    for s, levels in ((1.2, 65), (2.0, 17), (0.5, 128), (1.0, 2)):
        cdf, thr = csx.zipf_table(s, levels)
        assert (cdf == O.zipf_thresholds(s, levels)).all()
        assert thr[0] == 0 and thr[-1] == 0xFFFFFFFF


def test_pattern_count_matches_oracle(O):
    for pat, L in ((0xB5000000, 8), (0xDEADBEEF, 32), (0x6C000000, 7), (0x0, 5)):
        for n in (0, 1, 7, 8, 1000, 12345):
            mask, _ = O.mask_pattern((n + 7) // 8, pat & ((0xFFFFFFFF << (32 - L)) & 0xFFFFFFFF), L)
            assert csx.pattern_count(n, pat, L) == O.mask_count(mask, n)
    with pytest.raises(csx.CsxError):
        csx.pattern_count(10, 1, 33)


def test_server_and_arena_argument_errors():
    """Serving entry points reject bad arguments with a message, before any device work."""
    import gpu_compressstore2_amd.serving as sv

    L = csx.lib()
    h = ctypes.c_void_p()
    cfg = sv.ServerConfig(0, 8, 1 << 20)
    assert L.csx_server_create(ctypes.c_void_p(ctypes.addressof(cfg)), ctypes.byref(h)) == 1
    assert "streams" in L.csx_last_error().decode()
    cfg = sv.ServerConfig(17, 8, 1 << 20)
    assert L.csx_server_create(ctypes.c_void_p(ctypes.addressof(cfg)), ctypes.byref(h)) == 1
    assert L.csx_server_create(None, ctypes.byref(h)) == 1
    assert L.csx_arena_create(0, ctypes.byref(h)) == 1 and "zero" in L.csx_last_error().decode()
    assert L.csx_arena_alloc(None, 16, ctypes.byref(h)) == 1
    assert L.csx_arena_free(None, None) == 1
    assert L.csx_arena_used(None) == 0
    assert L.csx_server_submit(None, None) == 1
    assert L.csx_server_step(None, None) == 1
    n = ctypes.c_uint32()
    assert L.csx_server_poll(None, None, 0, ctypes.byref(n)) == 1
    assert L.csx_server_queued(None) == 0 and L.csx_server_inflight(None) == 0


def test_last_launch_reports_nothing_on_a_fresh_thread():
    """csx_last_launch reads the calling thread's record: (0, 0) before any
    launch on that thread; null arguments refused (no device work)."""
    import threading
    got = []
    t = threading.Thread(target=lambda: got.append(csx.last_launch()))
    t.start()
    t.join()
    assert got == [(0, 0)]
    assert csx.lib().csx_last_launch(None, None) == 1
