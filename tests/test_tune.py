"""CPU checks of the tuning table: text round trip, nearest-cell lookup,
rejected lines, and configuration checks that fail before any device work."""
import ctypes

import pytest

import gpu_compressstore2_amd as csx


@pytest.fixture(autouse=True)
def clean_table():
    """An empty table for the test; the shipped table again afterwards, so
    that later tests run the tuned configurations whatever the order."""
    csx.tune_clear()
    yield
    csx.tune_clear()
    csx.tune_load(csx.TUNED_TABLE)


def test_default_without_measurements():
    c = csx.tune_lookup(4, 1 << 28, 0.5)
    assert (c.pipeline, c.chunk_len, c.vpt) == (csx.PIPELINE_THREE_PHASE, 1024, 8)


def test_save_load_and_nearest_cell(tmp_path):
    p = tmp_path / "t.txt"
    p.write_text(
        "# a table\n"
        "compact 4 20 0.500000 1 256 2 0.050\n"
        "compact 4 28 0.500000 1 1024 8 0.280\n"
        "compact 4 28 0.990000 2 1024 8 0.380\n"
        "compact 1 28 0.500000 1 4096 4 0.300\n"
        "batch 4 4 2048\n"
    )
    csx.tune_load(str(p))
    c = csx.tune_lookup(4, 1 << 20, 0.5)
    assert (c.pipeline, c.chunk_len, c.vpt) == (1, 256, 2)
    c = csx.tune_lookup(4, 3 << 27, 0.97)  # log2 n = 28, nearest selectivity 0.99
    assert (c.pipeline, c.chunk_len, c.vpt) == (2, 1024, 8)
    c = csx.tune_lookup(4, 1 << 27, -1)  # unknown selectivity: 0.5; nearest size 28
    assert (c.pipeline, c.chunk_len, c.vpt) == (1, 1024, 8)
    c = csx.tune_lookup(1, 1 << 10, 0.1)  # the only cell of width 1
    assert (c.chunk_len, c.vpt) == (4096, 4)
    c = csx.tune_lookup(2, 1 << 28, 0.5)  # nothing for width 2
    assert (c.chunk_len, c.vpt) == (1024, 8)
    q = tmp_path / "u.txt"
    csx.tune_save(str(q))
    text = q.read_text()
    assert text.count("\ncompact ") == 4 and "batch 4 4 2048" in text
    csx.tune_clear()
    csx.tune_load(str(q))
    assert csx.tune_lookup(4, 1 << 20, 0.5).vpt == 2


@pytest.mark.parametrize(
    "line",
    [
        "compact 3 20 0.5 1 1024 8 0.1",  # width
        "compact 4 20 0.5 1 1000 8 0.1",  # chunk not a power of two
        "compact 1 20 0.5 1 1024 16 0.1",  # vpt 16 needs 2-byte elements
        "compact 4 20 0.5 2 1024 4 0.1",  # two-pass uses vpt 8
        "compact 4 20 0.5 4 1024 16 0.1",  # the single launch takes 2, 4 or 8
        "compact 4 20 0.5 4 1024 4 512 0.1",  # ... in 256-thread tiles
        "compact 4 20 1.5 1 1024 8 0.1",  # selectivity
        "batch 4 16 4096",  # batched vpt
        "hello",
    ],
)
def test_bad_lines_rejected(tmp_path, line):
    p = tmp_path / "bad.txt"
    p.write_text(line + "\n")
    with pytest.raises(csx.CsxError):
        csx.tune_load(str(p))
    assert csx.tune_lookup(4, 1 << 20, 0.5).vpt == 8  # nothing half-loaded


def test_block_size_rows_and_version_one_rows(tmp_path):
    """Rows with a workgroup size (format 2) and without (format 1, meaning
    256) load side by side; a save writes the sizes back."""
    p = tmp_path / "t.txt"
    p.write_text("compact 4 20 0.500000 1 1024 2 512 0.02\ncompact 4 24 0.500000 1 256 4 0.03\n"
                 "batch 4 4 8192 128\nbatch 8 8 4096\n")
    csx.tune_load(str(p))
    c = csx.tune_lookup(4, 1 << 20, 0.5)
    assert (c.pipeline, c.chunk_len, c.vpt, c.block_size) == (1, 1024, 2, 512)
    c = csx.tune_lookup(4, 1 << 24, 0.5)
    assert (c.vpt, c.block_size) == (4, 256)
    out = tmp_path / "saved.txt"
    csx.tune_save(str(out))
    text = out.read_text()
    assert "compact 4 20 0.500000 1 1024 2 512 " in text and "compact 4 24 0.500000 1 256 4 256 " in text
    assert "batch 4 4 8192 128" in text and "batch 8 8 4096 256" in text


@pytest.mark.parametrize("line", [
    "compact 4 20 0.5 1 1024 2 384 0.02",  # no such workgroup size
    "compact 4 20 0.5 2 1024 8 512 0.02",  # only the three-phase writeout takes another size
    "batch 4 4 8192 64",
])
def test_bad_block_sizes_rejected(tmp_path, line):
    p = tmp_path / "bad.txt"
    p.write_text(line + "\n")
    with pytest.raises(csx.CsxError):
        csx.tune_load(str(p))


def test_single_launch_rows_take_a_tile_size(tmp_path):
    p = tmp_path / "t.txt"
    p.write_text("compact 4 20 0.500000 4 1024 4 256 0.008\ncompact 2 18 0.500000 4 1024 2 0.006\n")
    csx.tune_load(str(p))
    c = csx.tune_lookup(4, 1 << 20, 0.5)
    assert (c.pipeline, c.vpt, c.block_size) == (csx.PIPELINE_SINGLE, 4, 256)
    assert csx.tune_lookup(2, 1 << 18, 0.5).vpt == 2


def test_config_checks_before_device_work():
    L = csx.lib()
    for pipeline, chunk, vpt, w in [(1, 1024, 3, 4), (1, 1024, 16, 1), (2, 1024, 4, 4), (7, 1024, 8, 4), (1, 100, 8, 4), (4, 1024, 16, 4)]:
        c = csx.Config(pipeline, chunk, vpt, 0)
        assert L.csx_compact_config(None, None, None, 10, w, 0, ctypes.byref(c), None, None, None, 0, None) == 1
    assert L.csx_compact_config(None, None, None, 10, 4, 0, None, None, None, None, 0, None) == 1
    for bad_block in (64, 384, 1024):
        c = csx.Config(1, 1024, 2, 0, bad_block)
        assert L.csx_compact_config(None, None, None, 10, 4, 0, ctypes.byref(c), None, None, None, 0, None) == 1
        assert b"workgroup size" in L.csx_last_error()
    c = csx.Config(csx.PIPELINE_TWO_PASS, 1024, 8, 0, 512)  # only the three-phase writeout takes another size
    assert L.csx_compact_config(None, None, None, 10, 4, 0, ctypes.byref(c), None, None, ctypes.c_void_p(4096), 1 << 20, None) == 1
    c = csx.Config(1, 0, 4, 0, 96)
    assert L.csx_compact_batch_config(None, 1, 10, 4, 0, ctypes.byref(c), None, None, 0, None) == 1
    c = csx.Config(1, 0, 16, 0)
    assert L.csx_compact_batch_config(None, 1, 10, 4, 0, ctypes.byref(c), None, None, 0, None) == 1
    ms = ctypes.c_double()
    assert L.csx_autotune(4, 0, 0.5, 3, ctypes.byref(c), ctypes.byref(ms), None) == 1
    assert L.csx_autotune(4, 100, 1.5, 3, ctypes.byref(c), ctypes.byref(ms), None) == 1
    fake = ctypes.c_void_p(1 << 20)
    # tuning on the caller's column: null buffers, zero reps, bad width, short workspace
    assert L.csx_autotune_inputs(None, None, None, 100, 4, 0, 3, fake, 1 << 20, ctypes.byref(c), ctypes.byref(ms), None, None) == 1
    assert L.csx_autotune_inputs(fake, fake, fake, 100, 4, 0, 0, fake, 1 << 20, ctypes.byref(c), ctypes.byref(ms), None, None) == 1
    assert L.csx_autotune_inputs(fake, fake, fake, 100, 3, 0, 3, fake, 1 << 20, ctypes.byref(c), ctypes.byref(ms), None, None) == 1
    assert L.csx_autotune_inputs(fake, fake, fake, 1 << 20, 4, 0, 3, fake, 64, ctypes.byref(c), ctypes.byref(ms), None, None) == 2
    assert csx.auto_workspace_size(1 << 20, 4) == csx.workspace_size(1 << 20, 4, 256)


def test_tile_grid_bound_checked_for_the_chosen_tile():
    """2^38 one-byte elements fit one launch with 8-vector tiles but not with
    2-vector tiles (2^25 workgroups of 256 threads); the call refuses before
    any device work."""
    L = csx.lib()
    c = csx.Config(csx.PIPELINE_THREE_PHASE, 1 << 17, 2, 0)
    # disjoint fake ranges (2^38 input bytes, 2^35 mask bytes, 2^38 output bytes), never dereferenced
    fin, fmask, fout, fws = (ctypes.c_void_p(a) for a in (1 << 40, 1 << 41, 1 << 42, 1 << 20))
    st = L.csx_compact_config(fin, fmask, fout, 1 << 38, 1, 0, ctypes.byref(c), None, None, fws, 2**63, None)
    assert st == 4 and "exceed one launch" in L.csx_last_error().decode()
