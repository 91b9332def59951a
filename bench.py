"""Benchmark: single-column bitmask compaction on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]     the benchmark line
    python bench.py --only headline                          its headline alone
    python bench.py --sweep phases|widths|shapes|chunks|sizes|ceiling

Workload: 2^28 four-byte elements (float32 uniform in [0,1), moved as opaque
32-bit words) with packed MSB-first Bernoulli masks at 1 %, 50 % and 99 %
selectivity.  A step is one compaction call over one input set, inputs already
resident in HBM; three input sets rotate so the 256 MB Infinity Cache cannot
serve a step from the previous one.

Metric (effective GB/s) = (data bytes read + mask bytes read + selected bytes
written) / time = (4N + N/8 + 4k) / t.  `value` is the 50 % point of
csx_compact_auto (the configuration the autotuner measured for this width,
size and selectivity: the shipped tuning table), measured first and its output
compared with torch's masked_select.  Every selectivity's three-phase,
two-pass, fused, sparse and tuned forms, rocPRIM's select, the batched
4096-column workload, the predicate path, the serving loop and a plain device
copy are reported under "sweep" (low-selectivity entries also with the bytes
physically moved).  "roofline" prices the dominant kernel (the staged
writeout) against the 8 TB/s HBM peak with its duration from a HIP event pair
launched with the kernel on its stream (hipExtLaunchKernelGGL), so it reads
the same kernel time as rocprofv3's kernel trace; "cpu_baseline" times the
reference's CPU compaction (generate_validation, compiled from its sources
into oracle/_ref) on the host.
"""
import argparse
import subprocess
import json
import os
import platform
import sys
import time

import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import gpu_compressstore2_amd as csx  # noqa: E402
from gpu_compressstore2_amd import baseline  # noqa: E402

METRIC = "effective GB/s (data+mask+out) & % HBM peak, 2^28 u32 at 1/50/99% sel"
# counter summary (HBM bytes per launch of the headline's writeout kernel, calibrated
# against csx_copy) of two `rocprofv3 --pmc` runs of `python bench.py --only headline`
TRAFFIC_FILE = "profiles/traffic_headline_2p28_u32_p50.json"
# kernel-trace summary of `python bench.py --only headline` (the writeout's average duration)
PROFILE_FILE = "profiles/headline_2p28_u32_p50.md"
HBM_PEAK_GBS = 8000.0
ROTATE = 3


def pack_msb_first(bits):
    """bool[n] (n % 8 == 0) -> uint8[n/8], element i at bit 7-(i%8)."""
    w = torch.tensor([128, 64, 32, 16, 8, 4, 2, 1], dtype=torch.uint8, device=bits.device)
    return (bits.view(-1, 8).to(torch.uint8) * w).sum(1, dtype=torch.uint8)


def unpack_msb_first(mask, n):
    sh = torch.arange(7, -1, -1, dtype=torch.uint8, device=mask.device)
    return ((mask.view(-1, 1) >> sh) & 1).view(-1)[:n].bool()


def make_mask(n, p, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    bits = torch.rand(n, device="cuda", generator=g) < p
    mask = pack_msb_first(bits)
    k = int(bits.sum().item())
    del bits
    return mask, k


def make_data(n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.rand(n, device="cuda", generator=g, dtype=torch.float32)


_WARM = []


def device_warm(ms=20.0):
    """Sustained untimed load right before a timed window: the GPU needs ~10 ms
    of load to reach its steady clocks, more than a few warmup steps give (the
    first window after them measured ~10 % slow).  Copies between two scratch
    buffers of its own, so no bench input is left in the Infinity Cache."""
    if not _WARM:
        _WARM.extend(torch.empty(1 << 27, dtype=torch.int32, device="cuda") for _ in range(2))
    for _ in range(int(ms / 0.17)):  # one 512 MiB copy takes ~0.17 ms
        csx.copy(_WARM[1], _WARM[0])


def time_steps(fn, steps, warmup):
    for i in range(warmup):
        fn(i)
    device_warm()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for i in range(steps):
        fn(i)
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b)


def bench_batch(steps, warmup, seed=2024, nseg=4096, sparse_share=0.0, batched_only=False):
    """4096 u32 columns, lengths log-uniform in [2^10, 2^24], per-column
    Bernoulli selectivity uniform in [0.01, 0.99] (with sparse_share, that
    share of the columns at selectivity uniform in [0.001, 0.01] instead):
    one batched call against one csx_compact call per column (both eager, and
    both graph-replayed; batched_only: the batched call alone)."""
    import ctypes

    import numpy as np

    rng = np.random.default_rng(seed)
    lengths = np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 24), nseg)).astype(np.int64)
    sels = rng.uniform(0.01, 0.99, nseg)
    if sparse_share > 0:
        pick = rng.random(nseg) < sparse_share
        sels[pick] = rng.uniform(0.001, 0.01, int(pick.sum()))
    starts = np.zeros(nseg, np.int64)
    mstarts = np.zeros(nseg, np.int64)
    pos = mpos = 0
    for i, n in enumerate(lengths):
        starts[i], mstarts[i] = pos, mpos
        pos += (int(n) + 63) // 64 * 64           # 256-byte aligned column starts
        mpos += ((int(n) + 7) // 8 + 255) // 256 * 256
    total = int(lengths.sum())
    data = csx.gen_values(pos, 4, csx.VALUES_INTEGER, seed=seed)
    out = torch.empty(pos, dtype=torch.int32, device="cuda")
    masks = torch.empty(mpos, dtype=torch.uint8, device="cuda")
    ins, ms, outs = [], [], []
    for i in range(nseg):
        n = int(lengths[i])
        mv = masks[mstarts[i]: mstarts[i] + (n + 7) // 8]
        csx.gen_mask_bernoulli(n, float(sels[i]), seed=seed + i, out=mv)
        ins.append(data[starts[i]: starts[i] + n])
        ms.append(mv)
        outs.append(out[starts[i]: starts[i] + n])
    table = csx.segment_table(ins, ms, outs, lengths)
    counts = torch.zeros(nseg, dtype=torch.int64, device="cuda")
    ws = csx.alloc_workspace(csx.batch_workspace_size(nseg, total, 4))
    csx.compact_batch(table, 4, total, counts=counts, workspace=ws)
    torch.cuda.synchronize()
    ks = counts.cpu().numpy()
    step_bytes = float(4 * total + total / 8 + 4 * ks.sum())
    res = {"columns": nseg, "total_elements": total, "algorithmic_bytes": int(step_bytes)}

    # the default tile (8 vectors per thread), then the autotuner on this table
    cfg8 = csx.Config(csx.PIPELINE_THREE_PHASE, 0, 8, 4096)
    ms_d = time_steps(lambda i: csx.compact_batch_config(table, 4, total, cfg8, counts=counts, workspace=ws), steps, warmup)
    res["default_config_GB/s"] = round(steps * step_bytes / (ms_d * 1e-3) / 1e9, 1)
    tuned, _ = csx.autotune_batch(table, 4, total, counts, ws, reps=5)
    res["tuned_config"] = tuned.as_dict()
    ms_b = time_steps(lambda i: csx.compact_batch(table, 4, total, counts=counts, workspace=ws), steps, warmup)
    res["batched_GB/s"] = round(steps * step_bytes / (ms_b * 1e-3) / 1e9, 1)
    res["batched_ms"] = round(ms_b / steps, 4)
    torch.cuda.synchronize()
    if not np.array_equal(counts.cpu().numpy(), ks):
        print("[bench] MISMATCH batched counts with the tuned configuration", file=sys.stderr)
        sys.exit(1)
    if batched_only:
        res["sparse_share"] = sparse_share
        csx.compact_batch(table, 4, total, counts=counts, workspace=ws)
        torch.cuda.synchronize()
        for i in (0, nseg // 2, int(np.argmax(lengths)), int(np.argmin(sels))):
            n = int(lengths[i])
            expect = torch.masked_select(ins[i], unpack_msb_first(ms[i], n))
            k = int(counts[i].item())
            if k != expect.numel() or not torch.equal(outs[i][:k], expect):
                print(f"[bench] batch MISMATCH in column {i}", file=sys.stderr)
                sys.exit(1)
        print(f"[bench] batch (sparse share {sparse_share}): {res}", file=sys.stderr, flush=True)
        del data, out, masks, table, ws
        torch.cuda.empty_cache()
        return res

    # one csx_compact per column, through the C ABI with prebuilt arguments
    L = csx.lib()
    cws = csx.alloc_workspace(csx.workspace_size(int(lengths.max()), 4, 1024))
    one = torch.empty(1, dtype=torch.int64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = [(ctypes.c_void_p(ins[i].data_ptr()), ctypes.c_void_p(ms[i].data_ptr()), ctypes.c_void_p(outs[i].data_ptr()), int(lengths[i]))
            for i in range(nseg)]
    wsp, wsn, cnt = ctypes.c_void_p(cws.data_ptr()), cws.numel(), ctypes.c_void_p(one.data_ptr())

    def per_column(_i, stream=stream):
        for a_in, a_m, a_o, n in args:
            L.csx_compact(a_in, a_m, a_o, n, 4, 1024, 0, csx.PIPELINE_THREE_PHASE, cnt, None, wsp, wsn, stream)

    ms_c = time_steps(per_column, max(2, steps // 5), 1)
    res["per_column_GB/s"] = round(max(2, steps // 5) * step_bytes / (ms_c * 1e-3) / 1e9, 1)
    res["per_column_ms"] = round(ms_c / max(2, steps // 5), 4)

    # both forms captured into graphs (no host launch cost at all)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    gb, gc = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(gb, stream=s):
            csx.compact_batch(table, 4, total, counts=counts, workspace=ws, stream=s)
        with torch.cuda.graph(gc, stream=s):
            per_column(0, stream=ctypes.c_void_p(s.cuda_stream))
    torch.cuda.synchronize()
    ms_gb = time_steps(lambda i: gb.replay(), steps, warmup)
    ms_gc = time_steps(lambda i: gc.replay(), steps, warmup)
    res["batched_graph_GB/s"] = round(steps * step_bytes / (ms_gb * 1e-3) / 1e9, 1)
    res["per_column_graph_GB/s"] = round(steps * step_bytes / (ms_gc * 1e-3) / 1e9, 1)
    res["speedup_vs_per_column"] = round(ms_c / max(2, steps // 5) / (ms_b / steps), 2)
    res["speedup_vs_per_column_graph"] = round(ms_gc / ms_gb, 2)
    # correctness of the replayed batch: three columns against torch's masked_select
    csx.compact_batch(table, 4, total, counts=counts, workspace=ws)
    torch.cuda.synchronize()
    for i in (0, nseg // 2, int(np.argmax(lengths))):
        n = int(lengths[i])
        expect = torch.masked_select(ins[i], unpack_msb_first(ms[i], n))
        k = int(counts[i].item())
        if k != expect.numel() or not torch.equal(outs[i][:k], expect):
            print(f"[bench] batch MISMATCH in column {i}", file=sys.stderr)
            sys.exit(1)
    print(f"[bench] batch: {res}", file=sys.stderr, flush=True)
    del data, out, masks, table, ws, cws, gb, gc
    torch.cuda.empty_cache()
    return res


def touched_bytes(mask, n, w, k, line=128):
    """Bytes a low-selectivity call actually moves: the mask twice (popcount
    and writeout), every `line`-byte data line holding a selected element once,
    and the k selected elements written.  MSB-first mask, n a multiple of 8."""
    sh = torch.arange(7, -1, -1, dtype=torch.uint8, device=mask.device)
    epl = line // w
    m = n // epl * epl
    bits = ((mask[: m // 8].view(-1, 1) >> sh) & 1).view(-1, epl)
    lines = int(bits.any(dim=1).sum().item())
    if m < n:
        lines += 1
    return 2 * ((n + 7) // 8) + line * lines + k * w


def headline(n, K, W, datas, masks, ks, out, count):
    """The metric's configuration: csx_compact_auto over 2^28 4-byte elements
    at p = 0.5.  Returns (GB/s, ms per step, writeout kernel ms, bytes per
    step, tuned configuration); the output of the last step is compared with
    torch's masked_select."""
    w = 4
    k_mean = sum(ks) / len(ks)
    step_bytes = w * n + n / 8 + w * k_mean
    ws = csx.alloc_workspace(csx.auto_workspace_size(n, w))

    def step(i):
        csx.compact_auto(datas[i % ROTATE], masks[i % ROTATE], 0.5, n=n, out=out, count=count, workspace=ws, host_count=False)

    ms = time_steps(step, K, W)
    # the dominant kernel alone: its start and end recorded by HIP events
    # launched with it on its stream
    device_warm()
    csx.profile_begin()
    for i in range(K):
        step(i)
    kern_ms, launches = csx.profile_end()
    kern_avg = kern_ms / max(launches, 1)
    r = (K - 1) % ROTATE
    step(K - 1)
    torch.cuda.synchronize()
    expect = torch.masked_select(datas[r], unpack_msb_first(masks[r], n))
    kk = int(count.item())
    if kk != ks[r] or kk != expect.numel() or not torch.equal(out[:kk].view(torch.int32), expect.view(torch.int32)):
        print("[bench] MISMATCH in the headline output", file=sys.stderr)
        sys.exit(1)
    del expect
    gbs = K * step_bytes / (ms * 1e-3) / 1e9
    cfg = csx.tune_lookup(w, n, 0.5).as_dict()
    print(f"[bench] headline auto p=0.5: {gbs:.1f} GB/s ({ms / K:.4f} ms/step, writeout kernel {kern_avg:.4f} ms) {cfg}", file=sys.stderr,
          flush=True)
    return gbs, ms / K, kern_avg, step_bytes, cfg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--log2n", type=int, default=None, help="log2 of the element count (28; a sweep's own default otherwise)")
    ap.add_argument("--chunk", type=int, default=1024)
    ap.add_argument("--only", choices=["headline"], default=None,
                    help="headline: only the metric's configuration (the command the committed kernel-trace profile runs)")
    ap.add_argument("--sweep", choices=list(SWEEPS), default=None, help="run one sweep instead of the benchmark line")
    ap.add_argument("--reps", type=int, default=10, help="timed repetitions per point of a sweep (median)")
    ap.add_argument("--json", default=None, help="a sweep's rows also as JSON")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the 4096-column batched measurement")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline")
    ap.add_argument("--sels", default="0.01,0.5,0.99", help="selectivities to sweep")
    ap.add_argument("--pipelines", default="three_phase,two_pass,fused,sparse", help="pipelines to sweep")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="length of the serving-loop run (0 skips it)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the compaction shards "
              "by contiguous ranges with one exchange of per-GPU counts)", file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    if args.sweep:
        rows = SWEEPS[args.sweep](args)
        if args.json:
            with open(args.json, "w") as f:
                json.dump(rows, f, indent=1)
        return

    sels = sorted(set(float(x) for x in args.sels.split(",")))
    all_pipes = {"three_phase": csx.PIPELINE_THREE_PHASE, "two_pass": csx.PIPELINE_TWO_PASS, "fused": csx.PIPELINE_FUSED, "sparse": csx.PIPELINE_SPARSE}
    pipes = [(k, all_pipes[k]) for k in all_pipes if k in args.pipelines.split(",")]

    log2n = args.log2n if args.log2n else 28
    n = 1 << log2n
    w = 4
    K, W = args.steps, args.warmup
    print(f"[bench] n=2^{log2n} steps={K} warmup={W}", file=sys.stderr, flush=True)
    datas = [make_data(n, 100 + r) for r in range(ROTATE)]
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    count = torch.empty(1, dtype=torch.int64, device="cuda")

    # the headline first, on a quiet device
    masks, ks = zip(*[make_mask(n, 0.5, 1000 * r + 50) for r in range(ROTATE)])
    main_result = headline(n, K, W, datas, masks, ks, out, count)
    del masks

    sweep = {}
    if args.only != "headline":
        ws = csx.alloc_workspace(csx.workspace_size(n, w, args.chunk))
        for p in sels:
            masks, ks = zip(*[make_mask(n, p, 1000 * r + int(p * 100)) for r in range(ROTATE)])
            k_mean = sum(ks) / len(ks)
            step_bytes = 4 * n + n / 8 + 4 * k_mean
            # low selectivity: the bytes such a call physically moves (whole
            # 128-byte data lines around the selected elements), beside the
            # metric's algorithmic bytes
            phys = touched_bytes(masks[0], n, w, ks[0]) if p <= 0.1 else None

            def record(name, ms, kern_avg=None, extra=None):
                gbs = K * step_bytes / (ms * 1e-3) / 1e9
                e = {"GB/s": round(gbs, 1), "pct_peak": round(100 * gbs / HBM_PEAK_GBS, 2), "ms_per_step": round(ms / K, 5), "selected": int(k_mean)}
                if kern_avg is not None:
                    e["kernel_ms"] = round(kern_avg, 5)
                if phys is not None:
                    pg = phys / (ms / K * 1e-3) / 1e9
                    e["physical_bytes"] = int(phys)
                    e["physical_GB/s"] = round(pg, 1)
                    e["physical_pct_peak"] = round(100 * pg / HBM_PEAK_GBS, 2)
                if extra:
                    e.update(extra)
                sweep[name] = e
                print(f"[bench] {name}: {gbs:.1f} GB/s ({ms / K:.4f} ms/step)", file=sys.stderr, flush=True)

            def check(name, r):
                torch.cuda.synchronize()
                expect = torch.masked_select(datas[r], unpack_msb_first(masks[r], n))
                kk = int(count.item())
                if kk != expect.numel() or not torch.equal(out[:kk].view(torch.int32), expect.view(torch.int32)):
                    print(f"[bench] MISMATCH {name}", file=sys.stderr)
                    sys.exit(1)

            for name, pipeline in pipes:

                def step(i, pipeline=pipeline):
                    csx.compact(datas[i % ROTATE], masks[i % ROTATE], n=n, chunk_len=args.chunk, pipeline=pipeline, out=out, count=count,
                                workspace=ws, host_count=False)

                ms = time_steps(step, K, W)
                csx.profile_begin()
                for i in range(K):
                    step(i)
                kern_ms, launches = csx.profile_end()
                record(f"{name}@{p}", ms, kern_ms / max(launches, 1))
                step(K - 1)
                check(f"{name}@{p}", (K - 1) % ROTATE)
            if p != 0.5:
                ws_auto = csx.alloc_workspace(csx.auto_workspace_size(n, w))

                def auto_step(i):
                    csx.compact_auto(datas[i % ROTATE], masks[i % ROTATE], p, n=n, out=out, count=count, workspace=ws_auto, host_count=False)

                ms = time_steps(auto_step, K, W)
                record(f"auto@{p}", ms, extra={"config": csx.tune_lookup(w, n, p).as_dict()})
                auto_step(K - 1)
                check(f"auto@{p}", (K - 1) % ROTATE)
                del ws_auto
            # vendor baseline on the same inputs: rocPRIM flagged select
            temp = torch.empty(max(baseline.temp_size(w, n), 16), dtype=torch.uint8, device="cuda")
            ms = time_steps(lambda i: baseline.rocprim_select(datas[i % ROTATE], masks[i % ROTATE], n=n, out=out, count=count, temp=temp), K, W)
            record(f"rocprim_select@{p}", ms)
            check(f"rocprim_select@{p}", (K - 1) % ROTATE)
            # ... and with the flags pre-expanded to one byte per element (not timed)
            flags = [unpack_msb_first(m, n).to(torch.uint8) for m in masks]
            temp = torch.empty(max(baseline.temp_size_bytes(w, n), 16), dtype=torch.uint8, device="cuda")
            ms = time_steps(lambda i: baseline.rocprim_select_bytes(datas[i % ROTATE], flags[i % ROTATE], n=n, out=out, count=count, temp=temp), K, W)
            record(f"rocprim_select_byteflags@{p}", ms)
            check(f"rocprim_select_byteflags@{p}", (K - 1) % ROTATE)
            del masks, temp, flags
        del ws

        if not args.no_batch:
            del datas
            torch.cuda.empty_cache()
            sweep["batch_4096_columns"] = bench_batch(args.steps, args.warmup)
            sweep["batch_4096_columns_half_sparse"] = bench_batch(args.steps, args.warmup, sparse_share=0.5, batched_only=True)
            datas = [make_data(n, 100 + r) for r in range(ROTATE)]

        sweep["predicate_f32_gt_0.5"] = bench_predicate(datas, n, K, W, out, count)

        if args.serve_seconds > 0:
            sweep["serving_poisson_20k"] = bench_serving(args.serve_seconds)

    # plain device copy of the input size (csx_copy): what this HBM sustains,
    # and the known-size kernel the counter runs calibrate against
    copy_steps = K if args.only != "headline" else 5
    copy_ms = time_steps(lambda i: csx.copy(out, datas[i % ROTATE]), copy_steps, min(W, 2))
    sweep["copy_4N"] = {"GB/s": round(copy_steps * 2 * 4 * n / (copy_ms * 1e-3) / 1e9, 1), "ms_per_step": round(copy_ms / copy_steps, 5)}

    gbs, ms_step, kern_avg, step_bytes, cfg = main_result
    achieved = step_bytes / (kern_avg * 1e-3) / 1e9
    if cfg["pipeline"] == csx.PIPELINE_THREE_PHASE:
        kernel = f"csx::k_tile<4, 0, true, 0, true, {cfg['vpt']}, {cfg['block_size']}>"
    else:
        kernel = f"the dominant kernel of pipeline {cfg['pipeline']}"
    # HBM bytes per launch of that kernel, from the committed counter summary
    # (null when it was taken for another tile shape)
    traffic = None
    tpath = os.path.join(ROOT, TRAFFIC_FILE)
    if os.path.isfile(tpath):
        with open(tpath) as f:
            t = json.load(f)
        if t.get("kernel") == kernel:
            traffic = t.get("hbm_bytes_per_launch")
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernel": kernel + (" (phase 3, staged writeout)" if cfg["pipeline"] == csx.PIPELINE_THREE_PHASE else ""),
        "kernel_ms": round(kern_avg, 5),
        "algorithmic_bytes_per_launch": int(step_bytes),
        "profile": PROFILE_FILE,
    }

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(datas[0], n, args.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: float32 uniform [0,1) moved as u32 words, seeded torch generators, MSB-first Bernoulli masks, 3 rotating input sets",
        "config": {
            "workload": "single-column compaction of 2^28 4-byte elements, Bernoulli(0.5) mask, tuned configuration (csx_compact_auto: "
                        f"block popcount with unit offsets, block prefixes, staged writeout; chunk {cfg['chunk_len']}, {cfg['vpt']} vectors per thread, "
                        f"{cfg['block_size']}-thread tiles)",
            "n": n,
            "elem_bytes": w,
            "selectivity": 0.5,
            "chunk_len": cfg["chunk_len"],
            "tile_vectors_per_thread": cfg["vpt"],
            "tile_workgroup_size": cfg["block_size"],
            "pct_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 2),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if "three_phase@0.5" in sweep:
        line["config"]["default_config_GB/s"] = sweep["three_phase@0.5"]["GB/s"]
    if "rocprim_select@0.5" in sweep:
        line["vs_rocprim_select"] = round(gbs / sweep["rocprim_select@0.5"]["GB/s"], 3)
        line["vs_rocprim_select_byteflags"] = round(gbs / sweep["rocprim_select_byteflags@0.5"]["GB/s"], 3)
    if sweep:
        line["sweep"] = sweep
    print(json.dumps(line), flush=True)


def bench_predicate(datas, n, K, W, out, count):
    """Column -> compacted column for `value > 0.5` over float32 in [0, 1):
    fused (csx_compact_predicate: predicate + chunk counts, scan, writeout)
    against two calls (csx_predicate_mask, then csx_compact), and the mask
    step alone.  Algorithmic bytes: n*4 read + k*4 written (the mask is
    internal); the mask step alone: n*4 read + n/8 written."""
    mask = torch.empty((n + 7) // 8, dtype=torch.uint8, device="cuda")
    ws_pred = csx.alloc_workspace(csx.predicate_workspace_size(n, 8))
    ws = csx.alloc_workspace(csx.workspace_size(n, 4, 1024))
    mcount = torch.empty(1, dtype=torch.int64, device="cuda")
    fused_ms = time_steps(lambda i: csx.compact_predicate(datas[i % ROTATE], ">", 0.5, out=out, count=count, workspace=ws_pred, host_count=False), K, W)
    k = int(count.item())

    def two(i):
        csx.predicate_mask(datas[i % ROTATE], ">", 0.5, mask=mask, count=mcount)
        csx.compact(datas[i % ROTATE], mask, n=n, out=out, count=count, workspace=ws, host_count=False)

    two_ms = time_steps(two, K, W)
    mask_ms = time_steps(lambda i: csx.predicate_mask(datas[i % ROTATE], ">", 0.5, mask=mask, count=mcount), K, W)
    # the fused result against torch on the last input
    r = (K - 1) % ROTATE
    _, kk = csx.compact_predicate(datas[r], ">", 0.5, out=out, count=count, workspace=ws_pred)
    ref = torch.masked_select(datas[r], datas[r] > 0.5)
    if kk != ref.numel() or not torch.equal(out[:kk], ref):
        print("[bench] MISMATCH predicate compaction", file=sys.stderr)
        sys.exit(1)
    alg = 4 * n + 4 * k
    res = {
        "selected": k,
        "fused_GB/s": round(K * alg / (fused_ms * 1e-3) / 1e9, 1),
        "fused_ms": round(fused_ms / K, 5),
        "two_calls_GB/s": round(K * alg / (two_ms * 1e-3) / 1e9, 1),
        "two_calls_ms": round(two_ms / K, 5),
        "mask_step_GB/s": round(K * (4 * n + n / 8) / (mask_ms * 1e-3) / 1e9, 1),
        "mask_step_ms": round(mask_ms / K, 5),
    }
    print(f"[bench] predicate f32 > 0.5: fused {res['fused_GB/s']} GB/s ({res['fused_ms']} ms), two calls {res['two_calls_ms']} ms, "
          f"mask step {res['mask_step_GB/s']} GB/s", file=sys.stderr, flush=True)
    return res


def bench_serving(seconds):
    """The serving loop (gpu_compressstore2_amd/csx_serve, its own process):
    Poisson arrivals at 20k requests/s of u32/f32/u64 requests, lengths
    log-uniform in [2^12, 2^26], Bernoulli masks with p uniform in [0, 1],
    64 GiB device budget; every 64th request is recomputed alone and compared."""
    exe = os.path.join(ROOT, "gpu_compressstore2_amd", "csx_serve")
    print(f"[bench] serving loop, {seconds:g} s at 20000 requests/s", file=sys.stderr, flush=True)
    r = subprocess.run([exe, "-t", str(seconds)], capture_output=True, text=True, timeout=seconds + 240)
    if r.returncode != 0:
        print(f"[bench] serving loop failed ({r.returncode}): {r.stderr[-2000:]}", file=sys.stderr)
        sys.exit(1)
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print(f"[bench] serving: {res['achieved_rate']:.0f} req/s, {res['GB/s']} GB/s, p50 {res['p50_ms']} ms, p99 {res['p99_ms']} ms", file=sys.stderr,
          flush=True)
    return res


def cpu_baseline(data_dev, n, budget_s):
    """The reference CPU path on the host, one thread: the same 2^28 workload at
    50 % (the reported value), then the same data at 1 % and 99 % and its first
    2^24 elements at 50 % (the CPU-runnable configuration), each the median of
    up to 5 runs within its share of the time budget."""
    import numpy as np

    import oracle

    data = data_dev.cpu().numpy().view(np.uint32)
    if oracle.reference_available():
        impl, kind = oracle.Reference(), "reference"
    else:
        impl, kind = oracle.Oracle(), "port"

    def timed(d, m, budget):
        times, k = [], 0
        t_all = time.perf_counter()
        while len(times) < 5 and (not times or time.perf_counter() - t_all < budget):
            t0 = time.perf_counter()
            _, k = impl.generate_validation(d, m)
            times.append(time.perf_counter() - t0)
        t = sorted(times)[len(times) // 2]
        return (4 * d.size + d.size / 8 + 4 * k) / t / 1e9, t, len(times)

    mask = make_mask(n, 0.5, 7)[0].cpu().numpy()
    gbs, t, runs = timed(data, mask, budget_s * 0.5)
    also = {}
    for p in (0.01, 0.99):
        g, _, _ = timed(data, make_mask(n, p, 8)[0].cpu().numpy(), budget_s * 0.2)
        also[f"2^28 u32 p={p}"] = round(g, 3)
    n24 = 1 << 24
    g, _, _ = timed(np.ascontiguousarray(data[:n24]), np.ascontiguousarray(mask[: n24 // 8]), budget_s * 0.1)
    also["2^24 u32 p=0.5"] = round(g, 3)
    cpu_model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    cpu_model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": round(gbs, 3),
        "unit": "GB/s",
        "cores": 1,
        "kind": kind,
        "sample": f"full 2^28 u32 workload at p=0.5, generate_validation single-threaded, median of {runs} runs "
                  f"({t:.3f} s each); host nproc={os.cpu_count()}, cpu={cpu_model}",
        "also_GB/s": also,
    }


# ---------------------------------------------------------------- sweeps
# `python bench.py --sweep NAME [--reps R] [--log2n L] [--json FILE]`: one CSV
# line per point on stdout, each the per-call time of the whole call (or of
# one phase), every count checked.

def median_ms(fn, reps):
    """Per-call time as the device sees it in a stream of calls: R calls
    enqueued back to back between two HIP events, three rounds after a
    warm-up, the median round divided by R (the way the autotuner times its
    candidates; one call alone between events would also time the host's
    launch path)."""
    fn()
    times = []
    for _ in range(3):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b) / reps)
    times.sort()
    return times[1]


def graph_ms(fn, reps):
    """Per-call device time without the host: R calls captured in one graph,
    replayed in three rounds after a warm-up, the median round divided by R.
    `fn(stream)` enqueues one call on the given stream."""
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        fn(s)
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            for _ in range(reps):
                fn(s)
    torch.cuda.synchronize()
    g.replay()
    times = []
    for _ in range(3):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        g.replay()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b) / reps)
    times.sort()
    del g
    return times[1]


def _count_check(count, k, what):
    if int(count.item()) != k:
        print(f"[bench] count mismatch: {what}", file=sys.stderr)
        sys.exit(1)


def sweep_phases(args):
    """2^28 float32 in [0, 1); Bernoulli masks at 1/10/50/90/99 % and
    alternating runs of geometric mean 256 (50 %); chunk lengths 256, 1024,
    4096.  Per writeout variant (staged, direct, thread per chunk, sparse) the
    three phases timed separately -- popcount (csx_popc), scan (csx_scan),
    writeout (csx_writeout) -- the way the reference reports time_popc /
    time_pss / time_proc (src/benchmarks.cuh:10-52); the whole-call pipelines
    and the tuned call timed as one.  GB/s counts n*4 + n/8 + k*4 bytes."""
    n = 1 << (args.log2n if args.log2n else 28)
    g = torch.Generator(device="cuda").manual_seed(11)
    data = torch.rand(n, device="cuda", generator=g)
    out = torch.empty_like(data)
    masks = {f"bernoulli_{p}": csx.gen_mask_bernoulli(n, p, seed=7) for p in (0.01, 0.1, 0.5, 0.9, 0.99)}
    masks["runs_256"] = csx.gen_mask_runs(n, 256, seed=7)
    variants = {"staged": csx.WRITEOUT_STAGED, "direct": csx.WRITEOUT_DIRECT, "thread_per_chunk": csx.WRITEOUT_THREAD_PER_CHUNK,
                "sparse": csx.WRITEOUT_SPARSE}
    pipelines = {"three_phase": csx.PIPELINE_THREE_PHASE, "two_pass": csx.PIPELINE_TWO_PASS, "fused": csx.PIPELINE_FUSED, "sparse": csx.PIPELINE_SPARSE}
    count = torch.empty(1, dtype=torch.int64, device="cuda")
    rows = []
    print("mask;chunk_len;form;time_popc_ms;time_scan_ms;time_writeout_ms;time_total_ms;selected;gbps;pct_peak", flush=True)

    def emit(mname, chunk, form, tp, ts, tw, total, k):
        gbps = (4 * n + n / 8 + 4 * k) / (total * 1e-3) / 1e9
        row = {"mask": mname, "chunk_len": chunk, "form": form, "popc_ms": tp, "scan_ms": ts, "writeout_ms": tw, "total_ms": round(total, 5),
               "selected": k, "GB/s": round(gbps, 1), "pct_peak": round(100 * gbps / HBM_PEAK_GBS, 2)}
        rows.append(row)
        f = lambda v: "" if v is None else f"{v:.5f}"
        print(f"{mname};{chunk};{form};{f(tp)};{f(ts)};{f(tw)};{total:.5f};{k};{gbps:.1f};{row['pct_peak']}", flush=True)

    for mname, mask in masks.items():
        for chunk in (256, 1024, 4096):
            counts = csx.popc(mask, n, chunk)
            offsets, total = csx.scan(counts)
            torch.cuda.synchronize()
            k = int(total.item())
            if k != int(counts.to(torch.int64).sum().item()):
                print(f"[bench] count mismatch {mname} chunk {chunk}", file=sys.stderr)
                sys.exit(1)
            scan_ws = csx.alloc_workspace(csx.scan_workspace_size(counts.numel()))
            tp = median_ms(lambda: csx.popc(mask, n, chunk, counts=counts), args.reps)
            ts = median_ms(lambda: csx.scan(counts, offsets=offsets, total=total, workspace=scan_ws), args.reps)
            for vname, v in variants.items():
                reps = args.reps if v != csx.WRITEOUT_THREAD_PER_CHUNK else max(2, args.reps // 5)
                tw = median_ms(lambda: csx.writeout(data, mask, offsets, n, chunk, variant=v, counts=counts, out=out), reps)
                emit(mname, chunk, f"phases_{vname}", round(tp, 5), round(ts, 5), round(tw, 5), tp + ts + tw, k)
            ws = csx.alloc_workspace(csx.workspace_size(n, 4, chunk))
            for pname, pl in pipelines.items():
                tt = median_ms(lambda: csx.compact(data, mask, n=n, chunk_len=chunk, pipeline=pl, out=out, count=count, workspace=ws, host_count=False),
                               args.reps)
                _count_check(count, k, f"{mname} chunk {chunk} {pname}")
                emit(mname, chunk, pname, None, None, None, tt, k)
            del ws, scan_ws
        wsa = csx.alloc_workspace(csx.auto_workspace_size(n, 4))
        ta = median_ms(lambda: csx.compact_auto(data, mask, k / n, n=n, out=out, count=count, workspace=wsa, host_count=False), args.reps)
        _count_check(count, k, f"{mname} tuned")
        emit(mname, 0, "tuned", None, None, None, ta, k)
        del wsa
    return rows


WIDTHS = (("u8", 1, 0), ("u16", 2, 0), ("u32", 4, 0), ("u64", 8, 0), ("f64", 8, 1))


def sweep_widths(args):
    """2^29 elements (default) of u8, u16, u32, u64 and f64, each with a
    Bernoulli(0.25) mask, a periodic seeded 32-bit pattern and Zipf(s = 1.2)
    per-4096-block selectivity over {0, 1/64, ..., 1}: the default
    three-phase call, the tuned call (selectivity k/n) and, for the periodic
    mask, the mask-free pattern kernel."""
    n = 1 << (args.log2n if args.log2n else 29)
    pattern = int(csx.gen_values(1, 4, seed=29).view(torch.int32).item()) & 0xFFFFFFFF
    masks = {"bernoulli_0.25": csx.gen_mask_bernoulli(n, 0.25, seed=3), "periodic": csx.gen_mask_pattern(n, pattern, 32),
             "zipf_blocks": csx.gen_mask_zipf_blocks(n, 4096, 1.2, 65, seed=5)}
    count = torch.empty(1, dtype=torch.int64, device="cuda")
    rows = []
    print("type;mask;selected;default_ms;default_GB/s;tuned_ms;tuned_GB/s;tuned_config;pattern_kernel_ms", flush=True)
    for name, w, kind in WIDTHS:
        data = csx.gen_values(n, w, kind, seed=100 + w + kind)
        out = torch.empty_like(data)
        ws = csx.alloc_workspace(csx.auto_workspace_size(n, w))
        for mname, mask in masks.items():
            _, k = csx.compact(data, mask, n=n, out=out, workspace=ws)
            td = median_ms(lambda: csx.compact(data, mask, n=n, out=out, count=count, workspace=ws, host_count=False), args.reps)
            _count_check(count, k, f"{name} {mname} default")
            ta = median_ms(lambda: csx.compact_auto(data, mask, k / n, n=n, out=out, count=count, workspace=ws, host_count=False), args.reps)
            _count_check(count, k, f"{name} {mname} tuned")
            cfg = csx.tune_lookup(w, n, k / n).as_dict()
            tp = None
            if mname == "periodic":
                tp = median_ms(lambda: csx.pattern_compact(data, pattern, 32, n=n, out=out, count=count), args.reps)
                _count_check(count, k, f"{name} pattern kernel")
            b = n * w + n / 8 + k * w
            row = {"type": name, "mask": mname, "selected": k, "default_ms": round(td, 5), "default_GB/s": round(b / (td * 1e6), 1),
                   "tuned_ms": round(ta, 5), "tuned_GB/s": round(b / (ta * 1e6), 1), "tuned_config": cfg,
                   "pattern_kernel_ms": None if tp is None else round(tp, 5)}
            rows.append(row)
            print(f"{name};{mname};{k};{td:.5f};{row['default_GB/s']};{ta:.5f};{row['tuned_GB/s']};{cfg};{'' if tp is None else f'{tp:.5f}'}",
                  flush=True)
        del data, out, ws
        torch.cuda.empty_cache()
    return rows


def sweep_shapes(args):
    """Tile size against mask shape: 2^28 elements of every width; Bernoulli
    masks at 10-90 %, alternating runs of geometric mean 16-4096 (50 %) and
    Zipf blocks; the three-phase call (chunk 1024) with 2/4/8/16-vector tiles
    of 256 threads, 2-vector tiles of 128 and 4-vector tiles of 512 threads,
    and the two-pass call, each row with its mask's occupied 64-byte-line
    ratio against a Bernoulli mask of the same density."""
    n = 1 << (args.log2n if args.log2n else 28)
    masks = {f"bernoulli_{p}": csx.gen_mask_bernoulli(n, p, seed=11) for p in (0.1, 0.25, 0.5, 0.75, 0.9)}
    masks.update({f"runs_{r}": csx.gen_mask_runs(n, r, seed=7) for r in (16, 64, 256, 1024, 4096)})
    masks["zipf_blocks"] = csx.gen_mask_zipf_blocks(n, 4096, 1.2, 65, seed=5)
    forms = [("vpt2", csx.PIPELINE_THREE_PHASE, 2, 256), ("vpt4", csx.PIPELINE_THREE_PHASE, 4, 256), ("vpt8", csx.PIPELINE_THREE_PHASE, 8, 256),
             ("vpt16", csx.PIPELINE_THREE_PHASE, 16, 256), ("vpt2_b128", csx.PIPELINE_THREE_PHASE, 2, 128),
             ("vpt4_b512", csx.PIPELINE_THREE_PHASE, 4, 512), ("two_pass", csx.PIPELINE_TWO_PASS, 8, 256)]
    count = torch.empty(1, dtype=torch.int64, device="cuda")
    rows = []
    print("width;mask;selected;line_ratio;" + ";".join(f"{f[0]}_us" for f in forms) + ";best", flush=True)
    for w in (1, 2, 4, 8):
        data = csx.gen_values(n, w, seed=3)
        out = torch.empty_like(data)
        ws = csx.alloc_workspace(csx.auto_workspace_size(n, w))
        for mname, mask in masks.items():
            _, k = csx.compact(data, mask, n=n, out=out, workspace=ws)
            epl = 64 // w
            sh = torch.arange(7, -1, -1, dtype=torch.uint8, device="cuda")
            lines = int(((mask[: n // 8].view(-1, 1) >> sh) & 1).view(-1, epl).any(dim=1).sum().item())
            expected = (n // epl) * (1.0 - (1.0 - k / n) ** epl)
            row = {"width": w, "mask": mname, "selected": k, "line_ratio": round(lines / expected if expected else 1.0, 3)}
            for fname, pipeline, vpt, block in forms:
                if vpt == 16 and w == 1:
                    row[f"{fname}_us"] = None
                    continue
                cfg = csx.Config(pipeline, 1024, vpt, 0, block)
                t = median_ms(lambda: csx.compact_config(data, mask, cfg, n=n, out=out, count=count, workspace=ws, host_count=False), args.reps)
                _count_check(count, k, f"{w} {mname} {fname}")
                row[f"{fname}_us"] = round(t * 1000, 1)
            row["best"] = min((v, f) for f, v in row.items() if f.endswith("_us") and v is not None)[1][:-3]
            rows.append(row)
            print(f"{w};{mname};{k};{row['line_ratio']};" + ";".join("" if row[f'{f[0]}_us'] is None else str(row[f'{f[0]}_us']) for f in forms)
                  + f";{row['best']}", flush=True)
        del data, out, ws
        torch.cuda.empty_cache()
    return rows


def sweep_chunks(args):
    """Chunk length against time: 2^28 u32, Bernoulli 0.5, the three-phase
    call with the tuned 2-vector tile and the default 8-vector tile at chunk
    lengths 2^8 .. 2^17 (the reference's driver accepts any power of two
    >= 32, src/main.cu:35-36,236-245), and the sparse call at 1 %."""
    n = 1 << (args.log2n if args.log2n else 28)
    data = csx.gen_values(n, 4, seed=3)
    out = torch.empty_like(data)
    count = torch.empty(1, dtype=torch.int64, device="cuda")
    rows = []
    print("mask;form;chunk_len;us;GB/s", flush=True)
    for p, forms in ((0.5, (("three_phase_vpt2", csx.PIPELINE_THREE_PHASE, 2), ("three_phase_vpt8", csx.PIPELINE_THREE_PHASE, 8))),
                     (0.01, (("sparse", csx.PIPELINE_SPARSE, 8),))):
        mask = csx.gen_mask_bernoulli(n, p, seed=5)
        _, k = csx.compact(data, mask, n=n)
        for fname, pl, vpt in forms:
            for lc in (8, 10, 12, 14, 16, 17):
                chunk = 1 << lc
                ws = csx.alloc_workspace(csx.workspace_size(n, 4, chunk))
                cfg = csx.Config(pl, chunk, vpt, 0)
                t = median_ms(lambda: csx.compact_config(data, mask, cfg, n=n, out=out, count=count, workspace=ws, host_count=False), args.reps)
                _count_check(count, k, f"{fname} chunk {chunk}")
                gbs = (4 * n + n / 8 + 4 * k) / (t * 1e-3) / 1e9
                rows.append({"p": p, "form": fname, "chunk_len": chunk, "us": round(t * 1000, 1), "GB/s": round(gbs, 1)})
                print(f"bernoulli_{p};{fname};{chunk};{t * 1000:.1f};{gbs:.1f}", flush=True)
                del ws
    return rows


def sweep_sizes(args):
    """Call time against column length: u32 at Bernoulli 0.5 from 2^12 to
    2^26 elements, every pipeline (chunk 1024, default tile) and the tuned
    call, each timed as R calls replayed from one captured graph (device time
    per call, no host launch path); at these sizes launches, not bytes, set
    the time."""
    count = torch.empty(1, dtype=torch.int64, device="cuda")
    rows = []
    pipes = (("three_phase", csx.PIPELINE_THREE_PHASE), ("two_pass", csx.PIPELINE_TWO_PASS), ("fused", csx.PIPELINE_FUSED),
             ("single", csx.PIPELINE_SINGLE))
    print("log2n;" + ";".join(f"{p}_us" for p, _ in pipes) + ";tuned_us;tuned_config", flush=True)
    for ln in range(12, 27, 2):
        n = 1 << ln
        data = csx.gen_values(n, 4, seed=ln)
        mask = csx.gen_mask_bernoulli(n, 0.5, seed=ln + 1)
        out = torch.empty_like(data)
        ws = csx.alloc_workspace(csx.auto_workspace_size(n, 4))
        _, k = csx.compact(data, mask, n=n)
        row = {"log2n": ln, "selected": k}
        for pname, pl in pipes:
            if pl == csx.PIPELINE_SINGLE and n > csx.SINGLE_MAX_ELEMENTS:
                row[f"{pname}_us"] = None
                continue
            t = graph_ms(lambda st: csx.compact(data, mask, n=n, pipeline=pl, out=out, count=count, workspace=ws, host_count=False, stream=st),
                         args.reps)
            _count_check(count, k, f"2^{ln} {pname}")
            row[f"{pname}_us"] = round(t * 1000, 2)
        t = graph_ms(lambda st: csx.compact_auto(data, mask, 0.5, n=n, out=out, count=count, workspace=ws, host_count=False, stream=st), args.reps)
        _count_check(count, k, f"2^{ln} tuned")
        row["tuned_us"] = round(t * 1000, 2)
        row["tuned_config"] = csx.tune_lookup(4, n, 0.5).as_dict()
        rows.append(row)
        print(f"{ln};" + ";".join(str(row[f'{p}_us']) for p, _ in pipes) + f";{row['tuned_us']};{row['tuned_config']}", flush=True)
    return rows


def sweep_ceiling(args):
    """What this HBM sustains for plain access mixes in the writeout's access
    shape (csx_stream_probe over 1 GiB): read, write, copy and the 2:1
    read/write mix of a 50 % compaction, with 1 or 4 vectors per thread and
    blocks in plain order or XCD runs of 16 / 32.  GB/s counts bytes read plus
    bytes written."""
    nbytes = 1 << 30
    src = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    dst = torch.zeros(nbytes // 4, dtype=torch.int32, device="cuda")
    kinds = (("read", csx.PROBE_READ, 1.0), ("write", csx.PROBE_WRITE, 1.0), ("copy", csx.PROBE_COPY, 2.0), ("mix2:1", csx.PROBE_MIX, 1.5))
    rows = []
    print("vectors_per_thread;xcd_run;" + ";".join(f"{k}_GB/s" for k, _, _ in kinds), flush=True)
    for vpt in (1, 4):
        for xr in (0, 16, 32):
            row = {"vectors_per_thread": vpt, "xcd_run": xr}
            for kname, kind, moved in kinds:
                t = median_ms(lambda: csx.stream_probe(kind, dst, src, nbytes, vpt, xr), args.reps)
                row[f"{kname}_GB/s"] = round(moved * nbytes / (t * 1e-3) / 1e9, 1)
            rows.append(row)
            print(f"{vpt};{xr};" + ";".join(str(row[f'{k}_GB/s']) for k, _, _ in kinds), flush=True)
    return rows


SWEEPS = {"phases": sweep_phases, "widths": sweep_widths, "shapes": sweep_shapes, "chunks": sweep_chunks, "sizes": sweep_sizes,
          "ceiling": sweep_ceiling}


if __name__ == "__main__":
    main()
