"""Sweeps of the single-column compaction.

    python -m gpu_compressstore2_amd.sweep [--log2n 28] [--reps 10] [--json FILE]
    python -m gpu_compressstore2_amd.sweep --widths [--log2n 29] [--reps 10] [--json FILE]
    python -m gpu_compressstore2_amd.sweep --shapes [--log2n 28] [--reps 15] [--json FILE]

Default: the per-phase sweep over 2^28 float32.

Masks: Bernoulli at 1, 10, 50, 90 and 99 % selectivity, and alternating
1-runs / 0-runs with geometric lengths of mean 256 (50 %).  Chunk lengths 256,
1024 and 4096.  For every writeout variant (staged, direct, thread per chunk)
the three phases -- popcount, scan, writeout -- are timed separately with HIP
events on the launching stream (median of `reps`), the way the reference's
benchmarks report time_popc / time_pss / time_proc (src/benchmarks.cuh:10-52);
the whole-call forms (three-phase, two-pass, fused, tuned) are timed as one.
Every configuration's count is checked against the mask's popcount.  One CSV
line per configuration; GB/s counts n*4 + n/8 + k*4 bytes.
"""
import argparse
import json
import sys

import torch

import gpu_compressstore2_amd as csx

VARIANTS = {"staged": csx.WRITEOUT_STAGED, "direct": csx.WRITEOUT_DIRECT, "thread_per_chunk": csx.WRITEOUT_THREAD_PER_CHUNK, "sparse": csx.WRITEOUT_SPARSE}
PIPELINES = {"three_phase": csx.PIPELINE_THREE_PHASE, "two_pass": csx.PIPELINE_TWO_PASS, "fused": csx.PIPELINE_FUSED, "sparse": csx.PIPELINE_SPARSE}


def median_ms(fn, reps):
    fn()
    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    times.sort()
    return times[len(times) // 2]


WIDTHS = (("u8", 1, csx.VALUES_INTEGER), ("u16", 2, csx.VALUES_INTEGER), ("u32", 4, csx.VALUES_INTEGER), ("u64", 8, csx.VALUES_INTEGER),
          ("f64", 8, csx.VALUES_FLOAT))


def widths_sweep(args):
    """Element widths and mask shapes: 2^29 elements (default) of u8, u16, u32,
    u64 and f64, each with a Bernoulli(0.25) mask, a periodic seeded 32-bit
    pattern and Zipf(s = 1.2) per-4096-block selectivity over {0, 1/64, ..., 1}.
    Per row: the default three-phase call, the tuned call (selectivity = k/n)
    and, for the periodic mask, the mask-free pattern kernel; each the median
    of `reps` HIP-event timings, its count checked against the first call's."""
    n = 1 << args.log2n
    pattern = int(csx.gen_values(1, 4, seed=29).view(torch.int32).item()) & 0xFFFFFFFF  # a seeded 32-bit pattern
    masks = {
        "bernoulli_0.25": csx.gen_mask_bernoulli(n, 0.25, seed=3),
        "periodic": csx.gen_mask_pattern(n, pattern, 32),
        "zipf_blocks": csx.gen_mask_zipf_blocks(n, 4096, 1.2, 65, seed=5),
    }
    count = torch.empty(1, dtype=torch.int64, device="cuda")
    rows = []
    print("type;mask;selected;default_ms;default_GB/s;tuned_ms;tuned_GB/s;tuned_config;pattern_kernel_ms", flush=True)
    for name, w, kind in WIDTHS:
        data = csx.gen_values(n, w, kind, seed=100 + w + kind)
        out = torch.empty_like(data)
        ws = csx.alloc_workspace(csx.auto_workspace_size(n, w))
        for mname, mask in masks.items():
            _, k = csx.compact(data, mask, n=n, out=out, workspace=ws)

            def checked(fn):
                t = median_ms(fn, args.reps)
                if int(count.item()) != k:
                    print(f"count mismatch {name} {mname}", file=sys.stderr)
                    sys.exit(1)
                return t

            td = checked(lambda: csx.compact(data, mask, n=n, out=out, count=count, workspace=ws, host_count=False))
            ta = checked(lambda: csx.compact_auto(data, mask, k / n, n=n, out=out, count=count, workspace=ws, host_count=False))
            cfg = csx.tune_lookup(w, n, k / n).as_dict()
            tp = None
            if mname == "periodic":
                tp = checked(lambda: csx.pattern_compact(data, pattern, 32, n=n, out=out, count=count))
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


def line_ratio(mask, n, w, p):
    """Occupied 64-byte data lines (64 / w consecutive elements holding a
    selected element) over the count a Bernoulli(p) mask occupies on average,
    ceil(n / (64/w)) (1 - (1-p)^(64/w)); counted on the device with torch,
    MSB-first bits, n a multiple of 64."""
    epl = 64 // w
    sh = torch.arange(7, -1, -1, dtype=torch.uint8, device=mask.device)
    bits = ((mask[: n // 8].view(-1, 1) >> sh) & 1).view(-1, epl)
    lines = int(bits.any(dim=1).sum().item())
    expected = (n // epl) * (1.0 - (1.0 - p) ** epl)
    return lines / expected if expected else 1.0


def shapes_sweep(args):
    """Tile size against mask shape: 2^28 elements (default) of every width;
    Bernoulli masks at 10-90 %, alternating runs of geometric mean length
    16-4096 (50 %) and Zipf(s = 1.2) per-4096-block selectivity; the
    three-phase call (chunk 1024) with tiles of 2, 4, 8 and 16 vectors per
    thread, and the two-pass call.  Each time the median of `reps` HIP-event
    timings of the whole call; every count checked against the default
    call's.  Per row also the mask's occupied-line ratio (line_ratio).  Then,
    for 4- and 8-byte elements, the tuned call on Bernoulli 0.5, runs and
    Zipf masks with the shipped table, and again after csx_autotune_inputs on
    that column and mask."""
    n = 1 << args.log2n
    masks = {f"bernoulli_{p}": csx.gen_mask_bernoulli(n, p, seed=11) for p in (0.1, 0.25, 0.5, 0.75, 0.9)}
    masks.update({f"runs_{r}": csx.gen_mask_runs(n, r, seed=7) for r in (16, 64, 256, 1024, 4096)})
    masks["zipf_blocks"] = csx.gen_mask_zipf_blocks(n, 4096, 1.2, 65, seed=5)
    forms = [("vpt2", csx.PIPELINE_THREE_PHASE, 2), ("vpt4", csx.PIPELINE_THREE_PHASE, 4), ("vpt8", csx.PIPELINE_THREE_PHASE, 8),
             ("vpt16", csx.PIPELINE_THREE_PHASE, 16), ("two_pass", csx.PIPELINE_TWO_PASS, 8)]
    count = torch.empty(1, dtype=torch.int64, device="cuda")
    rows = []
    print("width;mask;selected;line_ratio;" + ";".join(f"{f}_us" for f, _, _ in forms) + ";best", flush=True)
    for w in (1, 2, 4, 8):
        data = csx.gen_values(n, w, seed=3)
        out = torch.empty_like(data)
        ws = csx.alloc_workspace(csx.auto_workspace_size(n, w))
        for _ in range(40):  # clocks up before the first timed row
            csx.compact(data, masks["bernoulli_0.5"], n=n, out=out, workspace=ws, count=count, host_count=False)
        for mname, mask in masks.items():
            _, k = csx.compact(data, mask, n=n, out=out, workspace=ws)
            row = {"width": w, "mask": mname, "selected": k, "line_ratio": round(line_ratio(mask, n, w, k / n), 3)}
            for fname, pipeline, vpt in forms:
                if vpt == 16 and w == 1:
                    row[f"{fname}_us"] = None
                    continue
                cfg = csx.Config(pipeline, 1024, vpt, 0)
                t = median_ms(lambda: csx.compact_config(data, mask, cfg, n=n, out=out, count=count, workspace=ws, host_count=False), args.reps)
                if int(count.item()) != k:
                    print(f"count mismatch {w} {mname} {fname}", file=sys.stderr)
                    sys.exit(1)
                row[f"{fname}_us"] = round(t * 1000, 1)
            row["best"] = min((v, f) for f, v in row.items() if f.endswith("_us") and v is not None)[1][:-3]
            rows.append(row)
            print(f"{w};{mname};{k};{row['line_ratio']};" + ";".join("" if row[f'{f}_us'] is None else str(row[f'{f}_us']) for f, _, _ in forms)
                  + f";{row['best']}", flush=True)
        del data, out, ws
        torch.cuda.empty_cache()

    # the tuned call (selectivity k/n) with the shipped table, then after
    # csx_autotune_inputs on the same column and mask
    tuned = []
    print("width;mask;table_config;table_us;column_tuned_config;column_tuned_us", flush=True)
    for w in (4, 8):
        data = csx.gen_values(n, w, seed=3)
        out = torch.empty_like(data)
        ws = csx.alloc_workspace(csx.auto_workspace_size(n, w))
        for mname in ("bernoulli_0.5", "runs_256", "runs_1024", "zipf_blocks"):
            mask = masks[mname]
            csx.tune_clear()
            csx.tune_load(csx.TUNED_TABLE)
            _, k = csx.compact(data, mask, n=n, out=out, workspace=ws)
            sel = k / n

            def auto():
                csx.compact_auto(data, mask, sel, n=n, out=out, count=count, workspace=ws, host_count=False)

            before_cfg = csx.tune_lookup(w, n, sel).as_dict()
            before = median_ms(auto, args.reps)
            cfg, _, kk = csx.autotune_inputs(data, mask, n=n, reps=5, out=out, workspace=ws)
            after = median_ms(auto, args.reps)
            if kk != k or int(count.item()) != k:
                print(f"count mismatch {w} {mname} (tuned on the column)", file=sys.stderr)
                sys.exit(1)
            row = {"width": w, "mask": mname, "table_config": before_cfg, "table_us": round(before * 1000, 1), "column_tuned_config": cfg.as_dict(),
                   "column_tuned_us": round(after * 1000, 1)}
            tuned.append(row)
            print(f"{w};{mname};{before_cfg};{row['table_us']};{cfg.as_dict()};{row['column_tuned_us']}", flush=True)
        del data, out, ws
        torch.cuda.empty_cache()
    csx.tune_clear()
    csx.tune_load(csx.TUNED_TABLE)
    return {"tiles": rows, "tuned_on_column": tuned}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--log2n", type=int, default=None, help="log2 of the element count (28; 29 with --widths)")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--widths", action="store_true", help="the element-width and mask-shape sweep instead of the per-phase sweep")
    ap.add_argument("--shapes", action="store_true", help="tile size against mask shape, every width")
    ap.add_argument("--json", default=None, help="also write the rows as JSON")
    args = ap.parse_args()
    if args.shapes:
        args.log2n = 28 if args.log2n is None else args.log2n
        rows = shapes_sweep(args)
        if args.json:
            with open(args.json, "w") as f:
                json.dump(rows, f, indent=1)
        return
    if args.widths:
        args.log2n = 29 if args.log2n is None else args.log2n
        rows = widths_sweep(args)
        if args.json:
            with open(args.json, "w") as f:
                json.dump(rows, f, indent=1)
        return
    args.log2n = 28 if args.log2n is None else args.log2n
    n = 1 << args.log2n
    g = torch.Generator(device="cuda").manual_seed(11)
    data = torch.rand(n, device="cuda", generator=g)
    out = torch.empty_like(data)
    masks = {f"bernoulli_{p}": csx.gen_mask_bernoulli(n, p, seed=7) for p in (0.01, 0.1, 0.5, 0.9, 0.99)}
    masks["runs_256"] = csx.gen_mask_runs(n, 256, seed=7)
    count = torch.empty(1, dtype=torch.int64, device="cuda")
    rows = []
    print("mask;chunk_len;form;time_popc_ms;time_scan_ms;time_writeout_ms;time_total_ms;selected;gbps;pct_peak", flush=True)

    def emit(mname, chunk, form, tp, ts, tw, total, k):
        gbps = (4 * n + n / 8 + 4 * k) / (total * 1e-3) / 1e9
        row = {"mask": mname, "chunk_len": chunk, "form": form, "popc_ms": tp, "scan_ms": ts, "writeout_ms": tw, "total_ms": round(total, 5),
               "selected": k, "GB/s": round(gbps, 1), "pct_peak": round(gbps / 80.0, 2)}
        rows.append(row)
        f = lambda v: "" if v is None else f"{v:.5f}"
        print(f"{mname};{chunk};{form};{f(tp)};{f(ts)};{f(tw)};{total:.5f};{k};{gbps:.1f};{gbps / 80.0:.2f}", flush=True)

    for mname, mask in masks.items():
        for chunk in (256, 1024, 4096):
            counts = csx.popc(mask, n, chunk)
            offsets, total = csx.scan(counts)
            torch.cuda.synchronize()
            k = int(total.item())
            expect = int(counts.to(torch.int64).sum().item())
            if k != expect:
                print(f"count mismatch {mname} chunk {chunk}: scan {k}, popcount {expect}", file=sys.stderr)
                sys.exit(1)
            scan_ws = csx.alloc_workspace(csx.scan_workspace_size(counts.numel()))
            tp = median_ms(lambda: csx.popc(mask, n, chunk, counts=counts), args.reps)
            ts = median_ms(lambda: csx.scan(counts, offsets=offsets, total=total, workspace=scan_ws), args.reps)
            for vname, v in VARIANTS.items():
                reps = args.reps if v != csx.WRITEOUT_THREAD_PER_CHUNK else max(2, args.reps // 5)
                tw = median_ms(lambda: csx.writeout(data, mask, offsets, n, chunk, variant=v, counts=counts, out=out), reps)
                emit(mname, chunk, f"phases_{vname}", round(tp, 5), round(ts, 5), round(tw, 5), tp + ts + tw, k)
            ws = csx.alloc_workspace(csx.workspace_size(n, 4, chunk))
            for pname, pl in PIPELINES.items():
                tt = median_ms(lambda: csx.compact(data, mask, n=n, chunk_len=chunk, pipeline=pl, out=out, count=count, workspace=ws, host_count=False),
                               args.reps)
                if int(count.item()) != k:
                    print(f"count mismatch {mname} chunk {chunk} {pname}", file=sys.stderr)
                    sys.exit(1)
                emit(mname, chunk, pname, None, None, None, tt, k)
            del ws, scan_ws
        sel = k / n
        wsa = csx.alloc_workspace(csx.auto_workspace_size(n, 4))
        ta = median_ms(lambda: csx.compact_auto(data, mask, sel, n=n, out=out, count=count, workspace=wsa, host_count=False), args.reps)
        emit(mname, 0, "tuned", None, None, None, ta, k)
        del wsa
    if args.json:
        with open(args.json, "w") as f:
            json.dump(rows, f, indent=1)


if __name__ == "__main__":
    main()
