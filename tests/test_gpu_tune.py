"""Every configuration the tuner can choose gives the oracle's result; the
tuners pick one of them and the tuned entry points stay bit-exact."""
import numpy as np
import pytest
import torch

import gpu_compressstore2_amd as csx

pytestmark = pytest.mark.gpu

NP = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
VIEW = {1: torch.uint8, 2: torch.int16, 4: torch.int32, 8: torch.int64}


def candidates(w):
    out = []
    for chunk in (256, 1024, 4096):
        for vpt in (2, 4, 8, 16):
            if vpt != 16 or w >= 2:
                out.append(csx.Config(csx.PIPELINE_THREE_PHASE, chunk, vpt, 0, 256))
    for block in (128, 512):
        for vpt in (2, 4, 8, 16):
            if vpt != 16 or w >= 2:
                out.append(csx.Config(csx.PIPELINE_THREE_PHASE, 1024, vpt, 0, block))
    out.append(csx.Config(csx.PIPELINE_TWO_PASS, 1024, 8, 0))
    out.append(csx.Config(csx.PIPELINE_FUSED, 1024, 8, 0))
    out.append(csx.Config(csx.PIPELINE_SPARSE, 1024, 8, 0))
    out.append(csx.Config(csx.PIPELINE_SPARSE, 4096, 8, 0))
    for vpt in (2, 4, 8):
        out.append(csx.Config(csx.PIPELINE_SINGLE, 1024, vpt, 0))
    return out


@pytest.fixture(autouse=True)
def clean_table():
    """An empty table for the test; the shipped table again afterwards, so
    that later tests run the tuned configurations whatever the order."""
    csx.tune_clear()
    yield
    csx.tune_clear()
    csx.tune_load(csx.TUNED_TABLE)


@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_every_candidate_matches_oracle(O, w):
    for n, p in ((1, 1.0), (4099, 0.3), (70001, 0.5), (300007, 0.9), (1 << 20, 0.02)):
        data = O.fill_values(n, NP[w], seed=w * 1000 + n)
        for order in (0, 1):
            mask = O.mask_bernoulli(n, p, seed=n + order, bit_order=order)
            expect, k = O.compact(data, mask, n, bit_order=order)
            d = torch.from_numpy(data.view(np.uint8)).cuda()
            # element-aligned but not 16-byte aligned input on the LSB runs
            if order == 1:
                buf = torch.zeros(d.numel() + w, dtype=torch.uint8, device="cuda")
                buf[w:] = d
                d = buf[w:]
            din = d.view(VIEW[w])
            m = torch.from_numpy(mask).cuda()
            for cfg in candidates(w):
                out, kk = csx.compact_config(din, m, cfg, n=n, bit_order=order)
                assert kk == k, (w, n, order, cfg.as_dict())
                got = out[:k].cpu().numpy().view(NP[w])
                assert (got == expect[:k]).all(), (w, n, order, cfg.as_dict())


def test_autotune_then_auto(O):
    cfg, ms = csx.autotune(4, 1 << 22, 0.5, reps=3)
    assert ms > 0 and cfg.vpt in (2, 4, 8, 16) and cfg.pipeline in (0, 1, 2, 4)
    got = csx.tune_lookup(4, 1 << 22, 0.5)
    assert got.as_dict() == cfg.as_dict()
    n = (1 << 22) + 77
    data = O.fill_values(n, np.uint32, seed=5)
    mask = O.mask_bernoulli(n, 0.5, seed=6)
    expect, k = O.compact(data, mask, n)
    out, kk = csx.compact_auto(torch.from_numpy(data.view(np.int32)).cuda(), torch.from_numpy(mask).cuda(), 0.5, n=n)
    assert kk == k and (out[:k].cpu().numpy().view(np.uint32) == expect).all()


@pytest.mark.parametrize("w", [1, 4, 8])
def test_autotune_on_callers_column(O, w):
    """Tuning on the caller's own column and mask (runs of mean 256, LSB-first
    bits, ragged length): the count and the output left in `out` are the
    oracle's, the fastest candidate is recorded under (w, n, k/n), and the
    tuned call then uses it and stays bit-exact."""
    n = 300007
    data = O.fill_values(n, NP[w], seed=77 + w)
    mask = O.mask_geometric_runs(n, 256, seed=78 + w, bit_order=1)
    expect, k = O.compact(data, mask, n, bit_order=1)
    din = torch.from_numpy(data.view(np.uint8)).cuda().view(VIEW[w])
    m = torch.from_numpy(mask).cuda()
    cfg, ms, kk = csx.autotune_inputs(din, m, n=n, bit_order=1, reps=3)
    assert kk == k and ms > 0
    got = csx.tune_lookup(w, n, k / n)
    assert got.as_dict() == cfg.as_dict()
    out, kk = csx.compact_auto(din, m, k / n, n=n, bit_order=1)
    assert kk == k and (out[:k].cpu().numpy().view(NP[w]) == expect).all()
    out_after_tuning = torch.empty_like(din)
    csx.autotune_inputs(din, m, n=n, bit_order=1, reps=1, out=out_after_tuning)
    assert (out_after_tuning[:k].cpu().numpy().view(NP[w]) == expect).all()


def test_batch_configs_and_autotune(O):
    rng = np.random.default_rng(21)
    lens = [int(x) for x in np.exp(rng.uniform(np.log(1), np.log(200000), 120))] + [0, 1, 31]
    ins, masks, outs, exp = [], [], [], []
    for i, n in enumerate(lens):
        data = O.fill_values(n, np.uint32, seed=300 + i)
        mask = O.mask_bernoulli(n, float(rng.uniform(0.01, 0.99)), seed=400 + i)
        e, k = O.compact(data, mask, n)
        exp.append((e, k))
        ins.append(torch.from_numpy(data.view(np.int32)).cuda() if n else torch.zeros(1, dtype=torch.int32, device="cuda"))
        masks.append(torch.from_numpy(mask).cuda() if n else torch.zeros(1, dtype=torch.uint8, device="cuda"))
        outs.append(torch.zeros(max(n, 1), dtype=torch.int32, device="cuda"))
    total = sum(lens)
    table = csx.segment_table(ins, masks, outs, lens)
    ws = csx.alloc_workspace(csx.batch_workspace_size(len(lens), total, 4))
    counts = torch.zeros(len(lens), dtype=torch.int64, device="cuda")

    def check():
        c = counts.cpu().numpy()
        for i, (e, k) in enumerate(exp):
            assert c[i] == k, i
            assert (outs[i][:k].cpu().numpy().view(np.uint32) == e).all(), i

    for vpt in (2, 4, 8):
        for cb in (1, 64, 4096):
            for o in outs:
                o.zero_()
            csx.compact_batch_config(table, 4, total, csx.Config(1, 0, vpt, cb), counts=counts, workspace=ws)
            torch.cuda.synchronize()
            check()
    cfg, ms = csx.autotune_batch(table, 4, total, counts, ws, reps=2)
    assert cfg.vpt in (2, 4, 8) and ms > 0
    for o in outs:
        o.zero_()
    csx.compact_batch(table, 4, total, counts=counts, workspace=ws)  # now with the tuned configuration
    torch.cuda.synchronize()
    check()
