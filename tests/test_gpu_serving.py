"""Serving: the arena budget and the continuous-batching server; every
request's result is the same alone (csx_compact) and in the mix."""
import numpy as np
import pytest
import torch

import gpu_compressstore2_amd as csx
import oracle
from gpu_compressstore2_amd.serving import Arena, Server

pytestmark = pytest.mark.gpu

NP = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
VIEW = {1: np.uint8, 2: np.int16, 4: np.int32, 8: np.int64}


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(VIEW[a.dtype.itemsize])).cuda()


def test_arena_first_fit_and_coalescing():
    a = Arena(1 << 20)
    p = [a.alloc(1000) for _ in range(4)]
    assert all(x is not None for x in p) and a.used() == 4 * 1024
    assert a.alloc(1 << 20) is None  # no room
    a.free(p[1])
    a.free(p[2])
    q = a.alloc(2048)  # the two freed neighbours merged
    assert q == p[1]
    for x in (p[0], q, p[3]):
        a.free(x)
    assert a.used() == 0
    assert a.alloc(1 << 20) is not None  # everything coalesced back
    a.close()


def test_results_same_alone_and_in_the_mix(O):
    rng = np.random.default_rng(11)
    reqs = []
    for i in range(80):
        w = int(rng.choice([1, 2, 4, 8]))
        n = int(rng.integers(0, 300000)) if i % 7 else int(rng.integers(0, 70))
        data = O.fill_values(n, NP[w], seed=900 + i)
        mask = O.mask_bernoulli(n, float(rng.uniform(0, 1)), seed=1900 + i)
        expect, k = O.compact(data, mask, n)
        d_in = dev(data) if n else torch.zeros(1, dtype=dev(np.zeros(1, NP[w])).dtype, device="cuda")
        d_mask = dev(mask) if n else torch.zeros(1, dtype=torch.uint8, device="cuda")
        reqs.append((w, n, d_in, d_mask, expect, k))
    # alone
    alone = []
    for w, n, d_in, d_mask, expect, k in reqs:
        out, kk = csx.compact(d_in, d_mask, n=n)
        alone.append((out[:kk].clone(), kk))
    torch.cuda.synchronize()
    # in the mix: arrivals in waves while earlier batches are still running
    srv = Server(streams=3, max_batch=7, max_batch_elements=1 << 22)
    outs = [torch.full((max(r[1], 1) + 8,), -1, dtype=r[2].dtype, device="cuda") for r in reqs]
    done = {}
    nxt = 0
    while len(done) < len(reqs):
        for _ in range(int(rng.integers(1, 12))):
            if nxt < len(reqs):
                w, n, d_in, d_mask, _, _ = reqs[nxt]
                srv.submit(nxt, d_in, d_mask, outs[nxt], n=n)
                nxt += 1
        srv.step()
        for rid, cnt, lat in srv.poll():
            assert lat >= 0
            done[rid] = cnt
    srv.close()
    torch.cuda.synchronize()
    for i, (w, n, d_in, d_mask, expect, k) in enumerate(reqs):
        assert done[i] == k == alone[i][1], i
        got = outs[i].cpu().numpy().view(NP[w])
        assert (got[:k] == expect).all(), i
        assert (got[:k] == alone[i][0].cpu().numpy().view(NP[w])).all(), i
        assert (got[k:].view(np.uint8) == 0xFF).all(), i


def test_server_rejects_bad_requests():
    srv = Server(streams=1, max_batch=2, max_batch_elements=100)
    x = torch.zeros(200, dtype=torch.int32, device="cuda")
    m = torch.zeros(25, dtype=torch.uint8, device="cuda")
    with pytest.raises(csx.CsxError):
        srv.submit(0, x, m, x, n=200)  # larger than max_batch_elements
    srv.close()
