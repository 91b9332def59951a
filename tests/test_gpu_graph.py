"""Single-column entry points captured into a graph: replays on new inputs
written into the captured buffers give the oracle's results every time
(the graph holds kernels only: look-back state and counts are reset by
kernels, not memset nodes)."""
import numpy as np
import pytest
import torch

import gpu_compressstore2_amd as csx

pytestmark = pytest.mark.gpu

N = 300_007


def capture(fn):
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        fn(torch.cuda.current_stream())  # warm-up outside the graph
        torch.cuda.current_stream().synchronize()
        with torch.cuda.graph(g, stream=s):
            fn(s)
    torch.cuda.synchronize()
    return g


@pytest.mark.parametrize("form", ["three_phase", "two_pass", "fused", "sparse", "single", "tuned"])
def test_compact_graph_replays(O, form):
    data = torch.zeros(N, dtype=torch.int32, device="cuda")
    mask = torch.zeros((N + 7) // 8, dtype=torch.uint8, device="cuda")
    out = torch.empty(N, dtype=torch.int32, device="cuda")
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    pipes = {"three_phase": csx.PIPELINE_THREE_PHASE, "two_pass": csx.PIPELINE_TWO_PASS, "fused": csx.PIPELINE_FUSED, "sparse": csx.PIPELINE_SPARSE,
             "single": csx.PIPELINE_SINGLE}
    if form == "tuned":
        ws = csx.alloc_workspace(csx.auto_workspace_size(N, 4))
        g = capture(lambda s: csx.compact_auto(data, mask, 0.4, n=N, out=out, count=count, workspace=ws, stream=s, host_count=False))
    else:
        ws = csx.alloc_workspace(csx.workspace_size(N, 4, 1024))
        g = capture(lambda s: csx.compact(data, mask, n=N, pipeline=pipes[form], out=out, count=count, workspace=ws, stream=s, host_count=False))
    for rep in range(4):
        v = O.fill_values(N, np.uint32, seed=70 + rep)
        m = O.mask_bernoulli(N, 0.1 + 0.25 * rep, seed=80 + rep)
        expect, k = O.compact(v, m, N)
        data.copy_(torch.from_numpy(v.view(np.int32)))
        mask.copy_(torch.from_numpy(m))
        out.fill_(-1)
        count.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert int(count.item()) == k, (form, rep)
        assert (out[:k].cpu().numpy().view(np.uint32) == expect).all(), (form, rep)
        assert (out[k:].cpu().numpy() == -1).all(), (form, rep)


def test_predicate_graph_replays(O):
    col = torch.zeros(N, dtype=torch.float32, device="cuda")
    out = torch.empty(N, dtype=torch.float32, device="cuda")
    mask = torch.empty((N + 7) // 8, dtype=torch.uint8, device="cuda")
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    ws = csx.alloc_workspace(csx.predicate_workspace_size(N, 8))
    g = capture(lambda s: csx.compact_predicate(col, ">", 0.5, out=out, mask=mask, count=count, workspace=ws, stream=s, host_count=False))
    for rep in range(4):
        v = O.fill_values(N, np.float32, seed=90 + rep)
        m, k = O.predicate_mask_typed(v, 0, 0.5)
        expect, _ = O.compact(v, m, N)
        col.copy_(torch.from_numpy(v))
        out.fill_(-1.0)
        count.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert int(count.item()) == k, rep
        assert (mask.cpu().numpy()[: (N + 7) // 8] == m).all(), rep
        assert (out[:k].cpu().numpy().view(np.uint32) == expect.view(np.uint32)).all(), rep
