"""gt_dist_ws captured in a HIP graph (torch.cuda.CUDAGraph).

A caller that owns the workspace can capture the whole distance call: the
packed binary path (byte rows, matrix-core tiles with their staggered start)
and the split-precision euclidean path (column statistics and centres, the
split, the norms, the Gram tiles and their exact pass).  The library records
no memset nodes.  Each graph is replayed several times with eager work
between the replays and must reproduce the eager result bit for bit; a
replay after the input changes must equal an eager call on the new input.
"""
import pytest

torch = pytest.importorskip("torch")

from gputools_amd import _lib  # noqa: E402
from gputools_amd.dist import workspace_size  # noqa: E402
from gputools_amd import workloads as wl  # noqa: E402

pytestmark = pytest.mark.gpu

CASES = [("binary", "fp32", 12000, 100), ("euclidean", "bf16x3", 3000, 64),
         ("euclidean", "fp16x3", 2500, 300)]


def points(method, n, k, seed):
    return wl.bernoulli(n, k, seed=seed) if method == "binary" else wl.mixture(n, k, seed=seed)


@pytest.mark.parametrize("method,precision,n,k", CASES)
def test_dist_ws_graph_replays_match_eager(method, precision, n, k):
    print("case graph", method, precision, n, k, flush=True)
    L = _lib.lib()
    m = {"binary": _lib.BINARY, "euclidean": _lib.EUCLIDEAN}[method]
    prec = _lib.PRECISIONS[precision]
    xt = torch.from_numpy(points(method, n, k, seed=n)).cuda()
    ws_bytes = workspace_size(n, k, method, precision)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device="cuda")
    out = torch.empty(n * (n - 1) // 2, dtype=torch.float32, device="cuda")
    s = torch.cuda.Stream()

    def call(dst):
        st = L.gt_dist_ws(xt.data_ptr(), n, k, k, m, 2.0, prec, _lib.LAYOUT_PACKED, dst.data_ptr(), n,
                          ws.data_ptr(), ws_bytes, s.cuda_stream)
        assert st == _lib.GT_OK, _lib.last_error()

    with torch.cuda.stream(s):
        call(out)                      # eager, on the capture stream
    s.synchronize()
    eager = out.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        call(out)
    torch.cuda.synchronize()
    for rep in range(3):
        out.fill_(-1.0)                # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out.view(torch.int32), eager.view(torch.int32)), rep

    # the graph reads the current input
    xt.copy_(torch.from_numpy(points(method, n, k, seed=n + 1)).cuda())
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    replayed = out.clone()
    with torch.cuda.stream(s):
        call(out)
    s.synchronize()
    assert torch.equal(replayed.view(torch.int32), out.view(torch.int32))
    assert not torch.equal(replayed.view(torch.int32), eager.view(torch.int32))
