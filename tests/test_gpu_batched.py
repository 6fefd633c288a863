"""GPU parity of the batched persistent kernel, its graph replay and the
autotuner.  Bar: maximum bit-exact against the FP32-rounded oracle;
euclidean and manhattan max relative error <= 1e-5 against the FP64 oracle.
"""
import numpy as np
import pytest

from oracle import oracle as orc

torch = pytest.importorskip("torch")

import gputools_amd as gt  # noqa: E402
from gputools_amd import workloads as wl  # noqa: E402
from gputools_amd.autotune import tune_batched  # noqa: E402
from gputools_amd.batched import Batch  # noqa: E402

pytestmark = pytest.mark.gpu


def check_set(got, x, method):
    want = orc.dist(x, method)
    if method == "maximum":
        np.testing.assert_array_equal(got, want.astype(np.float32))
    else:
        w = want[want > 0]
        g = got[want > 0].astype(np.float64)
        assert np.all(got[want == 0] == 0)
        if w.size:
            assert np.max(np.abs(g - w) / w) <= 1e-5


def ragged_sets(count, seed, hi=300):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(0, hi, count)
    sizes[:4] = [0, 1, 2, 129]
    dims = rng.choice([3, 16, 33, 128], count)
    return [rng.standard_normal((int(n), int(d))).astype(np.float32) for n, d in zip(sizes, dims)]


@pytest.mark.parametrize("method", ["euclidean", "manhattan", "maximum"])
@pytest.mark.parametrize("tile", [1, 2, 3])
def test_batched_vs_oracle(method, tile):
    sets = ragged_sets(40, seed=tile * 10 + len(method))
    b = Batch(sets, method=method)
    b.run(tile)
    torch.cuda.synchronize()
    for s, x in enumerate(sets):
        check_set(b.result(s).cpu().numpy(), x, method)


def test_graph_replays_match_eager():
    sets = ragged_sets(30, seed=5)
    b = Batch(sets, method="euclidean")
    b.run(2)
    torch.cuda.synchronize()
    eager = b.out.clone()
    g = b.capture(2)
    for rep in range(3):
        b.out.fill_(-1.0)                      # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(b.out, eager), rep
    # the graph reads the current input: change it and replay
    b.x.mul_(2.0)
    g.replay()
    torch.cuda.synchronize()
    torch.testing.assert_close(b.out, 2.0 * eager, rtol=0, atol=0)


def test_autotune_picks_a_valid_tile_and_caches(tmp_path, monkeypatch):
    monkeypatch.setenv("GPUTOOLS_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    b = Batch(ragged_sets(50, seed=6, hi=600), method="manhattan")
    t = tune_batched(b, reps=2)
    assert t in (1, 2, 3)
    assert (tmp_path / "tune.json").exists()
    assert tune_batched(b, reps=2) == t
    out = gt.batched.dist_batched([b.set_points(s).cpu().numpy() for s in range(5)], "manhattan")
    assert len(out) == 5


def test_batched_4096_sets():
    """The batching workload: 4096 sets, sizes log-uniform in [32, 2048],
    d in {16, 32, 64, 128}, N(0,1); one launch vs per-set launches."""
    sizes, dims = wl.batched_sets(4096, seed=5)
    rng = np.random.default_rng(55)
    x = torch.from_numpy(rng.standard_normal(int((sizes * dims).sum()), dtype=np.float32)).cuda()
    for method in ("euclidean", "manhattan"):
        print("case batched 4096 sets", method, flush=True)
        b = Batch(x=x, sizes=sizes, dims=dims, method=method)
        b.run(2)
        torch.cuda.synchronize()
        batched = b.out.clone()
        b.out.zero_()
        b.run_per_set()
        torch.cuda.synchronize()
        # same values as one gt_dist call per set (FP32, within rounding of the
        # chunking) and as the FP64 oracle on a sample of sets
        torch.testing.assert_close(batched, b.out, rtol=2e-6, atol=0)
        for s in rng.choice(4096, 25, replace=False):
            check_set(batched[int(b.out_off[s]):int(b.out_off[s]) +
                              int(sizes[s] * (sizes[s] - 1) // 2)].cpu().numpy(),
                      b.set_points(int(s)).cpu().numpy(), method)
