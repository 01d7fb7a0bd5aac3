"""The no-allocation entry point (song_search_device_noalloc) against the allocating one
(song_search_device_ex with retry = 0) through the C ABI: ids, distances, flags and stats
identical on two calls in a row (the second with fewer queries, a smaller pq and a query
subset), through a captured graph replayed between eager work, and the size and mode
limits rejected with an error."""
import ctypes

import numpy as np
import pytest

import golden_cases
import song_amd

pytestmark = pytest.mark.gpu


def _case(name):
    import torch

    meta, base, query, _, graph = golden_cases.load(name)
    if graph is None:
        graph = song_amd.build_graph(base, meta["metric"], threads=8)
    ix = song_amd.Index(base, graph, meta["metric"])
    return ix, torch.from_numpy(np.ascontiguousarray(query)).cuda()


def _outputs(nq, k, stats=True):
    import torch

    dev = torch.device("cuda")
    return (torch.full((nq, k), -9, dtype=torch.int32, device=dev),
            torch.full((nq, k), -1.0, dtype=torch.float32, device=dev),
            torch.full((nq,), -9, dtype=torch.int32, device=dev),
            torch.full((nq, 8), -9, dtype=torch.int32, device=dev) if stats else None)


def _compare(ix, q, pq, k, mode, subset=None):
    import torch

    nq = q.shape[0]
    want = _outputs(nq, k)
    got = _outputs(nq, k)
    ix.search_device(q, pq, k, want[0], want[1], want[2], want[3], mode=mode, retry=0, subset=subset)
    song_amd.search_device_noalloc(ix, q, pq, k, got[0], got[1], got[2], got[3], mode=mode, subset=subset)
    torch.cuda.synchronize()
    for name, a, b in zip(("ids", "dists", "flags", "stats"), got, want):
        assert torch.equal(a, b), f"{name} differ (pq {pq}, k {k}, nq {nq})"
    return got


@pytest.mark.parametrize("name,mode", [("mix_l2_20k", song_amd.MODE_FP32), ("c1_cos_15k", song_amd.MODE_EXACT),
                                       ("tie_l2_3k", song_amd.MODE_EXACT)])
def test_noalloc_equals_allocating_search_on_two_calls(name, mode):
    import torch

    print(f"[noalloc] {name} mode {mode}", flush=True)
    ix, q = _case(name)
    nq = q.shape[0]
    _compare(ix, q, 100, 10, mode)
    # second call: fewer queries, a smaller pq, every other query through a subset
    q2 = q[: max(1, nq // 3)].contiguous()
    sub = torch.arange(0, q2.shape[0], 2, dtype=torch.int32, device=q.device)
    _compare(ix, q2, 40, 5, mode, subset=sub)


def test_noalloc_replays_in_a_captured_graph():
    import torch

    ix, q = _case("mix_l2_20k")
    nq, pq, k = q.shape[0], 64, 10
    want = _outputs(nq, k, stats=False)
    ix.search_device(q, pq, k, want[0], want[1], want[2], mode=song_amd.MODE_FP32, retry=0)
    got = _outputs(nq, k, stats=False)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        song_amd.search_device_noalloc(ix, q, pq, k, got[0], got[1], got[2], mode=song_amd.MODE_FP32)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        song_amd.search_device_noalloc(ix, q, pq, k, got[0], got[1], got[2], mode=song_amd.MODE_FP32)
    for rep in range(3):
        got[0].fill_(-3)
        x = torch.randn(256, 256, device=q.device)
        (x @ x).sum().item()
        g.replay()
        torch.cuda.synchronize()
        for name, a, b in zip(("ids", "dists", "flags"), got, want):
            assert torch.equal(a, b), f"replay {rep}: {name} differ"


def test_noalloc_rejects_sizes_and_modes_beyond_its_limits():
    ix, q = _case("tiny_l2_500")
    ids, ds, fl, _ = _outputs(q.shape[0], 10, stats=False)
    with pytest.raises(song_amd.SongError, match="no-allocation array"):
        song_amd.search_device_noalloc(ix, q, 100, 10, ids, ds, fl, mode=song_amd.MODE_EXACT, slack=400000)
    with pytest.raises(song_amd.SongError, match="mode EXACT or FP32"):
        song_amd.search_device_noalloc(ix, q, 100, 10, ids, ds, fl, mode=song_amd.MODE_AUTO)
