"""Blocked Bloom visited filter on the GPU (approximate; for recall-judged searches).

A Bloom filter's false positives skip unvisited vertices, so ids may differ from the exact
search; what is pinned: (1) a filter large enough to have no false positives on a query's
few thousand ids gives the exact ids for nearly every query; (2) the reference-sized filter
(auto_tune_bloom) keeps recall within a small delta of the exact search while its LDS slice
is smaller than the exact table's; (3) a batch tier with the filter equals the plain search
with the same filter (the same search path, so the same ids), graph replays included.
"""
import numpy as np
import pytest

import song_amd
from song_amd import datasets, groundtruth, tier

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mix():
    base, _ = datasets.mixture_l2(n_base=50000, n_query=1, dim=64, n_centers=200, seed=21)
    query = datasets.mixture_l2_queries(1000, dim=64, n_centers=200, seed=21, stream=1)
    graph = song_amd.build_graph(base, "l2", threads=8)
    ix = song_amd.Index(base, graph, "l2")
    _, gt_d = groundtruth.knn_fp64(base, query, 10, "l2")
    return base, query, ix, gt_d


def test_large_filter_gives_the_exact_ids(mix):
    base, query, ix, _ = mix
    exact, _, _ = ix.search(query, 128, 10, mode=song_amd.MODE_FP32)
    bl, _, _ = ix.search(query, 128, 10, mode=song_amd.MODE_FP32, bloom_k=3, hash_log2=15)
    same = (bl == exact).all(1).mean()
    print(f"[bloom] 2^15 words, k 3: {same:.4f} of the queries equal the exact search", flush=True)
    assert same >= 0.99


@pytest.mark.parametrize("pq", [64, 128, 256])
def test_reference_sized_filter_recall_and_lds(mix, pq):
    import torch

    base, query, ix, gt_d = mix
    exact, _, fe = ix.search(query, pq, 10, mode=song_amd.MODE_FP32)
    bl, _, fb = ix.search(query, pq, 10, mode=song_amd.MODE_FP32, bloom_k=-1)
    r_e = groundtruth.recall_at_k(exact, gt_d, 10, base, query)
    r_b = groundtruth.recall_at_k(bl, gt_d, 10, base, query)
    hlog, k = song_amd.bloom_size(pq)
    q = torch.from_numpy(query).cuda()
    be = song_amd.Batch(ix, [tier(pq)], len(query), mode=song_amd.MODE_FP32)
    bb = song_amd.Batch(ix, [tier(pq, bloom_k=-1)], len(query), mode=song_amd.MODE_FP32)
    (ge, _, le), (gb, _, lb) = be.grid(0), bb.grid(0)
    print(f"[bloom] pq {pq}: 2^{hlog} words k {k}: recall@10 {r_b:.4f} vs exact {r_e:.4f}; "
          f"LDS per wave {lb} vs {le} bytes, grid {gb} vs {ge} groups", flush=True)
    assert r_b >= r_e - 0.02
    assert lb < le and gb >= ge
    be.close()
    bb.close()


def test_batch_tier_with_filter_equals_plain_search_with_filter(mix):
    import torch

    base, query, ix, _ = mix
    q = torch.from_numpy(query).cuda()
    nq, k = q.shape[0], 10
    pq_h = np.where(np.arange(nq) % 2 == 0, 64, 192).astype(np.int32)
    pq_d = torch.from_numpy(pq_h).cuda()
    k_d = torch.full((nq,), k, dtype=torch.int32, device=q.device)
    want = torch.full((nq, k), -1, dtype=torch.int32, device=q.device)
    for p in (64, 192):
        sel = torch.from_numpy(np.nonzero(pq_h == p)[0]).cuda()
        out = torch.empty((sel.numel(), k), dtype=torch.int32, device=q.device)
        ix.search_device(q[sel].contiguous(), p, k, out, mode=song_amd.MODE_FP32, bloom_k=-1)
        want[sel] = out
    b = song_amd.Batch(ix, [tier(64, bloom_k=-1), tier(192, bloom_k=-1, waves_per_group=2)], nq,
                       mode=song_amd.MODE_FP32)
    ids = torch.full((nq, k), -7, dtype=torch.int32, device=q.device)
    fl = torch.empty(nq, dtype=torch.int32, device=q.device)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    for rep in range(3):
        ids.fill_(-7)
        b.run(q, ids, fl, pq_arr=pq_d, k_arr=k_d, graph=rep > 0, stream=side.cuda_stream)
        torch.cuda.synchronize()
        assert torch.equal(ids, want), f"run {rep}"
    b.close()


@pytest.mark.parametrize("pq", [64, 128])
def test_multiprobe_tier_recall(mix, pq):
    """Two expansions per step (the next queue head's neighbours evaluated in the same gather)
    on the sub-group kernel: an approximate expansion order, so recall is what is pinned; the
    exact sub-group tier equals the plain search."""
    import torch

    base, query, ix, gt_d = mix
    q = torch.from_numpy(query).cuda()
    nq, k = q.shape[0], 10
    plain, _, _ = ix.search(query, pq, k, mode=song_amd.MODE_FP32)
    r_e = groundtruth.recall_at_k(plain, gt_d, k, base, query)
    out = {}
    for mp in (1, 2):
        b = song_amd.Batch(ix, [tier(pq, queries_per_wave=4, multiprobe=mp)], nq, mode=song_amd.MODE_FP32)
        ids = torch.full((nq, k), -7, dtype=torch.int32, device=q.device)
        fl = torch.empty(nq, dtype=torch.int32, device=q.device)
        st = torch.empty((nq, 8), dtype=torch.int32, device=q.device)
        b.run(q, ids, fl, pq=pq, k=k, stats=st)
        torch.cuda.synchronize()
        b.close()
        out[mp] = (ids.cpu().numpy().astype(np.int64), st.cpu().numpy())
    assert np.array_equal(out[1][0], plain)
    r_m = groundtruth.recall_at_k(out[2][0], gt_d, k, base, query)
    print(f"[multiprobe] pq {pq}: recall@10 {r_m:.4f} vs exact {r_e:.4f}; expansions "
          f"{out[2][1][:, 0].mean():.1f} vs {out[1][1][:, 0].mean():.1f}", flush=True)
    assert r_m >= r_e - 0.02
