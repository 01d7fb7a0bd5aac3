"""Precision form on the GPU: BF16 / FP8-E4M3 rows with per-row scales, FP32
accumulation, exact FP32 re-rank of the final pq list.

Pins: the GPU codes equal the numpy restatement of the formats bit for bit;
BF16 with power-of-two scales is lossless on integer data, so its ids equal
the plain search's; after the re-rank every returned distance equals the exact
FP32 distance of its id; recall deltas against the plain form are reported.
"""
import numpy as np
import pytest

import lowprec
import pyoracle
import song_amd
from song_amd import datasets, groundtruth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mix():
    base, query = datasets.mixture_l2(n_base=20000, n_query=300, dim=128, seed=2)
    graph = pyoracle.build(base, "l2")
    ix = song_amd.Index(base, graph, "l2")
    ix.add_storage(song_amd.STORE_BF16)
    ix.add_storage(song_amd.STORE_FP8)
    return base, query, graph, ix


@pytest.fixture(scope="module")
def factor():
    base, query = datasets.factor_l2(n_base=8000, n_query=200, dim=96, seed=3)
    graph = song_amd.build_graph(base, "l2", threads=8)
    ix = song_amd.Index(base, graph, "l2")
    ix.add_storage(song_amd.STORE_BF16)
    ix.add_storage(song_amd.STORE_FP8)
    return base, query, graph, ix


@pytest.mark.parametrize("storage,name", [(song_amd.STORE_BF16, "bf16"), (song_amd.STORE_FP8, "fp8")])
def test_codes_match_format_restatement(factor, storage, name):
    base, _, _, ix = factor
    codes, scales = ix.get_storage(storage)
    rc, rs, _ = lowprec.quantize(base, name)
    assert np.array_equal(scales, rs)
    got = codes.view(np.uint16)[:, : base.shape[1]] if name == "bf16" else codes[:, : base.shape[1]]
    assert np.array_equal(got, rc)


def test_bf16_lossless_on_integer_data_equals_plain(mix):
    base, query, graph, ix = mix
    plain, pd, _ = ix.search(query, 100, 100, mode=song_amd.MODE_FP32)
    for rerank in (0, 1):
        ids, ds, _ = ix.search(query, 100, 100, storage=song_amd.STORE_BF16, rerank=rerank)
        assert np.array_equal(ids, plain)
        assert np.array_equal(ds, pd)


@pytest.mark.parametrize("storage", [song_amd.STORE_BF16, song_amd.STORE_FP8])
def test_rerank_gives_exact_fp32_distances(factor, storage):
    base, query, graph, ix = factor
    ids, ds, _ = ix.search(query, 64, 20, storage=storage)
    assert (ids >= 0).all()
    exact = np.array([[pyoracle.distance("l2", base[j], query[i]) for j in row] for i, row in enumerate(ids)])
    assert np.allclose(ds, exact, rtol=1e-5, atol=0)
    assert (np.diff(ds, axis=1) >= 0).all()


def test_precision_recall_delta_is_small(factor):
    base, query, graph, ix = factor
    gt_i, gt_d = groundtruth.knn_fp64(base, query, 10, "l2")
    plain, _, _ = ix.search(query, 64, 10, mode=song_amd.MODE_FP32)
    r_plain = groundtruth.recall_at_k(plain, gt_d, 10, base, query)
    for storage in (song_amd.STORE_BF16, song_amd.STORE_FP8):
        ids, _, _ = ix.search(query, 64, 10, storage=storage)
        r = groundtruth.recall_at_k(ids, gt_d, 10, base, query)
        print(f"[precision] storage={storage} recall@10={r:.4f} plain={r_plain:.4f}", flush=True)
        assert r >= r_plain - 0.05


@pytest.mark.parametrize("storage", [song_amd.STORE_BF16, song_amd.STORE_FP8])
@pytest.mark.parametrize("qpw", [2, 4])
def test_several_queries_per_wave_low_precision_equals_plain(factor, storage, qpw):
    """Sub-group search on BF16 / FP8 rows with the exact FP32 re-rank (float data, where the
    re-rank changes distances) returns the plain search's ids and distances."""
    import torch

    base, query, graph, ix = factor
    q = torch.from_numpy(query).cuda()
    nq, pq, k = q.shape[0], 64, 20
    want_i = torch.empty((nq, k), dtype=torch.int32, device=q.device)
    want_d = torch.empty((nq, k), dtype=torch.float32, device=q.device)
    ix.search_device(q, pq, k, want_i, want_d, storage=storage, mode=song_amd.MODE_FP32)
    b = song_amd.Batch(ix, [song_amd.tier(pq, queries_per_wave=qpw)], nq, mode=song_amd.MODE_FP32, storage=storage)
    ids = torch.full((nq, k), -7, dtype=torch.int32, device=q.device)
    ds = torch.zeros((nq, k), dtype=torch.float32, device=q.device)
    fl = torch.empty(nq, dtype=torch.int32, device=q.device)
    b.run(q, ids, fl, pq=pq, k=k, dists=ds)
    torch.cuda.synchronize()
    b.close()
    assert torch.equal(ids, want_i) and torch.equal(ds, want_d)
