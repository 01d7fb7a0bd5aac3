"""Wide rows on the GPU: 960-d (the precision workload's width) and 1,100-d rows.

960 float32 values are 240 16-byte units per row, more than the 64-lane shapes cover in
one pass of their smaller instantiations; 1,100 values (275 units) take the chunked
64-lane shape.  Every lanes-per-row shape (forced through a batch tier) must give the
exact search's ids, the exact mode must equal the oracle (ids bit for bit, distances
within 1e-5), FP32 accumulation must be exact on integer rows, and the BF16 / FP8 rows
with the exact FP32 re-rank must return FP32 distances with a bounded recall delta
against the float64 ground truth.
"""
import numpy as np
import pytest

import lowprec
import pyoracle
import song_amd
from song_amd import datasets, groundtruth, tier

pytestmark = pytest.mark.gpu

DIST_RTOL = 1e-5


def _int_rows(n, nq, dim, seed):
    """Integer rows in [0, 64]: every partial L2 sum stays below 2^24 at these widths, so
    FP32 accumulation is exact."""
    rng = np.random.default_rng(seed)
    centres = rng.integers(0, 65, size=(40, dim))
    lab = rng.integers(0, 40, size=n + nq)
    x = np.clip(centres[lab] + rng.integers(-6, 7, size=(n + nq, dim)), 0, 64).astype(np.float32)
    return x[:n], x[n:].copy()


@pytest.fixture(scope="module", params=[960, 1100])
def wide(request):
    dim = request.param
    print(f"[wide] dim {dim}", flush=True)
    base, query = datasets.factor_l2(n_base=4000, n_query=120, dim=dim, seed=3)
    graph = song_amd.build_graph(base, "l2", threads=8)
    ix = song_amd.Index(base, graph, "l2")
    ix.add_storage(song_amd.STORE_BF16)
    ix.add_storage(song_amd.STORE_FP8)
    return dim, base, query, graph, ix


def test_exact_mode_matches_oracle(wide):
    dim, base, query, graph, ix = wide
    pq = 48
    o_ids, o_ds = pyoracle.search_batch(base, graph, "l2", query, pq, threads=8)
    ids, ds, flags = ix.search(query, pq, mode=song_amd.MODE_EXACT)
    assert np.array_equal(ids, o_ids.astype(np.int64))
    ok = ids >= 0
    err = np.abs(ds[ok] - o_ds[ok]) / np.maximum(np.abs(o_ds[ok]), 1e-30)
    assert err.max() <= DIST_RTOL


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4])
def test_every_row_shape_gives_the_exact_ids(wide, shape):
    """Shapes 0-2 read a 960-d row over 4-30 passes of 4 lanes; shape 3 in one or two passes of
    16 lanes; shape 4 in one chunked pass of 64 lanes."""
    import torch

    dim, base, query, graph, ix = wide
    pq, k = 48, 20
    want, wd, _ = ix.search(query, pq, k, mode=song_amd.MODE_EXACT)
    q = torch.from_numpy(query).cuda()
    nq = q.shape[0]
    b = song_amd.Batch(ix, [tier(pq, shape=shape)], nq, mode=song_amd.MODE_EXACT)
    ids = torch.full((nq, k), -7, dtype=torch.int32, device=q.device)
    ds = torch.zeros((nq, k), dtype=torch.float32, device=q.device)
    fl = torch.empty(nq, dtype=torch.int32, device=q.device)
    b.run(q, ids, fl, pq=pq, k=k, dists=ds)
    torch.cuda.synchronize()
    b.close()
    assert np.array_equal(ids.cpu().numpy().astype(np.int64), want)
    assert np.array_equal(ds.cpu().numpy(), wd)


def test_fp32_accumulation_is_exact_on_integer_rows(wide):
    dim = wide[0]
    base, query = _int_rows(3000, 80, dim, seed=dim)
    graph = song_amd.build_graph(base, "l2", threads=8)
    ix = song_amd.Index(base, graph, "l2")
    assert ix.integer_valued
    exact, _, _ = ix.search(query, 40, mode=song_amd.MODE_EXACT)
    f32, _, _ = ix.search(query, 40, mode=song_amd.MODE_FP32)
    assert np.array_equal(exact, f32)
    o_ids, _ = pyoracle.search_batch(base, graph, "l2", query, 40, threads=8)
    assert np.array_equal(exact, o_ids.astype(np.int64))


@pytest.mark.parametrize("storage,name", [(song_amd.STORE_BF16, "bf16"), (song_amd.STORE_FP8, "fp8")])
def test_low_precision_rows_rerank_and_recall(wide, storage, name):
    dim, base, query, graph, ix = wide
    codes, scales = ix.get_storage(storage)
    rc, rs, _ = lowprec.quantize(base[:500], name)
    got = codes[:500].view(np.uint16)[:, :dim] if name == "bf16" else codes[:500, :dim]
    assert np.array_equal(got, rc) and np.array_equal(scales[:500], rs)
    pq, k = 64, 10
    ids, ds, _ = ix.search(query, pq, k, storage=storage)
    assert (ids >= 0).all()
    exact = np.array([[pyoracle.distance("l2", base[j], query[i]) for j in row] for i, row in enumerate(ids)])
    assert np.allclose(ds, exact, rtol=DIST_RTOL, atol=0)
    assert (np.diff(ds, axis=1) >= 0).all()
    _, gt_d = groundtruth.knn_fp64(base, query, k, "l2")
    plain, _, _ = ix.search(query, pq, k, mode=song_amd.MODE_FP32)
    r_plain = groundtruth.recall_at_k(plain, gt_d, k, base, query)
    r = groundtruth.recall_at_k(ids, gt_d, k, base, query)
    print(f"[wide] dim {dim} {name}+rerank recall@{k} {r:.4f} vs fp32 {r_plain:.4f}", flush=True)
    assert r >= r_plain - 0.05
