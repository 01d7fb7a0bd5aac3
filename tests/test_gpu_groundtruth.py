"""The float64 ground truth on the GPU (song_amd.groundtruth.knn_fp64), which decides every
recall figure, against the oracle's brute force (float64 products and sums on the CPU,
(distance, id) order), and the reference-shaped one-shot seam.

On the 1M x 128 integer-valued workload every float64 distance is exact in both, so ids
and distances must be identical; on float data (cosine, inner product) the two differ only
by float64 rounding, so distances agree to 1e-12 relative and ids agree wherever the
distances are not tied within that tolerance.
"""
import ctypes

import numpy as np
import pytest

import golden_cases
import pyoracle
import song_amd
from song_amd import datasets, groundtruth

pytestmark = pytest.mark.gpu


def test_knn_fp64_equals_oracle_on_the_1m_workload():
    base, query = datasets.mixture_l2()
    rng = np.random.default_rng(0)
    sel = np.sort(rng.choice(query.shape[0], 200, replace=False))
    q = query[sel]
    g_ids, g_ds = groundtruth.knn_fp64(base, q, 100, "l2")
    o_ids, o_ds = pyoracle.bruteforce_knn(base, "l2", q, 100, threads=16)
    assert np.array_equal(g_ds, o_ds)
    assert np.array_equal(g_ids, o_ids.astype(np.int64))


@pytest.mark.parametrize("name", ["c1_cos_15k", "ip_10k", "tie_l2_3k"])
def test_knn_fp64_equals_oracle_on_float_and_tied_data(name):
    meta, base, query, _, _ = golden_cases.load(name)
    metric = meta["metric"]
    k = 50
    g_ids, g_ds = groundtruth.knn_fp64(base, query, k, metric)
    o_ids, o_ds = pyoracle.bruteforce_knn(base, metric, query, k, threads=8)
    o_ids = o_ids.astype(np.int64)
    assert np.allclose(g_ds, o_ds, rtol=1e-12, atol=1e-12)
    # where ids differ, the two lists hold the same distances there (a tie within rounding)
    diff = g_ids != o_ids
    for i, j in zip(*np.nonzero(diff)):
        dg = pyoracle.distance(metric, base[g_ids[i, j]], query[i])
        do = pyoracle.distance(metric, base[o_ids[i, j]], query[i])
        assert abs(dg - do) <= 1e-6 * max(1.0, abs(do)), (name, i, j)
    print(f"[gt] {name}: {int(diff.sum())} of {diff.size} positions swapped within a tie", flush=True)
    # the tie-aware recall of the oracle's own lists against the GPU ground truth is 1
    assert groundtruth.recall_at_k(o_ids, g_ds, 10, base, query, metric) == 1.0


@pytest.mark.parametrize("name", ["tiny_l2_500", "ip_10k", "c1_cos_15k"])
def test_reference_seam_returns_reference_ids(name):
    """song_astar_multi_start_search_batch with the reference seam's argument list
    (warp_astar_accelerator.h:347): host queries, data and bfsg.graph rows in, pq ids per
    query out (UINT64_MAX padding), uploaded and freed inside the call."""
    meta, base, query, ref_ids, graph = golden_cases.load(name)
    if graph is None:
        graph = song_amd.build_graph(base, meta["metric"], threads=8)
    assert golden_cases.sha256(graph) == meta["graph_sha256"]
    pq = meta["pq"]
    n, dim = base.shape
    q = np.ascontiguousarray(query, dtype=np.float32)
    d = np.ascontiguousarray(base, dtype=np.float32)
    g = np.ascontiguousarray(graph, dtype=np.uint64)
    out = np.zeros((q.shape[0], pq), dtype=np.uint64)
    rc = song_amd.lib().song_astar_multi_start_search_batch(
        ctypes.c_void_p(q.ctypes.data), q.shape[0], pq, ctypes.c_void_p(out.ctypes.data),
        ctypes.c_void_p(d.ctypes.data), ctypes.c_void_p(g.ctypes.data), 5, n, dim,
        song_amd.metric_code(meta["metric"]))
    assert rc == 0, song_amd.lib().song_last_error()
    got = out.astype(np.int64)
    got[out == np.uint64(0xFFFFFFFFFFFFFFFF)] = -1
    assert np.array_equal(got, ref_ids)
