"""Pins the CPU oracle before it is trusted as the checker.

Checks: golden graph hashes and id lists written by the reference binary
(song.cpu, reference main.cc), closed-form distances, brute force on a
complete graph, and a hand-worked trace of the stop rule with ties.
"""
import numpy as np
import pytest

import golden_cases
import pyoracle

BUILD_CASES = ["tiny_cos_500", "tiny_l2_500", "tie_l2_3k", "tie_ip_2k", "ip_10k", "c1_cos_15k"]


@pytest.mark.parametrize("name", BUILD_CASES)
def test_oracle_build_matches_reference_graph_bytes(name):
    meta, base, _, _, graph = golden_cases.load(name)
    g = pyoracle.build(base, meta["metric"])
    assert golden_cases.sha256(g) == meta["graph_sha256"]
    assert golden_cases.sha256(base) == meta["data_sha256"]
    if graph is not None:
        assert np.array_equal(g, graph)


@pytest.mark.parametrize("name", golden_cases.names())
def test_oracle_search_matches_reference_ids(name):
    meta, base, query, ids, graph = golden_cases.load(name)
    if graph is None:
        graph = pyoracle.build(base, meta["metric"])
        assert golden_cases.sha256(graph) == meta["graph_sha256"]
    pq = meta["pq"]
    got, _ = pyoracle.search_batch(base, graph, meta["metric"], query, pq, threads=8)
    got = got.astype(np.int64)
    got[got == np.int64(-1)] = -1  # UINT64_MAX padding
    got[got > np.iinfo(np.int64).max // 2] = -1
    assert np.array_equal(got, ids)


def test_distance_closed_forms():
    rng = np.random.default_rng(0)
    x = rng.integers(0, 256, 128).astype(np.float32)
    q = rng.integers(0, 256, 128).astype(np.float32)
    assert pyoracle.distance("l2", x, q) == float(((x.astype(np.int64) - q.astype(np.int64)) ** 2).sum())
    assert pyoracle.distance("ip", x, q) == -float((x.astype(np.int64) * q.astype(np.int64)).sum())
    a = np.array([1.0, 0.0], np.float32)
    b = np.array([-2.0, 0.0], np.float32)
    # negative cosine: sign * cos^2 with sign = +1 when the dot product is negative
    assert pyoracle.distance("cos", a, b) == 1.0
    assert pyoracle.distance("cos", a, a) == -1.0
    c = np.array([3.0, 4.0], np.float32)
    assert abs(pyoracle.distance("cos", c, a) - (-(9.0 / 25.0))) < 1e-15


def test_distance_float_products_double_accumulation():
    # 1 + 2^-30 is not representable in float, so a float accumulator would
    # lose the second term; the reference adds float products into a double.
    x = np.array([1.0, 2.0 ** -15], np.float32)
    q = np.array([0.0, 0.0], np.float32)
    assert pyoracle.distance("l2", x, q) == 1.0 + 2.0 ** -30
    # the product itself is rounded to float before widening
    y = np.array([1.0 + 2.0 ** -23], np.float32)
    z = np.array([0.0], np.float32)
    assert pyoracle.distance("l2", y, z) == float(np.float32(y[0]) * np.float32(y[0]))


def test_complete_graph_equals_bruteforce():
    rng = np.random.default_rng(3)
    n, dim = 32, 5
    base = rng.integers(0, 3, size=(n, dim)).astype(np.float32)  # many ties
    graph = np.zeros((n, 32), np.uint64)
    for v in range(n):
        nb = [u for u in range(n) if u != v][:31]
        graph[v, 0] = len(nb)
        graph[v, 1:1 + len(nb)] = nb
    for metric in ("l2", "ip"):
        for qi in range(8):
            q = rng.integers(0, 3, size=dim).astype(np.float32)
            ids, ds, _ = pyoracle.search(base, graph, metric, q, n)
            d = np.array([pyoracle.distance(metric, base[i], q) for i in range(n)])
            order = sorted(range(n), key=lambda i: (d[i], i))
            assert ids.tolist() == order


def test_stop_rule_trace_with_ties():
    # Hand-worked 6-vertex path 0-1-2-3-4-5 in 1-d L2; query at 0.
    # dists: v0=0, v1=1, v2=1, v3=4, v4=9, v5=16  (v1 at +1, v2 at -1)
    base = np.array([[0.0], [1.0], [-1.0], [2.0], [3.0], [4.0]], np.float32)
    graph = np.zeros((6, 32), np.uint64)
    adj = {0: [1], 1: [2], 2: [3], 3: [4], 4: [5], 5: []}
    for v, nb in adj.items():
        graph[v, 0] = len(nb)
        graph[v, 1:1 + len(nb)] = nb
    q = np.array([0.0], np.float32)
    # k=2: pop v0 -> topk{0}; push v1. pop v1 -> topk{0,1}; push v2 (d=1).
    # now=v2 d=1: topk.max.d = 1, not < 1 -> expand v2 (tie), evicted at once; push v3.
    # now=v3 d=4 > 1 -> stop.  Expansions = 3, visited = {0,1,2,3}.
    ids, ds, (exp, vis) = pyoracle.search(base, graph, "l2", q, 2)
    assert ids.tolist() == [0, 1]
    assert (exp, vis) == (3, 4)
    # k=3: topk fills with v0,v1,v2; v3 d=4 > 1 -> stop before expanding v3.
    ids, ds, (exp, vis) = pyoracle.search(base, graph, "l2", q, 3)
    assert ids.tolist() == [0, 1, 2]
    assert (exp, vis) == (3, 4)


def test_parse_libsvm_lines():
    assert pyoracle.parse_libsvm_line("0 1:0.5 3:-2\n") == [(0, 0.5), (2, -2.0)]
    # %lf then cast to float: 0.1 becomes the float nearest 0.1
    (i, v), = pyoracle.parse_libsvm_line("7 2:0.1")
    assert i == 1 and v == float(np.float32(0.1))
    # a trailing space adds an odd token and is ignored
    assert pyoracle.parse_libsvm_line("1 4:3 \n") == [(3, 3.0)]
    assert pyoracle.parse_libsvm_line("1\n") == []
    assert pyoracle.parse_libsvm_line("0 1:2", one_based=False) == [(1, 2.0)]


def test_bruteforce_knn_matches_numpy():
    rng = np.random.default_rng(5)
    base = rng.standard_normal((300, 7)).astype(np.float32)
    qs = rng.standard_normal((5, 7)).astype(np.float32)
    ids, ds = pyoracle.bruteforce_knn(base, "l2", qs, 10, threads=2)
    d64 = ((base.astype(np.float64)[None] - qs.astype(np.float64)[:, None]) ** 2).sum(-1)
    ref = np.argsort(d64, axis=1, kind="stable")[:, :10]
    assert np.array_equal(ids.astype(np.int64), ref)
    ids, _ = pyoracle.bruteforce_knn(base, "ip", qs, 10, threads=2)
    ref = np.argsort(-(qs.astype(np.float64) @ base.astype(np.float64).T), axis=1, kind="stable")[:, :10]
    assert np.array_equal(ids.astype(np.int64), ref)
