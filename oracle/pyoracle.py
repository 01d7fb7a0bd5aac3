"""ctypes loader for the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker or the CPU baseline, never as the measured or
shipped path.  The library is built by ``make -C oracle`` (see build() in
__graft_entry__.py).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

METRICS = {"l2": 0, "ip": 1, "cos": 2}


class Stats(ctypes.Structure):
    _fields_ = [("expansions", ctypes.c_int64), ("visited", ctypes.c_int64)]


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        f, d, u64, i32 = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double), \
            ctypes.POINTER(ctypes.c_uint64), ctypes.c_int
        L.oracle_distance.restype = ctypes.c_double
        L.oracle_distance.argtypes = [i32, f, f, i32]
        L.oracle_pair_distance.restype = ctypes.c_double
        L.oracle_pair_distance.argtypes = [i32, f, f, i32]
        L.oracle_search.restype = i32
        L.oracle_search.argtypes = [f, ctypes.c_uint64, i32, u64, i32, i32, f, i32, u64, d,
                                    ctypes.POINTER(Stats)]
        L.oracle_search_batch.restype = i32
        L.oracle_search_batch.argtypes = [f, ctypes.c_uint64, i32, u64, i32, i32, f, ctypes.c_uint64,
                                          i32, u64, d, ctypes.POINTER(Stats), i32]
        L.oracle_build.restype = i32
        L.oracle_build.argtypes = [f, ctypes.c_uint64, i32, i32, u64]
        L.oracle_parse_libsvm_line.restype = i32
        L.oracle_parse_libsvm_line.argtypes = [ctypes.c_char_p, i32, ctypes.POINTER(ctypes.c_int), f, i32]
        L.oracle_bruteforce_knn.restype = i32
        L.oracle_bruteforce_knn.argtypes = [f, ctypes.c_uint64, i32, i32, f, ctypes.c_uint64, i32, u64,
                                            d, i32]
        _LIB = L
    return _LIB


def _metric(m):
    return METRICS[m] if isinstance(m, str) else int(m)


def distance(metric, x, q):
    x = np.ascontiguousarray(x, dtype=np.float32)
    q = np.ascontiguousarray(q, dtype=np.float32)
    return lib().oracle_distance(_metric(metric), _p(x, ctypes.c_float), _p(q, ctypes.c_float), x.shape[0])


def search(data, graph, metric, query, k, shift=5, n_visible=None):
    """One query; returns (ids uint64[m], dists float64[m], (expansions, visited))."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    graph = np.ascontiguousarray(graph, dtype=np.uint64)
    query = np.ascontiguousarray(query, dtype=np.float32)
    n = data.shape[0] if n_visible is None else n_visible
    ids = np.empty(k, dtype=np.uint64)
    ds = np.empty(k, dtype=np.float64)
    st = Stats()
    m = lib().oracle_search(_p(data, ctypes.c_float), n, data.shape[1], _p(graph, ctypes.c_uint64), shift,
                            _metric(metric), _p(query, ctypes.c_float), k, _p(ids, ctypes.c_uint64),
                            _p(ds, ctypes.c_double), ctypes.byref(st))
    return ids[:m], ds[:m], (st.expansions, st.visited)


def search_batch(data, graph, metric, queries, k, threads=1, shift=5, with_stats=False):
    """nq queries; ids uint64[nq,k] (UINT64_MAX padding), dists float64[nq,k]."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    graph = np.ascontiguousarray(graph, dtype=np.uint64)
    queries = np.ascontiguousarray(queries, dtype=np.float32)
    nq = queries.shape[0]
    ids = np.empty((nq, k), dtype=np.uint64)
    ds = np.empty((nq, k), dtype=np.float64)
    st = (Stats * max(nq, 1))()
    lib().oracle_search_batch(_p(data, ctypes.c_float), data.shape[0], data.shape[1],
                              _p(graph, ctypes.c_uint64), shift, _metric(metric),
                              _p(queries, ctypes.c_float), nq, k, _p(ids, ctypes.c_uint64),
                              _p(ds, ctypes.c_double), st, threads)
    if with_stats:
        exp = np.array([s.expansions for s in st[:nq]], dtype=np.int64)
        vis = np.array([s.visited for s in st[:nq]], dtype=np.int64)
        return ids, ds, exp, vis
    return ids, ds


def build(data, metric):
    """Reference insertion build; returns the uint64[n, 32] bfsg.graph image."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    n = data.shape[0]
    g = np.zeros((n, 32), dtype=np.uint64)
    lib().oracle_build(_p(data, ctypes.c_float), n, data.shape[1], _metric(metric), _p(g, ctypes.c_uint64))
    return g


def parse_libsvm_line(line: str, one_based=True, cap=4096):
    idx = (ctypes.c_int * cap)()
    val = np.empty(cap, dtype=np.float32)
    m = lib().oracle_parse_libsvm_line(line.encode(), 1 if one_based else 0, idx, _p(val, ctypes.c_float), cap)
    m = min(m, cap)
    return [(idx[i], float(val[i])) for i in range(m)]


def bruteforce_knn(data, metric, queries, k, threads=8):
    data = np.ascontiguousarray(data, dtype=np.float32)
    queries = np.ascontiguousarray(queries, dtype=np.float32)
    nq = queries.shape[0]
    ids = np.empty((nq, k), dtype=np.uint64)
    ds = np.empty((nq, k), dtype=np.float64)
    lib().oracle_bruteforce_knn(_p(data, ctypes.c_float), data.shape[0], data.shape[1], _metric(metric),
                                _p(queries, ctypes.c_float), nq, k, _p(ids, ctypes.c_uint64),
                                _p(ds, ctypes.c_double), threads)
    return ids, ds
