"""song_amd: MI355X (gfx950) batched graph search behind a C ABI.

The search runs only in the HIP library ``song_amd/lib/libsong_hip.so``
(declared in include/song_amd.h); this module is a thin ctypes binding over
that ABI.  Importing it does not need a GPU; calling a search without the
library or without a GPU raises ``SongError``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .datasets import METRIC_COS, METRIC_IP, METRIC_L2, METRIC_NAMES  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsong_hip.so")

MODE_AUTO, MODE_EXACT, MODE_FP32 = 0, 1, 2
STORE_F32, STORE_BF16, STORE_FP8 = 0, 1, 2
FLAG_VISITED_OVF, FLAG_QUEUE_OVF, FLAG_STEP_CAP = 1, 2, 4

_lib = None


class SongError(RuntimeError):
    pass


class QueryParams(ctypes.Structure):
    _fields_ = [("pq", ctypes.c_void_p), ("k", ctypes.c_void_p), ("entry", ctypes.c_void_p), ("n_entry", ctypes.c_int),
                ("steps", ctypes.c_void_p), ("subset", ctypes.c_void_p), ("n_subset", ctypes.c_uint64)]


class SearchOpts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("slack", ctypes.c_int), ("hash_log2", ctypes.c_int),
                ("max_steps", ctypes.c_int), ("retry", ctypes.c_int), ("storage", ctypes.c_int),
                ("rerank", ctypes.c_int), ("visited_bits", ctypes.c_int), ("bloom_k", ctypes.c_int),
                ("best_in_global", ctypes.c_int)]


class Tier(ctypes.Structure):
    _fields_ = [("pq_max", ctypes.c_int), ("hash_log2", ctypes.c_int), ("slack", ctypes.c_int), ("shape", ctypes.c_int),
                ("waves_per_group", ctypes.c_int), ("groups_per_cu", ctypes.c_int), ("visited_bits", ctypes.c_int),
                ("storage", ctypes.c_int), ("bloom_k", ctypes.c_int), ("queries_per_wave", ctypes.c_int),
                ("multiprobe", ctypes.c_int), ("bag_in_lds", ctypes.c_int), ("waves_per_query", ctypes.c_int),
                ("best_in_global", ctypes.c_int)]


MAX_TIERS = 8


class BatchConfig(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("storage", ctypes.c_int), ("rerank", ctypes.c_int),
                ("max_steps", ctypes.c_int), ("spill_log2", ctypes.c_int), ("n_tiers", ctypes.c_int),
                ("tier", Tier * MAX_TIERS), ("retry", Tier), ("parallel_tiers", ctypes.c_int),
                ("priority_first", ctypes.c_int)]


def lib():
    """Load the HIP library (raises SongError if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SongError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        # One HIP runtime per process: torch ships its own libamdhip64 (SONAME
        # libamdhip64.so.7).  Loading torch first lets this library's NEEDED entry bind
        # to that copy instead of starting a second runtime from /opt/rocm.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        L.song_last_error.restype = ctypes.c_char_p
        L.song_search_opts_default.argtypes = [ctypes.POINTER(SearchOpts)]
        L.song_index_create.argtypes = [vp, u64, i32, vp, i32, i32, ctypes.POINTER(vp)]
        L.song_index_destroy.argtypes = [vp]
        L.song_index_integer_valued.argtypes = [vp]
        L.song_search.argtypes = [vp, vp, u64, i32, i32, vp, vp, vp, ctypes.POINTER(SearchOpts)]
        L.song_search_device.argtypes = [vp, vp, u64, i32, i32, vp, vp, vp, vp, i32, vp, vp,
                                         ctypes.POINTER(SearchOpts), vp]
        L.song_astar_multi_start_search_batch.argtypes = [vp, u64, i32, vp, vp, vp, i32, u64, i32, i32]
        L.song_search_device_ex.argtypes = [vp, vp, u64, i32, i32, ctypes.POINTER(QueryParams), vp, vp, i32, vp,
                                            vp, ctypes.POINTER(SearchOpts), vp]
        L.song_index_add_storage.argtypes = [vp, i32]
        L.song_index_get_storage.argtypes = [vp, i32, vp, vp, ctypes.POINTER(ctypes.c_uint64)]
        L.song_batch_config_default.argtypes = [ctypes.POINTER(BatchConfig), i32]
        L.song_resolve_mode.argtypes = [vp, vp, u64, ctypes.POINTER(i32), vp]
        L.song_batch_create.argtypes = [vp, ctypes.POINTER(BatchConfig), u64, ctypes.POINTER(vp)]
        L.song_batch_destroy.argtypes = [vp]
        run_args = [vp, vp, u64, i32, i32, ctypes.POINTER(QueryParams), vp, vp, i32, vp, vp, vp]
        L.song_batch_run.argtypes = run_args
        L.song_batch_run_graph.argtypes = run_args
        L.song_batch_counts.argtypes = [vp, vp, vp]
        L.song_batch_grid.argtypes = [vp, i32, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.song_bloom_size.argtypes = [i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.song_search_device_noalloc.argtypes = [vp, vp, u64, i32, i32, ctypes.POINTER(QueryParams), vp, vp, i32,
                                                 vp, vp, ctypes.POINTER(SearchOpts), vp]
        _lib = L
    return _lib


EXPORTED_SYMBOLS = (
    "song_search_opts_default", "song_index_create", "song_index_destroy", "song_index_integer_valued",
    "song_search", "song_search_device", "song_astar_multi_start_search_batch", "song_last_error",
    "song_index_add_storage", "song_index_get_storage", "song_search_device_ex",
    "song_batch_config_default", "song_resolve_mode", "song_batch_create", "song_batch_destroy", "song_batch_run",
    "song_batch_run_graph", "song_batch_counts", "song_batch_grid", "song_search_device_noalloc",
    "song_bloom_size",
)
NOALLOC_SCRATCH_BYTES = 64 << 20
FLAG_BAD_PARAMS = 8


def _check(rc):
    if rc != 0:
        raise SongError(f"song error {rc}: {lib().song_last_error().decode(errors='replace')}")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _check_tensor(t, name, dtypes, rows, cols=None, device=None, ndim=None):
    """Argument check before a raw pointer goes to the library: dtype, contiguity, device and
    enough rows (and columns) for the kernel's writes and reads."""
    import torch

    if t is None:
        return
    if not isinstance(t, torch.Tensor):
        raise SongError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.dtype not in dtypes:
        raise SongError(f"{name}: dtype {t.dtype} not in {[str(d) for d in dtypes]}")
    if not t.is_cuda or (device is not None and t.device != device):
        raise SongError(f"{name}: must be on {device or 'a cuda device'}, is on {t.device}")
    if not t.is_contiguous():
        raise SongError(f"{name}: must be contiguous")
    if ndim is not None and t.dim() not in ndim:
        raise SongError(f"{name}: {t.dim()} dimensions, expected {ndim}")
    if t.shape[0] < rows:
        raise SongError(f"{name}: {t.shape[0]} rows, need at least {rows}")
    if cols is not None and (t.dim() < 2 or t.shape[1] < cols):
        raise SongError(f"{name}: shape {tuple(t.shape)}, need at least {cols} columns")


def _int32s():
    import torch

    return (torch.int32, torch.uint32)


def _f32():
    import torch

    return (torch.float32,)


def metric_code(metric):
    return METRIC_NAMES[metric] if isinstance(metric, str) else int(metric)


def make_opts(mode=MODE_AUTO, slack=128, hash_log2=0, max_steps=0, retry=1, storage=STORE_F32, rerank=1,
              visited_bits=0, bloom_k=0, best_in_global=-1):
    o = SearchOpts()
    lib().song_search_opts_default(ctypes.byref(o))
    o.mode, o.slack, o.hash_log2, o.max_steps, o.retry = mode, slack, hash_log2, max_steps, retry
    o.storage, o.rerank, o.visited_bits, o.bloom_k = storage, rerank, visited_bits, bloom_k
    o.best_in_global = best_in_global
    return o


def bloom_size(pq):
    """(hash_log2, bloom_k) of the reference's auto_tune_bloom sizing for budget pq: 2^hash_log2
    32-bit words, bloom_k bits per id."""
    h, k = ctypes.c_int(), ctypes.c_int()
    _check(lib().song_bloom_size(int(pq), ctypes.byref(h), ctypes.byref(k)))
    return h.value, k.value


class Index:
    """An uploaded (data, graph) pair.  data: float32 [n, dim]; graph: uint64 [n, 2**shift]."""

    def __init__(self, data, graph, metric, shift=5):
        self.data = np.ascontiguousarray(data, dtype=np.float32)
        g = np.ascontiguousarray(graph, dtype=np.uint64)
        self.n, self.dim = self.data.shape
        self.metric = metric_code(metric)
        h = ctypes.c_void_p()
        _check(lib().song_index_create(_ptr(self.data), self.n, self.dim, _ptr(g), shift, self.metric,
                                       ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().song_index_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def integer_valued(self):
        return bool(lib().song_index_integer_valued(self._h))

    def add_storage(self, storage):
        """Build the BF16 / FP8 copy of the rows on the GPU."""
        _check(lib().song_index_add_storage(self._h, storage))

    def get_storage(self, storage):
        """(codes uint8 [n, row_bytes], scales float32 [n]) of a built low-precision copy."""
        rb = ctypes.c_uint64()
        _check(lib().song_index_get_storage(self._h, storage, None, None, ctypes.byref(rb)))
        codes = np.empty((self.n, rb.value), dtype=np.uint8)
        scales = np.empty(self.n, dtype=np.float32)
        _check(lib().song_index_get_storage(self._h, storage, _ptr(codes), _ptr(scales), ctypes.byref(rb)))
        return codes, scales

    def search(self, queries, pq, k=None, mode=MODE_AUTO, slack=128, hash_log2=0, retry=1, storage=STORE_F32,
               rerank=1, visited_bits=0, bloom_k=0, best_in_global=-1):
        """Host buffers in and out: ids int64 [nq, k] (-1 padding), dists float32, flags uint32."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        k = pq if k is None else min(k, pq)
        ids = np.empty((nq, k), dtype=np.uint64)
        ds = np.empty((nq, k), dtype=np.float32)
        fl = np.empty(nq, dtype=np.uint32)
        o = make_opts(mode, slack, hash_log2, 0, retry, storage, rerank, visited_bits, bloom_k, best_in_global)
        _check(lib().song_search(self._h, _ptr(q), nq, pq, k, _ptr(ids), _ptr(ds), _ptr(fl), ctypes.byref(o)))
        out = ids.astype(np.int64)
        out[ids == np.uint64(0xFFFFFFFFFFFFFFFF)] = -1
        return out, ds, fl

    def search_device(self, queries, pq, k, ids, dists=None, flags=None, stats=None, pq_arr=None, k_arr=None,
                      mode=MODE_AUTO, slack=128, hash_log2=0, retry=1, stream=None, storage=STORE_F32, rerank=1,
                      max_steps=0, entry=None, steps=None, subset=None, visited_bits=0, bloom_k=0, best_in_global=-1):
        """Device buffers (torch tensors on cuda): queries float32 [nq, dim], ids int32/uint32 [nq, >=k].

        entry: int32 [nq] or [nq, E] start vertices (learned form; default vertex 0);
        steps: int32 [nq] per-query expansion caps; subset: int32 query indices to run;
        stats: int32 [nq, 8]."""
        import torch

        nq = queries.shape[0]
        dev = queries.device
        _check_tensor(queries, "queries", _f32(), nq, self.dim, ndim=(2,))
        if queries.shape[1] != self.dim:
            raise SongError(f"queries: {queries.shape[1]} columns, the index has dim {self.dim}")
        _check_tensor(ids, "ids", _int32s(), nq, k, dev, ndim=(2,))
        _check_tensor(dists, "dists", _f32(), nq, ids.shape[1], dev, ndim=(2,))
        _check_tensor(flags, "flags", _int32s(), nq, None, dev, ndim=(1,))
        _check_tensor(stats, "stats", _int32s(), nq, 8, dev, ndim=(2,))
        for name, t in (("pq_arr", pq_arr), ("k_arr", k_arr), ("steps", steps)):
            _check_tensor(t, name, _int32s(), nq, None, dev, ndim=(1,))
        _check_tensor(entry, "entry", _int32s(), nq, None, dev, ndim=(1, 2))
        if subset is not None:
            _check_tensor(subset, "subset", _int32s(), 0, None, dev, ndim=(1,))
            if subset.numel() and (int(subset.min().item()) < 0 or int(subset.max().item()) >= nq):
                raise SongError(f"subset: query indices outside [0, {nq})")
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        dp = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        o = make_opts(mode, slack, hash_log2, max_steps, retry, storage, rerank, visited_bits, bloom_k, best_in_global)
        qp = QueryParams()
        qp.pq, qp.k = (pq_arr.data_ptr() if pq_arr is not None else None), (k_arr.data_ptr() if k_arr is not None else None)
        qp.entry = entry.data_ptr() if entry is not None else None
        qp.n_entry = 1 if entry is None else (entry.shape[1] if entry.dim() == 2 else 1)
        qp.steps = steps.data_ptr() if steps is not None else None
        qp.subset = subset.data_ptr() if subset is not None else None
        qp.n_subset = subset.numel() if subset is not None else 0
        _check(lib().song_search_device_ex(self._h, dp(queries), nq, pq, k, ctypes.byref(qp), dp(ids), dp(dists),
                                           ids.shape[1], dp(flags), dp(stats), ctypes.byref(o),
                                           ctypes.c_void_p(stream)))


def search_device_noalloc(index, queries, pq, k, ids, dists=None, flags=None, stats=None, pq_arr=None, k_arr=None,
                          mode=MODE_EXACT, slack=128, hash_log2=0, storage=STORE_F32, rerank=1, max_steps=0,
                          entry=None, steps=None, subset=None, visited_bits=0, stream=None, bloom_k=0,
                          best_in_global=-1):
    """song_search_device_noalloc: Index.search_device without device allocation (scratch in the
    library's static per-GPU array); mode must be MODE_EXACT or MODE_FP32, overflowed queries keep
    their flags (no re-run).  Argument checks as search_device; no host synchronization."""
    import torch

    nq = queries.shape[0]
    dev = queries.device
    _check_tensor(queries, "queries", _f32(), nq, index.dim, ndim=(2,))
    _check_tensor(ids, "ids", _int32s(), nq, k, dev, ndim=(2,))
    _check_tensor(dists, "dists", _f32(), nq, ids.shape[1], dev, ndim=(2,))
    _check_tensor(flags, "flags", _int32s(), nq, None, dev, ndim=(1,))
    _check_tensor(stats, "stats", _int32s(), nq, 8, dev, ndim=(2,))
    for name, t in (("pq_arr", pq_arr), ("k_arr", k_arr), ("steps", steps)):
        _check_tensor(t, name, _int32s(), nq, None, dev, ndim=(1,))
    _check_tensor(entry, "entry", _int32s(), nq, None, dev, ndim=(1, 2))
    _check_tensor(subset, "subset", _int32s(), 0, None, dev, ndim=(1,))
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    dp = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    o = make_opts(mode, slack, hash_log2, max_steps, 0, storage, rerank, visited_bits, bloom_k, best_in_global)
    qp = _query_params(pq_arr, k_arr, entry, steps, subset)
    _check(lib().song_search_device_noalloc(index._h, dp(queries), nq, pq, k, ctypes.byref(qp), dp(ids), dp(dists),
                                            ids.shape[1], dp(flags), dp(stats), ctypes.byref(o),
                                            ctypes.c_void_p(stream)))


def _query_params(pq_arr=None, k_arr=None, entry=None, steps=None, subset=None):
    qp = QueryParams()
    qp.pq = pq_arr.data_ptr() if pq_arr is not None else None
    qp.k = k_arr.data_ptr() if k_arr is not None else None
    qp.entry = entry.data_ptr() if entry is not None else None
    qp.n_entry = 1 if entry is None else (entry.shape[1] if entry.dim() == 2 else 1)
    qp.steps = steps.data_ptr() if steps is not None else None
    qp.subset = subset.data_ptr() if subset is not None else None
    qp.n_subset = subset.numel() if subset is not None else 0
    return qp


def resolve_mode(index, queries, stream=None):
    """MODE_EXACT or MODE_FP32 as MODE_AUTO would pick for these device queries."""
    import torch

    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    m = ctypes.c_int()
    _check(lib().song_resolve_mode(index._h, ctypes.c_void_p(queries.data_ptr()), queries.shape[0], ctypes.byref(m),
                                   ctypes.c_void_p(stream)))
    return m.value


TIER_KEYS = ("pq_max", "hash_log2", "slack", "shape", "waves_per_group", "groups_per_cu", "visited_bits", "storage",
             "bloom_k", "queries_per_wave", "multiprobe", "bag_in_lds", "waves_per_query", "best_in_global")


def tier(pq_max, hash_log2=0, slack=32, shape=-1, waves_per_group=1, groups_per_cu=0, visited_bits=0, storage=-1,
         bloom_k=0, queries_per_wave=1, multiprobe=1, bag_in_lds=0, waves_per_query=1, best_in_global=-1):
    """One tier of a batch configuration (see song_tier in include/song_amd.h); storage -1 = the
    batch's row storage; bloom_k != 0: a Bloom filter in place of the exact visited table;
    queries_per_wave 2 / 4: sub-groups of 32 / 16 lanes each searching its own query; multiprobe 2
    (with queries_per_wave > 1): two expansions per step (approximate); bag_in_lds 1 (one query
    per wave, float32 rows): the queue's overflow bag in LDS; waves_per_query 2: a helper wave per query
    computes the predicted next expansion's distances ahead (FP32 mode, float32 rows);
    best_in_global 1 / 0: the best list's front in the slot's global scratch / in LDS (one query per
    wave, float32 rows, bag in global memory); -1 = 0 for a tier."""
    return dict(pq_max=pq_max, hash_log2=hash_log2, slack=slack, shape=shape, waves_per_group=waves_per_group,
                groups_per_cu=groups_per_cu, visited_bits=visited_bits, storage=storage, bloom_k=bloom_k,
                queries_per_wave=queries_per_wave, multiprobe=multiprobe, bag_in_lds=bag_in_lds,
                waves_per_query=waves_per_query, best_in_global=best_in_global)


class Batch:
    """Batching form over an Index: persistent per-tier kernels for queries with uneven k / pq.

    tiers: list of tier(...) dicts with ascending pq_max (a query goes to the first tier
    whose pq_max >= its pq); retry: a tier(...) dict for the re-run of overflowed queries
    (None: 4x the largest tier's filter and heap depth)."""

    def __init__(self, index, tiers, nq_max, mode=MODE_EXACT, storage=STORE_F32, rerank=1, max_steps=0,
                 spill_log2=14, retry=None, parallel_tiers=True, priority_first=True):
        if not 1 <= len(tiers) <= MAX_TIERS:
            raise SongError(f"{len(tiers)} tiers (1..{MAX_TIERS})")
        cfg = BatchConfig()
        lib().song_batch_config_default(ctypes.byref(cfg), int(tiers[-1]["pq_max"]))
        cfg.mode, cfg.storage, cfg.rerank, cfg.max_steps, cfg.spill_log2 = mode, storage, rerank, max_steps, spill_log2
        cfg.n_tiers = len(tiers)
        cfg.parallel_tiers = 1 if parallel_tiers else 0
        cfg.priority_first = 1 if priority_first else 0
        default = Tier()
        ctypes.pointer(default)[0] = cfg.tier[0]  # the library's default tier, before tier 0 is filled
        for i, t in enumerate(tiers):
            for key in TIER_KEYS:
                setattr(cfg.tier[i], key, int(t.get(key, getattr(default, key) if key != "pq_max" else 0)))
        if retry is not None:
            for key in TIER_KEYS:
                if key in retry:
                    setattr(cfg.retry, key, int(retry[key]))
        self.index, self.tiers, self.nq_max, self.cfg = index, [dict(t) for t in tiers], nq_max, cfg
        h = ctypes.c_void_p()
        _check(lib().song_batch_create(index._h, ctypes.byref(cfg), nq_max, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().song_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, queries, ids, flags, pq=None, k=10, pq_arr=None, k_arr=None, dists=None, stats=None, entry=None,
            steps=None, graph=False, stream=None):
        """Enqueue one run (torch cuda tensors): queries float32 [nq, dim]; ids int32 [nq, k_ld];
        flags int32 [nq]; pq_arr / k_arr / steps int32 [nq]; entry int32 [nq, E] (-1 pads);
        graph=True captures the run once per argument set and replays it."""
        import torch

        nq = queries.shape[0]
        dev = queries.device
        _check_tensor(queries, "queries", _f32(), nq, self.index.dim, ndim=(2,))
        if queries.shape[1] != self.index.dim:
            raise SongError(f"queries: {queries.shape[1]} columns, the index has dim {self.index.dim}")
        _check_tensor(ids, "ids", _int32s(), nq, k, dev, ndim=(2,))
        _check_tensor(flags, "flags", _int32s(), nq, None, dev, ndim=(1,))
        _check_tensor(dists, "dists", _f32(), nq, ids.shape[1], dev, ndim=(2,))
        _check_tensor(stats, "stats", _int32s(), nq, 8, dev, ndim=(2,))
        for name, t in (("pq_arr", pq_arr), ("k_arr", k_arr), ("steps", steps)):
            _check_tensor(t, name, _int32s(), nq, None, dev, ndim=(1,))
        _check_tensor(entry, "entry", _int32s(), nq, None, dev, ndim=(1, 2))
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        if pq is None:
            pq = self.tiers[-1]["pq_max"]
        dp = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        qp = _query_params(pq_arr, k_arr, entry, steps)
        fn = lib().song_batch_run_graph if graph else lib().song_batch_run
        _check(fn(self._h, dp(queries), queries.shape[0], int(pq), int(k), ctypes.byref(qp), dp(ids), dp(dists),
                  ids.shape[1], dp(flags), dp(stats), ctypes.c_void_p(stream)))

    def counts(self, stream=None):
        """Queries per tier and the re-run count of the last run (synchronizes)."""
        import torch

        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        out = np.zeros(len(self.tiers) + 1, dtype=np.uint32)
        _check(lib().song_batch_counts(self._h, _ptr(out), ctypes.c_void_p(stream)))
        return out

    def grid(self, t):
        """(workgroups, waves per group, LDS bytes per wave) of tier t (t = len(tiers): retry)."""
        g, w, l = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().song_batch_grid(self._h, t, ctypes.byref(g), ctypes.byref(w), ctypes.byref(l)))
        return g.value, w.value, l.value


# ---------------------------------------------------------------- host library (CPU side)

HOST_LIB_PATH = os.path.join(_HERE, "lib", "libsong_host.so")
_host = None


class BuildStats(ctypes.Structure):
    _fields_ = [("searches", ctypes.c_uint64), ("redone", ctypes.c_uint64), ("seconds", ctypes.c_double)]


HOST_EXPORTED_SYMBOLS = ("song_graph_build", "song_libsvm_parse", "song_libsvm_parse_mt", "song_libsvm_write",
                         "song_query_read", "song_file_write", "song_file_read", "song_host_last_error")
QUERY_LIBSVM, QUERY_F32, QUERY_FVECS, QUERY_FBIN = 0, 1, 2, 3


def host_lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise SongError(f"host library not built: {HOST_LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(HOST_LIB_PATH)
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        L.song_host_last_error.restype = ctypes.c_char_p
        L.song_graph_build.argtypes = [vp, u64, i32, i32, i32, vp, ctypes.POINTER(BuildStats)]
        L.song_libsvm_parse.argtypes = [ctypes.c_char_p, i32, u64, i32, vp, ctypes.POINTER(ctypes.c_uint64)]
        L.song_libsvm_parse_mt.argtypes = [ctypes.c_char_p, i32, u64, i32, i32, vp, ctypes.POINTER(ctypes.c_uint64)]
        L.song_libsvm_write.argtypes = [ctypes.c_char_p, vp, u64, i32, i32]
        L.song_query_read.argtypes = [ctypes.c_char_p, i32, i32, u64, vp, ctypes.POINTER(ctypes.c_uint64)]
        L.song_file_write.argtypes = [ctypes.c_char_p, vp, u64]
        L.song_file_read.argtypes = [ctypes.c_char_p, vp, u64]
        _host = L
    return _host


def _hcheck(rc):
    if rc != 0:
        raise SongError(f"song host error {rc}: {host_lib().song_host_last_error().decode(errors='replace')}")


def build_graph(data, metric, threads=0, return_stats=False):
    """Reference-identical insertion build (bfsg.graph image, uint64 [n, 32])."""
    d = np.ascontiguousarray(data, dtype=np.float32)
    n, dim = d.shape
    g = np.empty((n, 32), dtype=np.uint64)
    st = BuildStats()
    _hcheck(host_lib().song_graph_build(_ptr(d), n, dim, metric_code(metric), threads, _ptr(g), ctypes.byref(st)))
    if return_stats:
        return g, {"searches": st.searches, "redone": st.redone, "seconds": st.seconds}
    return g


def parse_libsvm(path, rows, dim, one_based=True, threads=0):
    """LIBSVM text -> (float32 [rows, dim], records read); threads <= 0: the library default."""
    out = np.empty((rows, dim), dtype=np.float32)
    got = ctypes.c_uint64()
    _hcheck(host_lib().song_libsvm_parse_mt(path.encode(), 1 if one_based else 0, rows, dim, threads, _ptr(out),
                                            ctypes.byref(got)))
    return out, int(got.value)


def write_libsvm(path, data, threads=0):
    """Rows as one-based LIBSVM text that parse_libsvm reads back exactly (native writer)."""
    d = np.ascontiguousarray(data, dtype=np.float32)
    _hcheck(host_lib().song_libsvm_write(path.encode(), _ptr(d), d.shape[0], d.shape[1], threads))


def read_queries(path, dim, fmt=None):
    """Query file -> float32 [rows, dim]; fmt QUERY_* or None = by extension (.f32, .fvecs, .fbin,
    else LIBSVM)."""
    if fmt is None:
        fmt = (QUERY_F32 if path.endswith(".f32") else QUERY_FVECS if path.endswith(".fvecs")
               else QUERY_FBIN if path.endswith(".fbin") else QUERY_LIBSVM)
    n = ctypes.c_uint64()
    _hcheck(host_lib().song_query_read(path.encode(), fmt, dim, 0, None, ctypes.byref(n)))
    out = np.zeros((max(1, n.value), dim), dtype=np.float32)
    _hcheck(host_lib().song_query_read(path.encode(), fmt, dim, out.shape[0], _ptr(out), ctypes.byref(n)))
    return out[: n.value]
