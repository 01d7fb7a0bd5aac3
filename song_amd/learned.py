"""Learned form: an MLP that predicts, per query, start vertices and a stop step.

No counterpart in the reference, whose search always starts at vertex 0
(graph.h:171-181, warp_astar_accelerator.h:166-176) and always runs to its
pq-based stop rule (graph.h:188).

Pipeline (all heavy work on the GPU):
  pivots   k-means (n_pivots centroids) over a sample of the base rows; each
           centroid maps to its nearest base vertex, the pivot's start vertex.
  traces   the plain search (from vertex 0) of the training queries; the label
           of a query is the centroid nearest to its best result, and its stop
           target is log2(1 + settle step) of a search started at the label's
           neighbourhood (settle step = the last expansion that changed the
           top-k, reported by the search kernel).
  network  song_learned.h: dim + 8 features -> hidden -> hidden -> n_pivots + 1,
           BF16 MFMA forward/backward kernels, Adam.
  search   confident queries (softmax mass of the best pivot's neighbourhood --
           it and its 3 nearest pivots -- at least the threshold) start at the
           neighbourhood's pivot vertices with budget pq and an expansion cap
           from the stop head; the rest fall back to the plain full-budget
           search from vertex 0, optionally under one expansion cap calibrated
           on training queries (calibrate_fallback_cap: a high quantile of the
           step at which the plain search's top-k settles), which cuts the
           latency tail of searches that keep expanding long after their top-k
           is final and leaves every query that settles within it with exactly
           the plain search's top-k.  LearnedSearch.search runs the two groups as
           two calls of song_search_device_ex with query subsets, on two
           streams; LearnedBatch turns the split into per-query (pq, entries,
           cap) arrays on the GPU and runs every query in one batched search
           (song_batch_run: one tier per budget), so the whole pipeline --
           features, fused inference, plan, search -- is one hipGraph replay.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import FLAG_QUEUE_OVF, FLAG_VISITED_OVF, MODE_FP32, Batch, SongError, _check, lib, tier


def _mlp_lib():
    L = lib()
    if not hasattr(L, "_mlp_bound"):
        vp, i32, u64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64
        L.song_mlp_last_error.restype = ctypes.c_char_p
        L.song_mlp_create.argtypes = [i32, i32, i32, i32, u64, ctypes.POINTER(vp)]
        L.song_mlp_destroy.argtypes = [vp]
        L.song_mlp_input_width.argtypes = [vp]
        L.song_mlp_param_count.argtypes = [vp, ctypes.POINTER(u64)]
        L.song_mlp_get_params.argtypes = [vp, vp]
        L.song_mlp_set_params.argtypes = [vp, vp]
        L.song_mlp_loss_grad.argtypes = [vp, vp, vp, vp, i32, ctypes.c_float, vp, vp, vp]
        L.song_mlp_adam_step.argtypes = [vp, ctypes.c_float, vp]
        L.song_mlp_infer.argtypes = [vp, vp, i32, vp, vp, vp, vp, vp]
        L.song_mlp_infer_nbr.argtypes = [vp, vp, i32, vp, i32, vp, vp, vp, vp, vp]
        L.song_mlp_features.argtypes = [vp, i32, i32, vp, vp, vp, vp, vp, i32, vp]
        f32 = ctypes.c_float
        L.song_mlp_plan.argtypes = [vp, vp, vp, i32, vp, i32, vp, i32, f32, f32, f32, i32, i32, i32, vp, vp, vp, vp, vp]
        L.song_gemm_bf16.argtypes = [i32, i32, i32, vp, i32, i32, vp, i32, i32, vp, i32, ctypes.c_float, vp]
        L._mlp_bound = True
    return L


def _mcheck(rc):
    if rc != 0:
        raise SongError(f"song mlp error {rc}: {_mlp_lib().song_mlp_last_error().decode(errors='replace')}")


def _dp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def gemm_bf16(a, b, sam, sak, sbk, sbn, M, N, K, alpha=1.0):
    """Test hook: C = alpha * A . B with bf16 tensors (int16 storage) and strides."""
    import torch

    c = torch.empty((M, N), dtype=torch.float32, device=a.device)
    _mcheck(_mlp_lib().song_gemm_bf16(M, N, K, _dp(a), sam, sak, _dp(b), sbk, sbn, _dp(c), N, alpha, _stream()))
    return c


class MLP:
    """ctypes binding of song_mlp (include/song_learned.h)."""

    def __init__(self, in_dim, hidden=256, n_classes=4096, max_batch=1024, seed=0):
        h = ctypes.c_void_p()
        _mcheck(_mlp_lib().song_mlp_create(in_dim, hidden, n_classes, max_batch, seed, ctypes.byref(h)))
        self._h = h
        self.in_dim, self.hidden, self.n_classes, self.max_batch = in_dim, hidden, n_classes, max_batch
        self.width = _mlp_lib().song_mlp_input_width(h)
        n = ctypes.c_uint64()
        _mcheck(_mlp_lib().song_mlp_param_count(h, ctypes.byref(n)))
        self.n_params = n.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _mlp_lib().song_mlp_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def get_params(self):
        p = np.empty(self.n_params, dtype=np.float32)
        _mcheck(_mlp_lib().song_mlp_get_params(self._h, ctypes.c_void_p(p.ctypes.data)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _mcheck(_mlp_lib().song_mlp_set_params(self._h, ctypes.c_void_p(p.ctypes.data)))

    def loss_grad(self, x, labels, stop_t, stop_weight=0.1, want_grads=False):
        loss = np.zeros(2, dtype=np.float32)
        g = np.empty(self.n_params, dtype=np.float32) if want_grads else None
        _mcheck(_mlp_lib().song_mlp_loss_grad(self._h, _dp(x), _dp(labels), _dp(stop_t), x.shape[0], stop_weight,
                                              ctypes.c_void_p(loss.ctypes.data),
                                              ctypes.c_void_p(g.ctypes.data) if g is not None else None, _stream()))
        return float(loss[0]), float(loss[1]), g

    def train_step(self, x, labels, stop_t, lr, stop_weight=0.1):
        _mcheck(_mlp_lib().song_mlp_loss_grad(self._h, _dp(x), _dp(labels), _dp(stop_t), x.shape[0], stop_weight,
                                              None, None, _stream()))
        _mcheck(_mlp_lib().song_mlp_adam_step(self._h, lr, _stream()))

    def infer(self, x, logits=False, nbr=None):
        """Fused inference: (top2 pivots, confidence, stop[, logits]).  With nbr (int32
        [n_classes, E]) the confidence is the mass of the best pivot's neighbourhood."""
        import torch

        B = x.shape[0]
        top2 = torch.empty((B, 2), dtype=torch.int32, device=x.device)
        conf = torch.empty(B, dtype=torch.float32, device=x.device)
        stop = torch.empty(B, dtype=torch.float32, device=x.device)
        lg = torch.empty((B, self.n_classes + 1), dtype=torch.float32, device=x.device) if logits else None
        _mcheck(_mlp_lib().song_mlp_infer_nbr(self._h, _dp(x), B, _dp(nbr), nbr.shape[1] if nbr is not None else 0,
                                              _dp(top2), _dp(conf), _dp(stop), _dp(lg), _stream()))
        return (top2, conf, stop, lg) if logits else (top2, conf, stop)


def kmeans(x, k, iters=10, seed=0, chunk=1 << 16):
    """Lloyd's k-means on a float32 cuda tensor; returns (centroids [k, d], counts [k])."""
    import torch

    g = torch.Generator(device="cpu").manual_seed(seed)
    c = x[torch.randperm(x.shape[0], generator=g)[:k].to(x.device)].clone()
    for _ in range(iters):
        cn = (c * c).sum(1)
        assign = torch.empty(x.shape[0], dtype=torch.int64, device=x.device)
        for s in range(0, x.shape[0], chunk):
            xs = x[s:s + chunk]
            assign[s:s + chunk] = torch.argmin(cn[None, :] - 2.0 * (xs @ c.T), dim=1)
        sums = torch.zeros_like(c).index_add_(0, assign, x)
        cnt = torch.bincount(assign, minlength=k).to(x.dtype)
        nz = cnt > 0
        c[nz] = sums[nz] / cnt[nz, None]
    return c, cnt


def nearest_rows(base, points, chunk=1 << 18):
    """Index of the nearest base row (L2) to each point; base/points float32 cuda tensors."""
    import torch

    pn = (points * points).sum(1)
    best_d = torch.full((points.shape[0],), float("inf"), device=points.device)
    best_i = torch.zeros(points.shape[0], dtype=torch.int64, device=points.device)
    for s in range(0, base.shape[0], chunk):
        b = base[s:s + chunk]
        d = pn[:, None] + (b * b).sum(1)[None, :] - 2.0 * (points @ b.T)
        v, i = torch.min(d, dim=1)
        upd = v < best_d
        best_d = torch.where(upd, v, best_d)
        best_i = torch.where(upd, i + s, best_i)
    return best_i


def nearest_centroids(points, c, k=2, chunk=1 << 15):
    import torch

    cn = (c * c).sum(1)
    out = torch.empty((points.shape[0], k), dtype=torch.int64, device=points.device)
    for s in range(0, points.shape[0], chunk):
        d = cn[None, :] - 2.0 * (points[s:s + chunk] @ c.T)
        out[s:s + chunk] = torch.topk(d, k, dim=1, largest=False).indices
    return out


class LearnedSearch:
    """Entry/stop predictor over one index (see the module docstring)."""

    def __init__(self, index, base, n_pivots=4096, hidden=256, seed=0, n_nbr=4):
        import torch

        self.index = index
        self.dev = torch.device("cuda:0")
        self.base_dev = torch.from_numpy(np.ascontiguousarray(base, dtype=np.float32)).to(self.dev)
        self.n_pivots, self.hidden, self.seed, self.n_nbr = n_pivots, hidden, seed, n_nbr
        self.dim = base.shape[1]
        self.mlp = None
        self.log = {}

    # ---------------------------------------------------------------- features
    def features(self, q_dev):
        import torch

        x = torch.empty((q_dev.shape[0], self.mlp.width), dtype=torch.int16, device=self.dev)
        fs = np.ascontiguousarray(self.fscale, dtype=np.float32)
        _mcheck(_mlp_lib().song_mlp_features(_dp(q_dev), q_dev.shape[0], self.dim, _dp(self.mean), _dp(self.inv_std),
                                             _dp(self.v0), ctypes.c_void_p(fs.ctypes.data), _dp(x), self.mlp.width,
                                             _stream()))
        return x

    def _feature_stats(self, q):
        import torch

        self.mean = q.mean(0).contiguous()
        self.inv_std = (1.0 / q.std(0).clamp_min(1e-6)).contiguous()
        self.v0 = self.base_dev[0].contiguous()
        f = torch.stack([q.norm(dim=1), (q - self.v0).norm(dim=1), q.mean(1), q.std(1, unbiased=False),
                         q.min(1).values, q.max(1).values, (q == 0).float().mean(1)], 1)
        sd = f.std(0).clamp_min(1e-6).cpu().numpy()
        self.fscale = np.concatenate([1.0 / sd, [1.0]]).astype(np.float32)

    # ---------------------------------------------------------------- training
    def fit(self, train_queries, pq_label=256, k=10, epochs=8, batch=1024, lr=2e-3, stop_weight=0.05,
            sample=200_000, verbose=True):
        import torch

        t0 = time.time()
        dev = self.dev
        g = torch.Generator(device="cpu").manual_seed(self.seed)
        n = self.base_dev.shape[0]
        samp = self.base_dev[torch.randperm(n, generator=g)[:min(sample, n)].to(dev)]
        self.centroids, _ = kmeans(samp, self.n_pivots, iters=10, seed=self.seed)
        self.pivot_vertex = nearest_rows(self.base_dev, self.centroids).to(torch.int32)
        # neighbourhood of each pivot: itself and its n_nbr - 1 nearest centroids
        self.nbr = nearest_centroids(self.centroids, self.centroids, self.n_nbr).to(torch.int32).contiguous()
        t1 = time.time()

        q = torch.from_numpy(np.ascontiguousarray(train_queries, dtype=np.float32)).to(dev)
        nt = q.shape[0]
        ids = torch.empty((nt, k), dtype=torch.int32, device=dev)
        self.index.search_device(q, pq_label, k, ids)
        top1 = self.base_dev[ids[:, 0].long()]
        labels = nearest_centroids(top1, self.centroids, 1)[:, 0].to(torch.int32).contiguous()
        entry = self.pivot_vertex[self.nbr[labels.long()].long()].contiguous()
        stats = torch.empty((nt, 8), dtype=torch.int32, device=dev)
        ids2 = torch.empty((nt, k), dtype=torch.int32, device=dev)
        self.index.search_device(q, pq_label, k, ids2, stats=stats, entry=entry)
        settle = stats[:, 4].float()
        stop_t = torch.log2(1.0 + settle).contiguous()
        t2 = time.time()

        self._feature_stats(q)
        self.mlp = MLP(self.dim + 8, self.hidden, self.n_pivots, max_batch=batch, seed=self.seed)
        X = self.features(q)
        nb = nt // batch
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(nt, generator=g).to(dev)
            lr_ep = lr * 0.5 * (1 + math.cos(math.pi * ep / epochs))
            for b in range(nb):
                sel = perm[b * batch:(b + 1) * batch]
                self.mlp.train_step(X[sel].contiguous(), labels[sel].contiguous(), stop_t[sel].contiguous(), lr_ep,
                                    stop_weight)
            xe, ms, _ = self.mlp.loss_grad(X[:batch].contiguous(), labels[:batch].contiguous(),
                                           stop_t[:batch].contiguous(), stop_weight)
            hist.append((xe, ms))
            if verbose:
                print(f"[learned] epoch {ep} xent={xe:.3f} stop_mse={ms:.3f}", flush=True)
        torch.cuda.synchronize()
        top2, conf, _ = self.mlp.infer(X[:min(nt, 20000)], nbr=self.nbr)
        acc = (top2[:, 0] == labels[:top2.shape[0]]).float().mean().item()
        self.log = {"kmeans_s": t1 - t0, "traces_s": t2 - t1, "train_s": time.time() - t2, "loss": hist,
                    "train_top1_acc": acc, "mean_conf": conf.mean().item(),
                    "mean_settle": settle.mean().item(), "pq_label": pq_label}
        return self.log

    # ---------------------------------------------------------------- search
    def predict(self, q_dev):
        """(start vertices [nq, n_nbr], confidence, stop output) of a batch."""
        x = self.features(q_dev)
        top2, conf, stop = self.mlp.infer(x, nbr=self.nbr)
        entry = self.pivot_vertex[self.nbr[top2[:, 0].long()].long()].contiguous()
        return entry, conf, stop

    def plan_arrays(self, q_dev, conf_threshold, step_scale, step_add, fallback_cap, pq_learned, pq_fallback):
        """(pq [nq], entries [nq, n_nbr] (-1 rows for the fallback), expansion caps [nq], confident
        mask) on the GPU: features, fused inference, then one plan kernel (song_mlp_plan)."""
        import torch

        x = self.features(q_dev)
        top2, conf, stop = self.mlp.infer(x, nbr=self.nbr)
        nq = q_dev.shape[0]
        pq = torch.empty(nq, dtype=torch.int32, device=self.dev)
        entry = torch.empty((nq, self.nbr.shape[1]), dtype=torch.int32, device=self.dev)
        caps = torch.empty(nq, dtype=torch.int32, device=self.dev)
        ok = torch.empty(nq, dtype=torch.bool, device=self.dev)
        _mcheck(_mlp_lib().song_mlp_plan(_dp(top2), _dp(conf), _dp(stop), nq, _dp(self.nbr), self.nbr.shape[1],
                                         _dp(self.pivot_vertex), self.pivot_vertex.shape[0], float(conf_threshold),
                                         float(step_scale), float(step_add), int(fallback_cap or 0), int(pq_learned),
                                         int(pq_fallback), _dp(pq), _dp(entry), _dp(caps), _dp(ok), _stream()))
        return pq, entry, caps, ok

    def plan(self, q_dev, conf_threshold=0.9, step_scale=1.5, step_add=8):
        """Entries, expansion caps and the confident / fallback split for a batch (entries and caps
        are used for the confident queries)."""
        import torch

        _, entry, caps, ok = self.plan_arrays(q_dev, conf_threshold, step_scale, step_add, 0, 1, 1)
        confident = torch.nonzero(ok).flatten().to(torch.int32)
        fallback = torch.nonzero(~ok).flatten().to(torch.int32)
        return entry, caps, confident, fallback

    def calibrate_fallback_cap(self, queries, pq, k=10, conf_threshold=0.9, quantile=0.999, max_queries=20000):
        """Expansion cap for fallback queries: the `quantile` of the step at which the plain
        search's top-k settles (its last change; the kernel's per-query stats), over training
        queries the predictor sends to the fallback (all of them if fewer than 500 do).  A query
        whose top-k settles within the cap returns exactly the plain search's top-k."""
        import torch

        qd = torch.from_numpy(np.ascontiguousarray(queries[:max_queries], dtype=np.float32)).to(self.dev)
        _, conf, _ = self.predict(qd)
        low = torch.nonzero(conf < conf_threshold).flatten()
        if low.numel() >= 500:
            qd = qd[low].contiguous()
        n = qd.shape[0]
        ids = torch.empty((n, k), dtype=torch.int32, device=self.dev)
        flags = torch.empty(n, dtype=torch.int32, device=self.dev)
        st = torch.empty((n, 8), dtype=torch.int32, device=self.dev)
        self.index.search_device(qd, pq, k, ids, flags=flags, stats=st, mode=MODE_FP32)
        torch.cuda.synchronize()
        settle = st[:, 4].cpu().numpy().astype(np.int64)
        cap = int(math.ceil(float(np.quantile(settle, quantile)))) if quantile < 1.0 else int(settle.max())
        self.fallback_cap = max(cap, 1)
        return {"cap": self.fallback_cap, "quantile": quantile, "queries": n,
                "low_confidence_only": bool(low.numel() >= 500), "settle_mean": round(float(settle.mean()), 1),
                "settle_max": int(settle.max()), "expansions_mean": round(float(st[:, 0].float().mean().item()), 1)}

    def search(self, q_dev, pq, k, ids, plan=None, pq_fallback=None, conf_threshold=0.9, step_scale=1.5, step_add=8,
               storage=0, dists=None, stats=None, fallback_cap=None, flags=None):
        """Learned search of a batch into ids (int32 cuda [nq, >=k]); returns the plan.
        fallback_cap: expansion cap of the fallback queries (None: uncapped plain search).
        flags (int32 cuda [nq], optional) receives each query's flags.  The two groups run
        side by side without host waits; queries whose visited table or queue overflowed are
        then searched again with the library's re-run (larger tables), so no result is
        truncated."""
        if plan is None:
            plan = self.plan(q_dev, conf_threshold, step_scale, step_add)
        import torch

        entry, steps, confident, fallback = plan
        nq = q_dev.shape[0]
        if flags is None:
            flags = torch.zeros(nq, dtype=torch.int32, device=self.dev)
        main = torch.cuda.current_stream()
        fsteps = None if fallback_cap is None else torch.full_like(steps, int(fallback_cap))
        if fallback.numel():
            # the fallback group runs beside the confident one on a second stream
            side = self._side_stream()
            side.wait_stream(main)
            with torch.cuda.stream(side):
                self.index.search_device(q_dev, pq_fallback or pq, k, ids, dists=dists, flags=flags, stats=stats,
                                         subset=fallback, storage=storage, retry=0, steps=fsteps)
        if confident.numel():
            self.index.search_device(q_dev, pq, k, ids, dists=dists, flags=flags, stats=stats, entry=entry,
                                     steps=steps, subset=confident, storage=storage, retry=0)
        if fallback.numel():
            main.wait_stream(self._side_stream())
        ovf = (flags & (FLAG_VISITED_OVF | FLAG_QUEUE_OVF)) != 0
        if bool(ovf.any().item()):
            conf_mask = torch.zeros(nq, dtype=torch.bool, device=self.dev)
            conf_mask[confident.long()] = True
            redo_f = torch.nonzero(ovf & ~conf_mask).flatten().to(torch.int32)
            redo_c = torch.nonzero(ovf & conf_mask).flatten().to(torch.int32)
            if redo_f.numel():
                self.index.search_device(q_dev, pq_fallback or pq, k, ids, dists=dists, flags=flags, stats=stats,
                                         subset=redo_f, storage=storage, retry=1, steps=fsteps)
            if redo_c.numel():
                self.index.search_device(q_dev, pq, k, ids, dists=dists, flags=flags, stats=stats, entry=entry,
                                         steps=steps, subset=redo_c, storage=storage, retry=1)
        return plan

    def _side_stream(self):
        import torch

        if getattr(self, "_side", None) is None:
            self._side = torch.cuda.Stream()
        return self._side


class LearnedBatch:
    """The learned form as one batched search over a fixed query buffer.

    Per query, on the GPU: confident -> (pq_learned, its neighbourhood's pivot vertices,
    expansion cap from the stop head); fallback -> (pq_fallback, vertex 0 -- entry list
    padded with -1 --, fallback_cap or no cap).  The batch has one tier per budget (tier knobs from the
    autotuner or defaults).  run() enqueues the pipeline; capture() records it once in a
    CUDA/HIP graph and replay() relaunches it."""

    def __init__(self, ls, nq, pq_learned, pq_fallback, k=10, conf_threshold=0.9, step_scale=1.5, step_add=8,
                 storage=0, mode=MODE_FP32, tiers=None, k_ld=None, fallback_cap=None, **batch_kw):
        import torch

        self.ls, self.nq, self.k = ls, nq, k
        self.pq_l, self.pq_f = int(pq_learned), int(pq_fallback)
        self.conf_threshold, self.step_scale, self.step_add = conf_threshold, step_scale, step_add
        self.fallback_cap = fallback_cap
        dev = ls.dev
        lo, hi = sorted({self.pq_l, self.pq_f}) if self.pq_l != self.pq_f else (self.pq_l, None)
        self.tiers = tiers or ([tier(lo)] + ([tier(hi)] if hi else []))
        self.batch = Batch(ls.index, self.tiers, nq, mode=mode, storage=storage, **batch_kw)
        self.k_ld = k_ld or k
        self.ids = torch.empty((nq, self.k_ld), dtype=torch.int32, device=dev)
        self.flags = torch.empty(nq, dtype=torch.int32, device=dev)
        self.k_arr = torch.full((nq,), k, dtype=torch.int32, device=dev)
        self.graph = None

    def plan(self, q_dev):
        """(pq [nq], entries [nq, E] with -1 padding, caps [nq], confident mask) on the GPU."""
        return self.ls.plan_arrays(q_dev, self.conf_threshold, self.step_scale, self.step_add, self.fallback_cap,
                                   self.pq_l, self.pq_f)

    def run(self, q_dev, stats=None, dists=None):
        """Enqueue the whole pipeline for q_dev (nq queries); returns the plan."""
        p = self.plan(q_dev)
        self.batch.run(q_dev, self.ids, self.flags, k=self.k, pq_arr=p[0], k_arr=self.k_arr, entry=p[1], steps=p[2],
                       stats=stats, dists=dists)
        return p

    def capture(self, q_dev):
        """Record run(q_dev) in a graph (warm-up first); replay() relaunches it on the
        current contents of q_dev."""
        import torch

        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2):
                self.run(q_dev)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._plan_out = self.run(q_dev)
        torch.cuda.synchronize()
        return self._plan_out

    def replay(self):
        self.graph.replay()
