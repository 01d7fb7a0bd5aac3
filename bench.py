#!/usr/bin/env python
"""Benchmark: QPS at recall@10 >= 0.95 on the 1M x 128 L2 workload, one GPU,
learned + BF16 batched form against the plain form.

Workload: 1,000,000 base vectors, dim 128, integer-valued mixture of 1,000
Gaussians (seed 2), 10,000 queries from the same mixture, L2, degree-31 graph
from the reference insertion build (byte-identical, built by the parallel
builder), top-10.  A "step" is one search of the whole 10,000-query batch with
queries, index and outputs resident in HBM.

Forms timed:
  plain          the bit-exact search from vertex 0, one launch per batch (one
                 workgroup per query); the smallest pq (a coarse grid, then a
                 bisection in steps of 16) whose recall@10 on the selection split
                 (tie-aware, against a float64 brute force on the GPU) reaches the
                 target.  The ratio's denominator.
  plain batched  the same pq through the batching form (autotuned persistent
                 kernel, hipGraph replay); ids checked equal to the plain form's.
  learned+BF16   BF16 rows with exact FP32 re-rank, MLP start vertices and stop
  batched        caps (trained here on 200,000 trace-labelled queries from the same
                 mixture, disjoint from the benchmark queries), confidence fallback
                 to the plain full-budget search; every query in one batched search
                 (per-query pq / entries / caps made on the GPU, autotuned tiers),
                 the whole pipeline -- features, MLP, plan, search -- one graph
                 replay per step.  `value` is its QPS at the setting chosen on the
                 selection split, with the fallback uncapped; the same setting with
                 the fallback under a training-calibrated expansion cap is reported
                 beside it.
Every setting (the plain pq, the learned pq and stop scale, the autotuned tiers) is chosen
on a held-out selection split -- 10,000 more queries of the same mixture on their own
random stream -- and then timed and scored on the test queries.

`--workload c4` measures the batching form on its own workload instead (2M x 96
inner product, 100,000 queries with uneven k and pq); `--workload c3` the precision
form (factor-model 960-d L2: recall deltas and distance errors of BF16 / FP8 rows);
`--workload c5` the C2 measurement on the same generator at 10,000,000 rows and
100,000 queries; `--workload c1` the 15k x 16 cosine workload with one and with several
queries per wavefront.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_search_l2_c2_1m.json")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def timed(fn, reps, torch):
    """median milliseconds of reps runs of fn, each between synchronizations"""
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(times))


def timed_steps(fn, steps, warmup, torch, stream=None):
    """W untimed calls, then K calls between synchronizations; ms per call from events on
    `stream` (the stream the work is launched on) and from the host clock."""
    stream = stream or torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return e0.elapsed_time(e1) / steps, wall


def alg_bytes(st, dim, row_bytes, nq):
    """Algorithmic bytes of one search launch from per-query counters: every distance
    reads one row, every expansion reads (1 + degree) u32 adjacency words, every query
    reads its vector."""
    st64 = st.astype(np.int64)
    return int(st64[:, 1].sum() * row_bytes + st64[:, 2].sum() * 4 + nq * dim * 4)


def load_traffic(pq):
    """HBM bytes per launch of the plain kernel from the committed rocprofv3 counter
    summary (the kernel's L2-to-fabric read and write requests, each times its size)
    when it was measured at this pq; else None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t if t.get("pq") == pq else None


def cpu_baseline(base, graph, metric, query, groups, seconds, check=None, repeats=3, one_core_queries=1000):
    """Oracle restatement of graph.h:169-216 on the host cores over a bounded sample, the median of
    `repeats` timed passes; plus a single-core figure over `one_core_queries` queries (median of
    the same repeats).  groups: list of (pq, query index array); the sample takes the same share of
    each.  Threads: min(16, cores) -- one GPU's share of the host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    nproc = os.cpu_count() or 1
    threads = min(16, nproc)
    total_q = sum(len(idx) for _, idx in groups)

    def one_pass(take_frac, nthreads, check_ids):
        done, tcpu, same = 0, 0.0, True
        for pq, idx in groups:
            take = idx[:max(nthreads, int(round(len(idx) * take_frac)))]
            tc = time.perf_counter()
            cids, _ = pyoracle.search_batch(base, graph, metric, query[take], pq, threads=nthreads)
            tcpu += time.perf_counter() - tc
            done += len(take)
            if check_ids and check is not None:
                same &= check(pq, take, cids)
        return done, tcpu, same

    # single core: the same share of every group, about one_core_queries queries in all
    f1 = min(1.0, one_core_queries / max(total_q, 1))
    one = [one_pass(f1, 1, False) for _ in range(repeats)]
    q1 = one[0][0]
    rate1 = float(np.median([d / t for d, t, _ in one]))
    # all threads: as much of the batch as fits `seconds` of CPU time at the single-core rate
    frac = min(1.0, seconds * threads * rate1 / max(total_q, 1))
    runs = [one_pass(frac, threads, i == 0) for i in range(repeats)]
    done = runs[0][0]
    rates = [d / t for d, t, _ in runs]
    return {"value": round(float(np.median(rates)), 1), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"{done} of the {total_q} queries (same share of every pq group), oracle restatement of "
                      f"graph.h:169-216 on {threads} threads (static split), median of {repeats} passes "
                      f"(passes: {', '.join(f'{r:.0f}' for r in rates)} q/s); single core: {q1} queries, "
                      f"median of {repeats} passes",
            "one_core_qps": round(rate1, 1), "repeats": repeats, "nproc": nproc, "cpu_model": cpu_model(),
            "ids_equal_gpu": bool(runs[0][2]) if check is not None else None}


def bench_c2(args, torch, song_amd):
    from song_amd import autotune, datasets, groundtruth, workloads

    dev = torch.device("cuda:0")
    t0 = time.time()
    if args.workload == "c5":
        # the same generator at 10M rows and 100k queries (the learned form's large workload)
        base, query = datasets.mixture_l2(n_base=args.c5_n, n_query=args.c5_nq)
        graph, bst = workloads.graph_for(base, "l2", verbose=True)
        metric = "l2"
    else:
        base, query, graph, metric, bst = workloads.load("c2_1m", verbose=True)
    log(f"workload ready in {time.time() - t0:.1f}s (graph cached={bst.get('cached')})")
    ix = song_amd.Index(base, graph, metric)
    nq, dim = query.shape
    k = args.k
    stream = torch.cuda.current_stream()

    if args.copy_mb:
        src = torch.ones(args.copy_mb << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {src.numel() * 4} bytes read, {dst.numel() * 4} bytes written")
        del src, dst

    # held-out selection split: queries from the same mixture as the test queries, on their own
    # stream (the learned form trains on stream 1); every setting is chosen on this split and then
    # timed and scored on the 10,000 test queries
    selq = datasets.mixture_l2_queries(args.sel_queries or len(query), seed=2, stream=3)
    gt_d = gs_d = None
    if not args.no_recall:
        t1 = time.time()
        _, gt_d = groundtruth.knn_fp64(base, query, k, metric)
        _, gs_d = groundtruth.knn_fp64(base, selq, k, metric)
        log(f"fp64 ground truth (test and selection queries) in {time.time() - t1:.1f}s")

    def recall(ids_t, gd, qs):
        if gd is None:
            return float("nan")
        return groundtruth.recall_at_k(ids_t.cpu().numpy().astype(np.int64), gd, k, base, qs, metric)

    q = torch.from_numpy(query).to(dev)
    qs = torch.from_numpy(selq).to(dev)
    ns = qs.shape[0]
    flags = torch.empty(nq, dtype=torch.int32, device=dev)
    stats = torch.empty((nq, 8), dtype=torch.int32, device=dev)
    ids = torch.empty((nq, k), dtype=torch.int32, device=dev)
    s_ids = torch.empty((ns, k), dtype=torch.int32, device=dev)
    s_flags = torch.empty(ns, dtype=torch.int32, device=dev)
    # integer-valued data and queries: FP32 accumulation is exact (same ids as the exact mode)
    mode = song_amd.resolve_mode(ix, q)

    def plain(qd, out, fl, pq, st=None, retry=0):
        ix.search_device(qd, pq, k, out, None, fl, st, mode=mode, retry=retry)

    # ---------------------------------------------------------------- plain form
    sweep = []

    def try_pq(pq):
        plain(qs, s_ids, s_flags, pq, retry=1)
        torch.cuda.synchronize()
        r = recall(s_ids, gs_d, selq)
        ms = timed(lambda: plain(qs, s_ids, s_flags, pq), 3, torch)
        sweep.append({"pq": pq, "recall@10_selection": round(r, 4), "qps_selection": round(ns / (ms / 1e3), 1)})
        log(f"plain pq={pq} selection recall@{k}={r:.4f} ({ns / (ms / 1e3):.0f} q/s)")
        return r >= sel_target

    # settings must reach the target on the selection split with a margin, so that the test split
    # (another sample of the same mixture) reaches it too
    sel_target = args.target_recall + args.selection_margin
    if args.pq:
        pq = args.pq
    elif args.no_recall:
        ap_error("--no-recall needs --pq")
    else:
        # coarse grid up to the first budget reaching the target on the selection split, then a
        # bisection on multiples of --pq-step between the last miss and that budget
        grid = [int(x) for x in args.pqs.split(",")]
        lo, hi = 0, None
        for g in grid:
            if try_pq(g):
                hi = g
                break
            lo = g
        if hi is None:
            pq = grid[-1]
            log(f"plain form did not reach the target recall on the selection split; timing pq={pq}")
        else:
            step = args.pq_step
            while hi - lo > step:
                mid = lo + max(step, ((hi - lo) // 2) // step * step)
                if mid >= hi:
                    break
                if try_pq(mid):
                    hi = mid
                else:
                    lo = mid
            pq = hi
    plain(q, ids, flags, pq, stats, retry=1)
    torch.cuda.synchronize()
    r = recall(ids, gt_d, query)
    st = stats.cpu().numpy()
    plain_ids = ids.clone()
    retry = 1 if int((flags.cpu().numpy() & 3).sum()) > 0 else 0

    ms_plain, wall_plain = timed_steps(lambda: plain(q, ids, flags, pq, retry=retry), args.steps, args.warmup,
                                       torch, stream)
    qps_plain = nq / (ms_plain / 1e3)
    plain_info = {"pq": pq, "recall@10": round(r, 4), "qps": round(qps_plain, 1), "ms_per_step": round(ms_plain, 4),
                  "wall_ms_per_step": round(wall_plain, 4), "mean_expansions": round(float(st[:, 0].mean()), 1),
                  "mean_distances": round(float(st[:, 1].mean()), 1),
                  "max_expansions": int(st[:, 0].max()), "mean_head_refills": round(float(st[:, 5].mean()), 2),
                  "mean_bag_truncations": round(float(st[:, 6].mean()), 2),
                  "arithmetic": "fp32 (exact on integer data)",
                  "pq_chosen_on": "selection split" if not args.pq else "--pq"}
    log(f"plain form: {qps_plain:.0f} q/s at pq={pq}, test recall={r:.4f}")
    alg_plain = alg_bytes(st, dim, dim * 4, nq)
    # host buffers in and out (song_search: device buffers allocated per call, queries up,
    # ids and distances down over PCIe) -- reported beside the resident-input rate
    ms_host = timed(lambda: ix.search(query, pq, k, mode=mode), 3, torch)
    plain_info["host_buffers_qps"] = round(nq / (ms_host / 1e3), 1)
    log(f"plain form with host buffers: {plain_info['host_buffers_qps']:.0f} q/s")

    # ---------------------------------------------------------------- plain, batched
    batched_info = None
    if not args.no_batch:
        # tuned on the selection queries, timed on the test queries
        pq_s = torch.full((ns,), pq, dtype=torch.int32, device=dev)
        k_s = torch.full((ns,), k, dtype=torch.int32, device=dev)
        tb = time.time()
        table = autotune.tune(ix, qs, pq_s, k_s, mode=mode, reps=3, max_queries_per_tier=min(ns, 20000), k_ld=k)
        log(f"plain batched: autotuned in {time.time() - tb:.1f}s: {table['tiers']}")
        b = song_amd.Batch(ix, table["tiers"], nq, mode=mode)
        side = torch.cuda.Stream()
        bids = torch.empty_like(ids)
        bfl = torch.empty_like(flags)
        side.wait_stream(stream)
        b.run(q, bids, bfl, pq=pq, k=k, graph=True, stream=side.cuda_stream)
        torch.cuda.synchronize()
        same = bool(torch.equal(bids, plain_ids))
        ms_b, wall_b = timed_steps(lambda: b.run(q, bids, bfl, pq=pq, k=k, graph=True, stream=side.cuda_stream),
                                   args.steps, args.warmup, torch, side)
        batched_info = {"pq": pq, "qps": round(nq / (ms_b / 1e3), 1), "ms_per_step": round(ms_b, 4),
                        "wall_ms_per_step": round(wall_b, 4), "ids_equal_plain": same, "tiers": table["tiers"],
                        "grid": [b.grid(t) for t in range(len(table["tiers"]) + 1)]}
        log(f"plain batched: {batched_info['qps']:.0f} q/s, ids equal plain: {same}")
        b.close()

    # ---------------------------------------------------------------- learned + BF16, batched
    learned_info = None
    if not args.no_learned and not args.no_recall:
        learned_info = bench_learned(args, torch, song_amd, ix, base, q, query, qs, selq, gt_d, gs_d, pq, k, dim,
                                     plain_ids, qps_plain, recall, stream, sel_target)

    # ---------------------------------------------------------------- the fallback tier alone
    # the learned step ends with its fallback tier (the plain search of the low-confidence
    # queries); here that tier's configuration runs alone on as many test queries (their
    # expansion counts are those of typical queries), with one and with two waves per query
    fb_alone = None
    if args.fallback_n > 0:
        fb_tier = None
        n_fb = args.fallback_n
        if learned_info and "tiers" in learned_info and len(learned_info["tiers"]) > 1:
            fb_tier = dict(learned_info["tiers"][1])
            n_fb = learned_info["fallback"]["queries"] or n_fb
        if fb_tier is None:
            fb_tier = song_amd.tier(pq, hash_log2=14, slack=128, waves_per_group=4, visited_bits=32,
                                    storage=song_amd.STORE_F32, waves_per_query=2)
        n_fb = min(n_fb, nq)
        fb_alone = {"queries": n_fb, "tier": fb_tier}
        side = torch.cuda.Stream()
        qf = q[:n_fb].contiguous()
        for wpq in (1, 2):
            # one query per workgroup in both (a two-wave tier runs one search wave per group)
            t = dict(fb_tier, waves_per_query=wpq, waves_per_group=1)
            b = song_amd.Batch(ix, [t], n_fb, mode=mode)
            fid = torch.empty((n_fb, k), dtype=torch.int32, device=dev)
            ffl = torch.empty(n_fb, dtype=torch.int32, device=dev)
            side.wait_stream(stream)
            b.run(qf, fid, ffl, pq=pq, k=k, stream=side.cuda_stream)
            torch.cuda.synchronize()
            same = bool(torch.equal(fid, plain_ids[:n_fb]))
            b.run(qf, fid, ffl, pq=pq, k=k, graph=True, stream=side.cuda_stream)
            torch.cuda.synchronize()
            same = same and bool(torch.equal(fid, plain_ids[:n_fb]))
            ms_f, _ = timed_steps(lambda: b.run(qf, fid, ffl, pq=pq, k=k, graph=True, stream=side.cuda_stream),
                                  args.steps, args.warmup, torch, side)
            b.close()
            fb_alone[f"waves_per_query_{wpq}"] = {"ms_per_step": round(ms_f, 4), "ids_equal_plain": same}
            log(f"fallback tier alone ({n_fb} queries, {wpq} wave(s) per query): {ms_f:.3f} ms, ids equal plain: {same}")
        fb_alone["max_expansions"] = int(st[:n_fb, 0].max())

    # ---------------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu:
        pids = plain_ids.cpu().numpy().astype(np.int64)
        cpu = cpu_baseline(base, graph, metric, query, [(pq, np.arange(nq))], args.cpu_seconds,
                           check=lambda p, take, cids: bool(np.array_equal(cids[:, :k].astype(np.int64), pids[take])))
        log(f"cpu baseline: {cpu['value']} q/s on {cpu['cores']} threads, {cpu['one_core_qps']} q/s on one")

    have_learned = bool(learned_info and "qps" in learned_info)
    if have_learned:
        value, ms_step, dtype = learned_info["qps"], learned_info["ms_per_step"], "bf16"
        form = ("learned + BF16, batched (MLP start vertices and stop caps, BF16 rows, exact FP32 re-rank; "
                "confidence fallback to the plain full-budget search on FP32 rows; persistent tiers, one graph "
                "replay per step; settings chosen on a held-out selection split)")
    else:
        value, ms_step, dtype = round(qps_plain, 1), ms_plain, "fp32"
        form = "plain (bit-exact ids)"
    achieved_plain = alg_plain / (ms_plain / 1e3)
    traffic = load_traffic(pq)
    n_rows = base.shape[0]
    size = f"{n_rows // 1_000_000}M" if n_rows % 1_000_000 == 0 else f"{n_rows:,}"
    out = {
        "metric": f"QPS @ recall@10>={args.target_recall:g}, {size} x128 L2, 1 GPU: learned/BF16 form vs plain",
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic: seeded Gaussian-mixture generator (integer-valued), graph = reference insertion build",
        "config": {"workload": f"{n_rows:,} x 128 L2 mixture (1,000 centres, sigma 20, clipped/rounded), "
                               f"{nq:,} queries, top-10",
                   "form": form, "k": k, "target_recall": args.target_recall,
                   "selection_queries": ns, "selection_target": round(sel_target, 4),
                   "plain": plain_info,
                   "plain_batched": batched_info,
                   "learned_bf16_batched": learned_info,
                   "learned_vs_plain": round(value / qps_plain, 3) if have_learned else None,
                   "fallback_tier_alone": fb_alone,
                   "plain_sweep_selection": sweep},
        "roofline": {"bound": "hbm", "achieved": round(achieved_plain / 1e9, 2), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved_plain / HBM_PEAK, 5),
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "kernel": ("search_kernel<L2, fp32, F32 rows, shape 1> (plain form, one launch per step; "
                                "a call of more queries than fit the device at once keeps the best lists' "
                                "fronts in global memory)"),
                     "kernel_ms": round(ms_plain, 4),
                     "algorithmic_bytes_per_launch": alg_plain,
                     "learned_achieved_GBps": (round(learned_info["algorithmic_bytes_per_step"]
                                                     / (learned_info["ms_per_step"] / 1e3) / 1e9, 2)
                                               if have_learned else None)},
        "cpu_baseline": cpu,
    }
    return out


def ap_error(msg):
    print(json.dumps({"error": msg}))
    sys.exit(2)


def bench_learned(args, torch, song_amd, ix, base, q, query, qs, selq, gt_d, gs_d, pq, k, dim, plain_ids,
                  qps_plain, recall, stream, sel_target):
    """Learned + BF16 batched form.  Trained on 200,000 queries of the mixture (stream 1); the
    learned budget and stop scale are chosen on the selection split (stream 3), the tiers are
    autotuned there (the confident tier may take a Bloom filter if the selection recall stays at
    the target), and the chosen setting is captured and timed on the test queries.  The headline
    runs the fallback as the plain full-budget search (no expansion cap); the same setting with
    the fallback under the training-calibrated cap is measured beside it."""
    from song_amd import autotune, datasets, learned

    dev = q.device
    nq, ns = q.shape[0], qs.shape[0]
    ix.add_storage(song_amd.STORE_BF16)
    train = datasets.mixture_l2_queries(args.train_queries, seed=2, stream=1)
    ls = learned.LearnedSearch(ix, base, n_pivots=4096, hidden=256, seed=0, n_nbr=8)
    t2 = time.time()
    fit = ls.fit(train, pq_label=256, k=k, epochs=8, batch=1024, lr=2e-3, verbose=False)
    log(f"learned predictor trained in {time.time() - t2:.1f}s: top1 acc {fit['train_top1_acc']:.3f} "
        f"mean conf {fit['mean_conf']:.3f} mean settle {fit['mean_settle']:.1f}")
    fcal, fcap = None, None
    if args.fallback_quantile > 0:
        fcal = ls.calibrate_fallback_cap(train, pq, k=k, quantile=args.fallback_quantile)
        fcap = fcal["cap"]
        log(f"fallback cap {fcap} expansions (settle quantile {args.fallback_quantile} over {fcal['queries']} "
            f"training queries, settle mean {fcal['settle_mean']}, max {fcal['settle_max']})")

    def ltiers(lpq):
        # the learned queries read BF16 rows (the form's storage); the fallback tier is the plain
        # search and reads its FP32 rows
        if lpq >= pq:
            return None
        return [song_amd.tier(lpq), song_amd.tier(pq, storage=song_amd.STORE_F32)]

    def make(n, lpq, scale, tiers, cap):
        return learned.LearnedBatch(ls, n, lpq, pq, k=k, conf_threshold=0.9, step_scale=scale,
                                    storage=song_amd.STORE_BF16, mode=song_amd.MODE_FP32, fallback_cap=cap,
                                    tiers=tiers)

    # sweep on the selection split, fallback uncapped (the headline form)
    lst = torch.empty((ns, 8), dtype=torch.int32, device=dev)
    lsweep, cands = [], []
    for lpq in [int(x) for x in args.learned_pqs.split(",")]:
        for sc in [float(x) for x in args.step_scales.split(",")]:
            scale = sc if sc > 0 else 1e6
            lb = make(ns, lpq, scale, ltiers(lpq), None)
            plan = lb.run(qs, stats=lst)
            torch.cuda.synchronize()
            rl = recall(lb.ids, gs_d, selq)
            ms = timed(lambda: lb.run(qs), 3, torch)
            stl = lst.cpu().numpy()
            fb = float((~plan[3]).float().mean().item())
            ent = {"pq": lpq, "step_scale": sc, "recall@10_selection": round(rl, 4),
                   "qps_selection": round(ns / (ms / 1e3), 1), "fallback_frac": round(fb, 4),
                   "mean_expansions": round(float(stl[:, 0].mean()), 1),
                   "mean_distances": round(float(stl[:, 1].mean()), 1)}
            lsweep.append(ent)
            log(f"learned+bf16 pq={lpq} scale={sc} selection recall={rl:.4f} {ms:.2f} ms "
                f"({ent['qps_selection']:.0f} q/s) fallback={fb:.3f} exp={ent['mean_expansions']}")
            if rl >= sel_target:
                cands.append((ent, lpq, scale))
            lb.batch.close()
    if not cands:
        return {"error": "no learned setting reached the target recall on the selection split", "sweep": lsweep}
    ent, lpq, scale = max(cands, key=lambda c: c[0]["qps_selection"])

    # autotune the tiers on the selection queries: the fallback tier must keep its results (it is
    # the plain search); the confident tier may take a Bloom filter if the selection recall stays
    # at the target
    tt = time.time()
    lb0 = make(ns, lpq, scale, ltiers(lpq), None)
    p = lb0.run(qs)
    torch.cuda.synchronize()
    base_ids = lb0.ids.clone()
    lb0.batch.close()
    gt_sel = gs_d

    def judge(tier_ids, rows):
        trial = base_ids.clone()
        trial[torch.from_numpy(rows.astype(np.int64)).to(dev)] = tier_ids[:, :k]
        return recall(trial, gt_sel, selq) >= sel_target

    approx = {0: judge} if ltiers(lpq) is not None and not args.no_bloom else None
    table = autotune.tune(ix, qs, p[0], torch.full((ns,), k, dtype=torch.int32, device=dev), tiers=ltiers(lpq),
                          mode=song_amd.MODE_FP32, storage=song_amd.STORE_BF16, reps=3, max_queries_per_tier=ns,
                          k_ld=k, entry=p[1], steps=p[2], freeze=("storage",), approx_ok=approx)
    log(f"learned tiers autotuned on the selection split in {time.time() - tt:.1f}s: {table['tiers']}")
    # the tuner times each tier alone; in the step the tiers run side by side and the step ends with
    # the fallback tier, which the learned tier's waves slow down.  The learned tier's grid is
    # therefore also chosen on the whole step (results do not depend on it)
    if ltiers(lpq) is not None:
        def step_ms(tiers):
            lb = make(ns, lpq, scale, tiers, None)
            lb.run(qs)
            ms = timed(lambda: lb.run(qs), 5, torch)
            lb.batch.close()
            return ms

        best_t, best_ms = table["tiers"], step_ms(table["tiers"])
        grid_log = [(table["tiers"][0]["groups_per_cu"], round(best_ms, 3))]
        for gpc in (1, 2, 4):
            if gpc == table["tiers"][0]["groups_per_cu"]:
                continue
            cand = [dict(t) for t in table["tiers"]]
            cand[0]["groups_per_cu"] = gpc
            ms = step_ms(cand)
            grid_log.append((gpc, round(ms, 3)))
            if ms < best_ms:
                best_t, best_ms = cand, ms
        table["tiers"] = best_t
        table["step_grid_trials"] = grid_log
        log(f"learned tier grid on the whole step (groups per CU, ms): {grid_log} -> "
            f"{best_t[0]['groups_per_cu']}")
        # the fallback tier's visited-table width and heap depth, also on the whole step: the tuner
        # times the tier alone, and beside the learned tier's waves the two choices rank
        # differently from run to run (results are the same either way)
        fb_log = []
        for key, choices in (("visited_bits", (0, 32)), ("slack", (32, 128))):
            for v in choices:
                if best_t[1].get(key) == v:
                    continue
                cand = [dict(t) for t in best_t]
                cand[1][key] = v
                try:
                    ms = step_ms(cand)
                except song_amd.SongError:
                    continue
                fb_log.append((key, v, round(ms, 3)))
                if ms < best_ms * 0.99:
                    best_t, best_ms = cand, ms
        table["tiers"] = best_t
        table["step_fallback_trials"] = fb_log
        log(f"fallback tier on the whole step (knob, value, ms): {fb_log} -> visited_bits "
            f"{best_t[1].get('visited_bits')}, slack {best_t[1].get('slack')}")

    def measure_test(cap):
        lb = make(nq, lpq, scale, table["tiers"], cap)
        st = torch.empty((nq, 8), dtype=torch.int32, device=dev)
        pe = lb.run(q, stats=st)
        torch.cuda.synchronize()
        eager_ids = lb.ids.clone()
        stl = st.cpu().numpy()
        plan_g = lb.capture(q)
        lb.replay()
        torch.cuda.synchronize()
        same = bool(torch.equal(lb.ids, eager_ids))
        fbm = ~plan_g[3]
        n_fb = int(fbm.sum().item())
        fb_capped = int(((lb.flags[fbm] & song_amd.FLAG_STEP_CAP) != 0).sum().item())
        fb_equal = int((lb.ids[fbm][:, :k] == plain_ids[fbm][:, :k]).all(1).sum().item())
        r_final = recall(lb.ids, gt_d, query)
        ms_l, wall_l = timed_steps(lb.replay, args.steps, args.warmup, torch, stream)
        # algorithmic bytes: confident queries read BF16 rows + scale, fallback queries FP32 rows
        conf = plan_g[3].cpu().numpy()
        alg = (alg_bytes(stl[conf], dim, dim * 2 + 4, int(conf.sum())) +
               alg_bytes(stl[~conf], dim, dim * 4, int((~conf).sum())))
        fb_st = stl[~conf]
        grid0 = lb.batch.grid(0)
        lb.batch.close()
        vf = None
        if cap is None and table["tiers"] and table["tiers"][0].get("bloom_k", 0) != 0:
            # the confident tier took a Bloom filter: the same setting with the exact visited table
            ex = [dict(t) for t in table["tiers"]]
            ex[0]["bloom_k"] = 0
            lbx = make(nq, lpq, scale, ex, cap)
            lbx.run(q)
            torch.cuda.synchronize()
            r_exact = recall(lbx.ids, gt_d, query)
            gx = lbx.batch.grid(0)
            lbx.batch.close()
            vf = {"learned_tier_bloom_k": table["tiers"][0]["bloom_k"], "lds_per_wave": grid0[2],
                  "exact_table": {"lds_per_wave": gx[2], "recall@10": round(r_exact, 4)},
                  "recall@10_delta": round(r_final - r_exact, 4)}
        return {"qps": round(nq / (ms_l / 1e3), 1), "ms_per_step": round(ms_l, 4), "wall_ms_per_step": round(wall_l, 4),
                "recall@10": round(r_final, 4), "graph_ids_equal_eager": same, "algorithmic_bytes_per_step": alg,
                "fallback": {"queries": n_fb, "cap": cap, "capped": fb_capped, "top10_ids_equal_plain": fb_equal,
                             "max_expansions": int(fb_st[:, 0].max()) if len(fb_st) else 0,
                             "mean_expansions": round(float(fb_st[:, 0].mean()), 1) if len(fb_st) else 0.0},
                "visited_filter": vf}

    head = measure_test(None)
    log(f"learned+bf16 batched pq={lpq} scale={ent['step_scale']} (uncapped fallback): {head['qps']:.0f} q/s "
        f"({head['qps'] / qps_plain:.2f}x plain), test recall {head['recall@10']:.4f}")
    capped = None
    if fcap:
        capped = measure_test(fcap)
        capped["calibration"] = fcal
        log(f"same setting, fallback capped at {fcap}: {capped['qps']:.0f} q/s, test recall {capped['recall@10']:.4f}, "
            f"{capped['fallback']['capped']} capped, {capped['fallback']['top10_ids_equal_plain']} of "
            f"{capped['fallback']['queries']} with the plain top-10")
    fit_out = {kk: vv for kk, vv in fit.items() if kk != "loss"}
    fit_out["final_loss"] = fit["loss"][-1]
    info = dict(head, pq=lpq, step_scale=ent["step_scale"], chosen_on="selection split", selection=ent,
                tiers=table["tiers"], step_grid_trials=table.get("step_grid_trials"),
                step_fallback_trials=table.get("step_fallback_trials"), capped_fallback=capped,
                training=fit_out, sweep_selection=lsweep)
    return info


def bench_c4(args, torch, song_amd):
    """Batching form on its workload: 2M x 96 unit-norm Gaussian rows, inner product,
    100,000 queries with k from {1, 10, 100} and pq from {32, 64, 128, 256, 512} (k capped at
    pq), FP32 accumulation.  Timed: the batched search (autotuned tiers, graph replay)
    against the same queries searched with one plain call per distinct pq."""
    from song_amd import autotune, datasets, groundtruth, workloads

    dev = torch.device("cuda:0")
    t0 = time.time()
    base, query = datasets.normalized_ip(n_base=args.c4_n, n_query=args.c4_nq)
    graph, bst = workloads.graph_for(base, "ip", verbose=True)
    log(f"c4 workload ready in {time.time() - t0:.1f}s (graph cached={bst.get('cached')})")
    ix = song_amd.Index(base, graph, "ip")
    nq, dim = query.shape
    k_h, pq_h = datasets.hetero_budgets(nq)
    q = torch.from_numpy(query).to(dev)
    k_d, pq_d = torch.from_numpy(k_h).to(dev), torch.from_numpy(pq_h).to(dev)
    kmax = int(k_h.max())
    mode = song_amd.MODE_FP32
    pq_values = sorted(set(pq_h.tolist()))
    groups = [(p, np.nonzero(pq_h == p)[0]) for p in pq_values]

    # plain: one call per distinct pq over its queries (what the plain API needs)
    pids = torch.full((nq, kmax), -1, dtype=torch.int32, device=dev)
    pfl = torch.zeros(nq, dtype=torch.int32, device=dev)
    gq = [(p, torch.from_numpy(idx).to(dev), q[torch.from_numpy(idx).to(dev)].contiguous()) for p, idx in groups]
    outs = [(torch.empty((len(i), min(p, kmax)), dtype=torch.int32, device=dev),
             torch.empty(len(i), dtype=torch.int32, device=dev)) for p, i, _ in gq]

    def plain_all():
        # retry=1: overflowed queries are re-run with larger tables (complete results, as the batch)
        for (p, _, qq), (o, f) in zip(gq, outs):
            ix.search_device(qq, p, min(p, kmax), o, None, f, mode=mode, retry=1)

    plain_all()
    torch.cuda.synchronize()
    for (p, i, _), (o, f) in zip(gq, outs):
        pids[i, :o.shape[1]] = o
        pfl[i] = f
    ms_plain, _ = timed_steps(plain_all, args.steps, args.warmup, torch)

    # batched, default tiers then autotuned tiers
    stats = torch.empty((nq, 8), dtype=torch.int32, device=dev)
    bids = torch.full((nq, kmax), -1, dtype=torch.int32, device=dev)
    bfl = torch.empty(nq, dtype=torch.int32, device=dev)
    side = torch.cuda.Stream()
    results = {}
    default = autotune.default_tiers(pq_h)
    tt = time.time()
    table = autotune.tune(ix, q, pq_d, k_d, mode=mode, reps=3, max_queries_per_tier=20000, k_ld=kmax)
    log(f"autotuned in {time.time() - tt:.1f}s: {table['tiers']}")
    col = torch.arange(kmax, device=dev)[None, :] < k_d[:, None]
    for name, tiers in (("default_tiers", default), ("autotuned_tiers", table["tiers"])):
        b = song_amd.Batch(ix, tiers, nq, mode=mode)
        b.run(q, bids, bfl, pq_arr=pq_d, k_arr=k_d, stats=stats)
        torch.cuda.synchronize()
        same = bool(torch.equal(torch.where(col, bids, -1), torch.where(col, pids, -1))) and bool(torch.equal(bfl, pfl))
        st = stats.cpu().numpy()
        side.wait_stream(torch.cuda.current_stream())
        ms, wall = timed_steps(lambda: b.run(q, bids, bfl, pq_arr=pq_d, k_arr=k_d, graph=True, stream=side.cuda_stream),
                               args.steps, args.warmup, torch, side)
        cnt = b.counts().tolist()
        results[name] = {"qps": round(nq / (ms / 1e3), 1), "ms_per_step": round(ms, 4),
                         "wall_ms_per_step": round(wall, 4), "ids_equal_plain": same, "tiers": tiers,
                         "per_tier_queries": cnt[:-1], "re_runs": cnt[-1],
                         "mean_distances": round(float(st[:, 1].mean()), 1),
                         "algorithmic_bytes": alg_bytes(st, dim, dim * 4, nq)}
        log(f"{name}: {results[name]['qps']:.0f} q/s, ids equal plain: {same}")
        b.close()

    # recall per k group on a sample, against the float64 brute force
    ns = min(nq, args.c4_recall_sample)
    _, gt_d = groundtruth.knn_fp64(base, query[:ns], kmax, "ip")
    rec = {}
    ph = pids.cpu().numpy().astype(np.int64)
    for kk in sorted(set(k_h.tolist())):
        sel = np.nonzero(k_h[:ns] == kk)[0]
        rec[f"recall@{kk}"] = round(groundtruth.recall_at_k(ph[:ns][sel], gt_d[sel], kk, base, query[:ns][sel], "ip"),
                                    4)

    cpu = None
    if not args.no_cpu:
        # FP32-mode GPU ids on float data are not compared with the float64 oracle (the
        # exact-mode parity is covered by the tests); the sample times the same searches
        cpu = cpu_baseline(base, graph, "ip", query, groups, args.cpu_seconds)
    # both tier tables are measured runs of the same batching form; the autotuner tunes each tier
    # alone, so the tiers running side by side can come out either way: report the faster
    best_name = max(results, key=lambda nm: results[nm]["qps"])
    best = results[best_name]
    ach = best["algorithmic_bytes"] / (best["ms_per_step"] / 1e3)
    return {
        "metric": "QPS, batching form: 2M x96 IP, 100k queries with k in {1,10,100} and pq in {32..512}, 1 GPU",
        "value": best["qps"], "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": best["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic: seeded unit-norm Gaussian rows, graph = reference insertion build",
        "config": {"workload": f"{args.c4_n:,} x {dim} inner product, {nq:,} queries, k uniform in {{1,10,100}}, "
                               f"pq uniform in {{32,64,128,256,512}}, k capped at pq",
                   "plain_per_pq_calls": {"qps": round(nq / (ms_plain / 1e3), 1), "ms_per_step": round(ms_plain, 4)},
                   "batched": results, "value_from": best_name, "recall_sample": ns, **rec},
        "roofline": {"bound": "hbm", "achieved": round(ach / 1e9, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK, 5), "traffic": None,
                     "kernel": "search_persistent<IP, fp32, F32 rows> tiers + bucket/retry (one batched run)"},
        "cpu_baseline": cpu,
    }


def bench_c1(args, torch, song_amd):
    """Small-row workload: 15,000 x 16 uniform rows, cosine, 5,000 queries, pq 100, top-10, exact
    mode (ids equal the reference CPU search).  Measures what several queries per wavefront buy on
    16-d rows (four 16-byte units, so one query leaves most of a 64-lane wave idle): the batched
    search autotuned with one query per wave, then with 2 or 4 queries per wave allowed, against the
    plain one-launch search."""
    from song_amd import autotune, datasets, groundtruth, workloads

    dev = torch.device("cuda:0")
    base, query, graph, metric, _ = workloads.load("c1_15k")
    if args.c1_repeat > 1:
        # the 5,000 queries repeated: the same per-query work at a batch that no longer fits the
        # device at once (throughput-bound instead of latency-bound)
        query = np.ascontiguousarray(np.tile(query, (args.c1_repeat, 1)))
    ix = song_amd.Index(base, graph, metric)
    nq, pq, k = query.shape[0], args.c1_pq, args.k
    q = torch.from_numpy(query).to(dev)
    mode = song_amd.MODE_EXACT
    ids = torch.empty((nq, k), dtype=torch.int32, device=dev)
    fl = torch.empty(nq, dtype=torch.int32, device=dev)
    ix.search_device(q, pq, k, ids, None, fl, mode=mode)
    torch.cuda.synchronize()
    plain_ids = ids.clone()
    ms_plain, _ = timed_steps(lambda: ix.search_device(q, pq, k, ids, None, fl, mode=mode, retry=0), args.steps,
                              args.warmup, torch)
    _, gt_d = groundtruth.knn_fp64(base, query, k, metric)
    r = groundtruth.recall_at_k(plain_ids.cpu().numpy().astype(np.int64), gt_d, k, base, query, metric)
    pq_d = torch.full((nq,), pq, dtype=torch.int32, device=dev)
    k_d = torch.full((nq,), k, dtype=torch.int32, device=dev)
    res = {}
    side = torch.cuda.Stream()
    runs = [("one_query_per_wave", ("queries_per_wave",), 1), ("autotuned_queries_per_wave", (), None),
            ("two_queries_per_wave", ("queries_per_wave",), 2), ("four_queries_per_wave", ("queries_per_wave",), 4)]
    for name, freeze, qpw in runs:
        start = [song_amd.tier(pq, queries_per_wave=qpw)] if qpw else None
        table = autotune.tune(ix, q, pq_d, k_d, tiers=start, mode=mode, reps=3, max_queries_per_tier=nq, k_ld=k,
                              freeze=freeze)
        b = song_amd.Batch(ix, table["tiers"], nq, mode=mode)
        bids = torch.empty_like(ids)
        bfl = torch.empty_like(fl)
        side.wait_stream(torch.cuda.current_stream())
        b.run(q, bids, bfl, pq=pq, k=k, graph=True, stream=side.cuda_stream)
        torch.cuda.synchronize()
        same = bool(torch.equal(bids, plain_ids))
        ms, _ = timed_steps(lambda: b.run(q, bids, bfl, pq=pq, k=k, graph=True, stream=side.cuda_stream), args.steps,
                            args.warmup, torch, side)
        res[name] = {"qps": round(nq / (ms / 1e3), 1), "ms_per_step": round(ms, 4), "ids_equal_plain": same,
                     "tiers": table["tiers"], "grid": b.grid(0)}
        log(f"c1 batched, {name}: {res[name]['qps']:.0f} q/s, ids equal plain: {same}, {table['tiers']}")
        b.close()
    cpu = None
    if not args.no_cpu:
        pids = plain_ids.cpu().numpy().astype(np.int64)
        cpu = cpu_baseline(base, graph, metric, query, [(pq, np.arange(nq))], args.cpu_seconds,
                           check=lambda p, take, cids: bool(np.array_equal(cids[:, :k].astype(np.int64), pids[take])))
    best = max(res.values(), key=lambda r: r["qps"])
    return {
        "metric": "QPS, 15k x16 cosine, pq 100, top-10, exact ids, 1 GPU",
        "value": best["qps"], "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": best["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 products, fp64 sums (exact)", "data": "synthetic: seeded uniform rows, graph = reference insertion build",
        "config": {"workload": (f"15,000 x 16 uniform [-1,1] cosine, {nq:,} queries"
                                + (f" (5,000 repeated {args.c1_repeat}x)" if args.c1_repeat > 1 else "")
                                + f", pq {pq}, top-{k}"),
                   "plain": {"qps": round(nq / (ms_plain / 1e3), 1), "ms_per_step": round(ms_plain, 4),
                             "recall@10": round(r, 4)},
                   "batched": res},
        "roofline": None, "cpu_baseline": cpu,
    }


def bench_c3(args, torch, song_amd):
    """Precision form on its workload: rank-64 Gaussian factor model + N(0, 0.05) noise, min-max
    scaled to [0,1], dim 960, L2, 1,000 queries.  Rows searched as FP32 (float64 accumulation =
    the exact form, and float32 accumulation), BF16 and FP8-E4M3 with per-row scales (FP32
    accumulation; exact FP32 re-rank of the final pq list, and without it to expose the raw
    low-precision distance error).  Reports recall@10 / recall@100 against the float64 brute force,
    the recall deltas against FP32 and FP64, the distance errors and the QPS of each."""
    from song_amd import datasets, groundtruth, workloads

    dev = torch.device("cuda:0")
    t0 = time.time()
    base, query = datasets.factor_l2(n_base=args.c3_n, n_query=args.c3_nq)
    graph, bst = workloads.graph_for(base, "l2", verbose=True)
    log(f"c3 workload ready in {time.time() - t0:.1f}s (graph cached={bst.get('cached')})")
    ix = song_amd.Index(base, graph, "l2")
    ix.add_storage(song_amd.STORE_BF16)
    ix.add_storage(song_amd.STORE_FP8)
    nq, dim = query.shape
    _, gt_d = groundtruth.knn_fp64(base, query, 100, "l2")
    q = torch.from_numpy(query).to(dev)
    pq = args.c3_pq
    ids = torch.empty((nq, 100), dtype=torch.int32, device=dev)
    ds = torch.empty((nq, 100), dtype=torch.float32, device=dev)
    fl = torch.empty(nq, dtype=torch.int32, device=dev)
    forms = [("fp64_accum", song_amd.MODE_EXACT, song_amd.STORE_F32, 1),
             ("fp32", song_amd.MODE_FP32, song_amd.STORE_F32, 1),
             ("bf16_rerank", song_amd.MODE_FP32, song_amd.STORE_BF16, 1),
             ("fp8_rerank", song_amd.MODE_FP32, song_amd.STORE_FP8, 1),
             ("bf16_raw", song_amd.MODE_FP32, song_amd.STORE_BF16, 0),
             ("fp8_raw", song_amd.MODE_FP32, song_amd.STORE_FP8, 0)]
    res = {}
    for name, mode, store, rerank in forms:
        def run():
            ix.search_device(q, pq, 100, ids, ds, fl, mode=mode, storage=store, rerank=rerank)
        run()
        torch.cuda.synchronize()
        i_h = ids.cpu().numpy().astype(np.int64)
        d_h = ds.cpu().numpy().astype(np.float64)
        exact = groundtruth.exact_distances(base, query, i_h, "l2")
        ok = i_h >= 0
        rel = np.abs(d_h - exact)[ok] / np.maximum(np.abs(exact[ok]), 1e-30)
        ms = timed(run, 5, torch)
        res[name] = {"recall@10": round(groundtruth.recall_at_k(i_h, gt_d, 10, base, query, "l2"), 4),
                     "recall@100": round(groundtruth.recall_at_k(i_h, gt_d, 100, base, query, "l2"), 4),
                     "max_rel_dist_err_vs_fp64": float(f"{rel.max():.3e}"),
                     "mean_rel_dist_err_vs_fp64": float(f"{rel.mean():.3e}"),
                     "qps": round(nq / (ms / 1e3), 1), "ms": round(ms, 4)}
        log(f"c3 {name}: {res[name]}")
    for name in ("bf16_rerank", "fp8_rerank", "bf16_raw", "fp8_raw"):
        for ref in ("fp32", "fp64_accum"):
            for kk in (10, 100):
                res[name][f"recall@{kk}_delta_vs_{ref}"] = round(res[name][f"recall@{kk}"] - res[ref][f"recall@{kk}"], 4)
    best = res["bf16_rerank"]
    return {
        "metric": "precision form: recall delta and distance error, factor-model 960-d L2, BF16 / FP8-E4M3 rows",
        "value": best["qps"], "unit": "queries/s", "n_gpus": 1, "steps": 5, "warmup": 1, "ms_per_step": best["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic: seeded rank-64 Gaussian factor model, graph = reference insertion build",
        "config": {"workload": f"{args.c3_n:,} x {dim} L2 in [0,1], {nq:,} queries, pq {pq}, k 10 and 100",
                   "forms": res},
        "roofline": None, "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("c1", "c2", "c3", "c4", "c5"), default="c2")
    ap.add_argument("--target-recall", type=float, default=0.95)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--pqs", default="256,384,512,640,768,896,1024")
    ap.add_argument("--selection-margin", type=float, default=0.004,
                    help="settings are chosen where the selection split's recall reaches the target plus this")
    ap.add_argument("--pq-step", type=int, default=16,
                    help="plain form: granularity of the bisection after the coarse grid")
    ap.add_argument("--sel-queries", type=int, default=0,
                    help="held-out selection queries (same mixture, own stream) on which settings are chosen; "
                         "0 = as many as test queries, so that tiers are tuned at the batch size they run at")
    ap.add_argument("--pq", type=int, default=0, help="plain form: skip the sweep and time this pq")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="no learned form")
    ap.add_argument("--no-batch", action="store_true", help="no plain batched form")
    ap.add_argument("--train-queries", type=int, default=200_000)
    ap.add_argument("--learned-pqs", default="64,96,128,192,256")
    ap.add_argument("--step-scales", default="1.5,2.5,0")
    ap.add_argument("--fallback-n", type=int, default=442,
                    help="C2: queries of the fallback-tier-alone measurement when the learned form does not run "
                         "(0 = skip it)")
    ap.add_argument("--no-bloom", action="store_true", help="learned form: exact visited tables only")
    ap.add_argument("--fallback-quantile", type=float, default=0.999,
                    help="capped variant beside the headline: fallback expansion cap = this quantile of the plain "
                         "search's top-k settle step over low-confidence training queries (<= 0: not measured)")
    ap.add_argument("--no-recall", action="store_true",
                    help="with --pq: skip the ground truth and recall (profiling runs)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run one device copy of this many MiB first (a known-size kernel that "
                         "calibrates rocprof FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--c4-n", type=int, default=2_000_000)
    ap.add_argument("--c4-nq", type=int, default=100_000)
    ap.add_argument("--c4-recall-sample", type=int, default=10_000)
    ap.add_argument("--c1-pq", type=int, default=100)
    ap.add_argument("--c1-repeat", type=int, default=1,
                    help="C1: repeat the 5,000 queries this many times (a throughput-bound batch)")
    ap.add_argument("--c5-n", type=int, default=10_000_000)
    ap.add_argument("--c5-nq", type=int, default=100_000)
    ap.add_argument("--c3-n", type=int, default=1_000_000)
    ap.add_argument("--c3-nq", type=int, default=1000)
    ap.add_argument("--c3-pq", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the search runs on one GPU; --gpus must be 1 (more GPUs would each search "
                                   "their own share of the queries against a replica of the index, with no "
                                   "communication)"}))
        sys.exit(2)
    if args.no_recall and not args.pq:
        ap.error("--no-recall needs --pq")

    import torch

    import song_amd

    out = {"c1": bench_c1, "c2": bench_c2, "c3": bench_c3, "c4": bench_c4,
           "c5": bench_c2}[args.workload](args, torch, song_amd)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
