#!/usr/bin/env python
"""Benchmark: QPS at recall@10 >= 0.95 on the 1M x 128 L2 workload, one GPU.

Workload: 1,000,000 base vectors, dim 128, integer-valued mixture of 1,000
Gaussians (seed 2), 10,000 queries from the same mixture, L2, degree-31 graph
from the reference insertion build (byte-identical, built here by the parallel
builder), top-10.  A "step" is one search of the whole 10,000-query batch with
queries, index and outputs resident in HBM.

The pq sweep picks the smallest pq whose recall@10 (tie-aware, against a
float64 brute force on the GPU) reaches the target; that pq is then timed.
Prints one JSON line (see the keys below).
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--target-recall", type=float, default=0.95)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--pqs", default="64,96,128,160,192,224,256,320,384,448,512,640,768")
    ap.add_argument("--pq", type=int, default=0, help="skip the sweep and time this pq")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-recall", action="store_true",
                    help="with --pq: skip the ground truth and recall (profiling runs)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run one device copy of this many MiB first (a known-size kernel that "
                         "calibrates rocprof FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.no_recall and not args.pq:
        ap.error("--no-recall needs --pq")
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (queries shard across GPUs "
                                   "with no communication, see DESIGN.md)"}))
        sys.exit(2)

    import torch

    import song_amd
    from song_amd import groundtruth, workloads

    dev = torch.device("cuda:0")
    t0 = time.time()
    base, query, graph, metric, bst = workloads.load("c2_1m", verbose=True)
    log(f"workload ready in {time.time() - t0:.1f}s (graph cached={bst.get('cached')})")
    ix = song_amd.Index(base, graph, metric)
    nq, dim = query.shape
    k = args.k

    if args.copy_mb:
        src = torch.ones(args.copy_mb << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {src.numel() * 4} bytes read, {dst.numel() * 4} bytes written")
        del src, dst

    if not args.no_recall:
        t1 = time.time()
        gt_i, gt_d = groundtruth.knn_fp64(base, query, k, metric)
        log(f"fp64 ground truth in {time.time() - t1:.1f}s")

    q = torch.from_numpy(query).to(dev)
    stream = torch.cuda.current_stream()

    def run(pq, ids, stats=None, retry=1):
        ix.search_device(q, pq, k, ids, None, flags, stats, retry=retry)

    flags = torch.empty(nq, dtype=torch.int32, device=dev)
    stats = torch.empty((nq, 8), dtype=torch.int32, device=dev)
    ids = torch.empty((nq, k), dtype=torch.int32, device=dev)

    sweep = []
    chosen = None
    pqs = [args.pq] if args.pq else [int(x) for x in args.pqs.split(",")]
    for pq in pqs:
        run(pq, ids, stats)
        torch.cuda.synchronize()
        r = (float("nan") if args.no_recall else
             groundtruth.recall_at_k(ids.cpu().numpy().astype(np.int64), gt_d, k, base, query, metric))
        st = stats.cpu().numpy()
        times = []
        for _ in range(3):
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            run(pq, ids, retry=0)
            ev1.record(stream)
            torch.cuda.synchronize()
            times.append(ev0.elapsed_time(ev1))
        ms = float(np.median(times))
        sweep.append({"pq": pq, "recall@10": round(r, 4), "qps": round(nq / (ms / 1e3), 1),
                      "mean_expansions": round(float(st[:, 0].mean()), 1),
                      "mean_distances": round(float(st[:, 1].mean()), 1),
                      "prefetch_hit": round(float(st[:, 3].sum() / max(1, st[:, 0].sum())), 4)})
        log(f"pq={pq} recall@{k}={r:.4f} median-of-3 {ms:.2f} ms ({nq / (ms / 1e3):.0f} q/s) "
            f"exp={st[:, 0].mean():.1f} dist={st[:, 1].mean():.1f} pf={st[:, 3].sum() / max(1, st[:, 0].sum()):.3f}")
        if chosen is None and r >= args.target_recall:
            chosen = (pq, r, st)
            if not args.pq:
                break
    if chosen is None:
        pq, r, st = pqs[-1], sweep[-1]["recall@10"], stats.cpu().numpy()
        log(f"target recall not reached; timing pq={pq}")
    else:
        pq, r, st = chosen

    # timed region: W untimed steps, then K steps between synchronizations
    fl_host = flags.cpu().numpy()
    overflow = int((fl_host & 3).sum() > 0)
    retry = 1 if overflow else 0
    for _ in range(args.warmup):
        run(pq, ids, retry=retry)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_start = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        run(pq, ids, retry=retry)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_total = ev0.elapsed_time(ev1)
    ms_step = ms_total / args.steps
    qps = nq * args.steps / (ms_total / 1e3)

    # algorithmic bytes of one launch: every distance reads one fp32 row, every expansion
    # reads (1 + degree) u32 adjacency words, every query reads its own vector once
    st64 = st.astype(np.int64)
    alg_bytes = int(st64[:, 1].sum() * dim * 4 + st64[:, 2].sum() * 4 + nq * dim * 4)
    achieved = alg_bytes / (ms_step / 1e3)

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle

        threads = min(16, os.cpu_count() or 1)
        # size the sample from a short single-query probe to about --cpu-seconds of work
        tp = time.perf_counter()
        pyoracle.search_batch(base, graph, metric, query[:50], pq, threads=1)
        per_q = (time.perf_counter() - tp) / 50
        sample = int(min(nq, max(threads * 20, args.cpu_seconds * threads / max(per_q, 1e-6))))
        tc = time.perf_counter()
        cids, _ = pyoracle.search_batch(base, graph, metric, query[:sample], pq, threads=threads)
        tcpu = time.perf_counter() - tc
        same = bool(np.array_equal(cids[:, :k].astype(np.int64), ids[:sample].cpu().numpy().astype(np.int64)))
        cpu = {"value": round(sample / tcpu, 1), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"first {sample} of the {nq} queries at pq={pq} (oracle restatement of graph.h:169-216, "
                         f"{threads} threads, static split); single-thread probe {1.0 / per_q:.1f} q/s",
               "cpu_model": cpu_model(), "ids_equal_gpu": same}
        log(f"cpu baseline: {cpu['value']} q/s on {threads} threads ({sample} queries)")

    out = {
        "metric": "QPS @ recall@10>=0.95, 1M x128 L2, 1 GPU",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded Gaussian-mixture generator (integer-valued), graph = reference insertion build",
        "config": {"workload": "1,000,000 x 128 L2 mixture (1,000 centres, sigma 20, clipped/rounded), "
                               "10,000 queries, top-10",
                   "form": "plain (exact ids; fp32 accumulation is exact on integer data)",
                   "pq": pq, "k": k, "recall@10": round(r, 4), "target_recall": args.target_recall,
                   "mean_expansions": round(float(st[:, 0].mean()), 1),
                   "mean_distances": round(float(st[:, 1].mean()), 1),
                   "wall_s_timed": round(wall, 4), "overflow_retry": bool(retry)},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 5), "traffic": None,
                     "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
        "sweep": sweep,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
