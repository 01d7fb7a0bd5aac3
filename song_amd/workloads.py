"""Seeded workloads with their graphs, built once per machine and cached.

A graph is the reference insertion build of the base vectors (song_graph_build,
byte-identical to `song build`).  Built graphs are cached under
``<repo>/_cache`` keyed by the SHA-256 of the base vectors, so a test session
and a benchmark on the same host share one build.
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

from . import build_graph, datasets

CACHE_DIR = os.environ.get("SONG_CACHE_DIR",
                           os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "_cache"))

WORKLOADS = {
    # name: (generator, kwargs, metric)
    "c2_1m": (datasets.mixture_l2, {}, "l2"),
    "c1_15k": (datasets.uniform_cos, {}, "cos"),
}


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def default_threads() -> int:
    """Worker threads for host-side work: SONG_THREADS, else min(16, cores) (one GPU's CPU share)."""
    env = os.environ.get("SONG_THREADS")
    return int(env) if env else min(16, os.cpu_count() or 1)


def graph_for(base: np.ndarray, metric: str, threads: int = 0, verbose: bool = False):
    """Return the bfsg.graph image (uint64 [n, 32]) of `base`, from the cache if present."""
    threads = threads or default_threads()
    key = sha256(base)[:20]
    path = os.path.join(CACHE_DIR, f"graph_{metric}_{base.shape[0]}x{base.shape[1]}_{key}.u64")
    n = base.shape[0]
    if os.path.exists(path) and os.path.getsize(path) == n * 32 * 8:
        return np.fromfile(path, dtype=np.uint64).reshape(n, 32), {"cached": True}
    t = time.time()
    if verbose:
        os.environ["SONG_BUILD_PROGRESS"] = "1"  # the builder reports on stderr every ~20 s
    g, st = build_graph(base, metric, threads=threads, return_stats=True)
    st["cached"] = False
    if verbose:
        print(f"[graph] built {n}x{base.shape[1]} {metric} in {time.time() - t:.1f}s ({st})", file=sys.stderr,
              flush=True)
    os.makedirs(CACHE_DIR, exist_ok=True)
    tmp = path + f".tmp{os.getpid()}"
    g.tofile(tmp)
    os.replace(tmp, path)
    return g, st


def load(name: str, threads: int = 0, verbose: bool = False):
    """(base, queries, graph, metric, build_stats) for a named workload."""
    gen, kw, metric = WORKLOADS[name]
    base, query = gen(**kw)
    graph, st = graph_for(base, metric, threads=threads, verbose=verbose)
    return base, query, graph, metric, st
