"""Loader for the golden fixtures written by tests/golden/make_golden.py.

A fixture is data only: the generator name + seed + sizes (to regenerate the
vectors), the SHA-256 of the reference-built bfsg.graph / bfsg.data, and the
id lists the reference song.cpu printed.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

from song_amd import datasets

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def tie_grid(n_base, n_query, dim, levels, seed):
    rng = np.random.default_rng(seed)
    base = rng.integers(0, levels, size=(n_base, dim)).astype(np.float32)
    query = rng.integers(0, levels, size=(n_query, dim)).astype(np.float32)
    return base, query


GENERATORS = {
    "uniform_cos": datasets.uniform_cos,
    "mixture_l2": datasets.mixture_l2,
    "normalized_ip": datasets.normalized_ip,
    "tie_grid": tie_grid,
}


def names():
    """Cases whose vectors, graph and ids fit a quick test (the 1M case has its own tests)."""
    return sorted(f[:-5] for f in os.listdir(GOLDEN)
                  if f.endswith(".json") and not f.startswith(("c2_1m", "auto_tune_bloom")))


def load_c2_1m():
    with open(os.path.join(GOLDEN, "c2_1m.json")) as f:
        meta = json.load(f)
    ids = {pq: np.load(os.path.join(GOLDEN, f"c2_1m_pq{pq}_ids.npy")).astype(np.int64) for pq in meta["pqs"]}
    return meta, ids


def load(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        meta = json.load(f)
    base, query = GENERATORS[meta["generator"]](**meta["kwargs"])
    query = query[: meta["n_query"]]
    ids = np.load(os.path.join(GOLDEN, name + "_ids.npy")).astype(np.int64)
    pad = np.iinfo(np.load(os.path.join(GOLDEN, name + "_ids.npy"), mmap_mode="r").dtype).max
    ids[ids == pad] = -1
    gpath = os.path.join(GOLDEN, name + "_graph.npy")
    graph = np.load(gpath).astype(np.uint64) if os.path.exists(gpath) else None
    return meta, base, query, ids, graph


def sha256(arr: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()
