"""Write golden fixtures from the reference's own CPU binary.

Runs only where the reference sources are present: ``make -C oracle ref``
compiles /root/reference/main.cc into oracle/_ref/song.cpu (the reference
Makefile's song.cpu line).  For each seeded case this script

  1. writes the base and query vectors as one-based LIBSVM text,
  2. runs ``song.cpu build <base> 0 0 <n> <dim> 0 <metric>`` (build_graph.sh:9),
  3. runs ``song.cpu test 0 <query> <pq> <n> <dim> <pq> <metric>`` (test_query.sh:17),
  4. stores the SHA-256 of bfsg.graph / bfsg.data and the printed id lists.

Outputs (small, committed): <case>.json (metadata + hashes) and
<case>_ids.npy ([nq, pq], uint16 when n < 65535 else uint32, all-ones
where fewer ids were printed).
Tiny cases also store the whole graph as <case>_graph.npy.

Usage: python tests/golden/make_golden.py [case ...]
       python tests/golden/make_golden.py c2_1m <dir with song.cpu's 1M bfsg.data/bfsg.graph>
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from song_amd import datasets  # noqa: E402

SONG_CPU = os.path.join(ROOT, "oracle", "_ref", "song.cpu")


def tie_grid(n_base, n_query, dim, levels, seed):
    rng = np.random.default_rng(seed)
    base = rng.integers(0, levels, size=(n_base, dim)).astype(np.float32)
    query = rng.integers(0, levels, size=(n_query, dim)).astype(np.float32)
    return base, query


# name: (generator, kwargs, metric, pq, n_query_kept, keep_graph)
CASES = {
    "c1_cos_15k": (datasets.uniform_cos, dict(n_base=15000, n_query=5000, dim=16, seed=1), "cos", 100, 400, False),
    "mix_l2_20k": (datasets.mixture_l2, dict(n_base=20000, n_query=300, dim=128, seed=2), "l2", 100, 300, False),
    "ip_10k": (datasets.normalized_ip, dict(n_base=10000, n_query=300, dim=96, seed=4), "ip", 64, 300, False),
    "tie_l2_3k": (tie_grid, dict(n_base=3000, n_query=300, dim=4, levels=4, seed=7), "l2", 50, 300, True),
    "tie_ip_2k": (tie_grid, dict(n_base=2000, n_query=200, dim=6, levels=3, seed=8), "ip", 40, 200, True),
    "tiny_cos_500": (datasets.uniform_cos, dict(n_base=500, n_query=100, dim=8, seed=9), "cos", 20, 100, True),
    "tiny_l2_500": (datasets.mixture_l2, dict(n_base=500, n_query=100, dim=12, n_centers=10, seed=10), "l2", 30, 100, True),
}


def sha256_file(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for blk in iter(lambda: f.read(1 << 20), b""):
            h.update(blk)
    return h.hexdigest()


def run_case(name):
    gen, kw, metric, pq, nkeep, keep_graph = CASES[name]
    base, query = gen(**kw)
    query = query[:nkeep]
    n, dim = base.shape
    with tempfile.TemporaryDirectory() as d:
        datasets.write_libsvm(os.path.join(d, "base.svm"), base)
        datasets.write_libsvm(os.path.join(d, "query.svm"), query)
        subprocess.run([SONG_CPU, "build", "base.svm", "0", "0", str(n), str(dim), "0", metric],
                       cwd=d, check=True, capture_output=True)
        out = subprocess.run([SONG_CPU, "test", "0", "query.svm", str(pq), str(n), str(dim), str(pq), metric],
                             cwd=d, check=True, capture_output=True, text=True).stdout
        lines = out.split("\n")[: len(query)]
        ids = np.full((len(query), pq), np.iinfo(np.uint32).max, dtype=np.uint32)
        for i, line in enumerate(lines):
            v = [int(t) for t in line.split()]
            ids[i, : len(v)] = v
        meta = {
            "case": name,
            "generator": gen.__name__,
            "kwargs": kw,
            "metric": metric,
            "pq": pq,
            "n_query": len(query),
            "graph_sha256": sha256_file(os.path.join(d, "bfsg.graph")),
            "data_sha256": sha256_file(os.path.join(d, "bfsg.data")),
            "source": "reference song.cpu (main.cc, g++ -std=c++11 -O3 -march=native)",
        }
        small = np.uint16 if n < 65535 else np.uint32
        ids_small = np.where(ids == np.iinfo(np.uint32).max, np.iinfo(small).max, ids).astype(small)
        np.save(os.path.join(HERE, f"{name}_ids.npy"), ids_small)
        if keep_graph:
            g = np.fromfile(os.path.join(d, "bfsg.graph"), dtype=np.uint64).reshape(n, 32)
            np.save(os.path.join(HERE, f"{name}_graph.npy"), g.astype(small))
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(name, meta["graph_sha256"][:16], flush=True)


def run_c2_1m(prebuilt_dir, n_query=200, pqs=(100, 200)):
    """The 1M x 128 mixture workload (seed 2).  Building it with song.cpu takes ~13 min,
    so this reuses a directory where ``song.cpu build`` already wrote bfsg.data/bfsg.graph
    from the same generator (checked through the data hash)."""
    base, query = datasets.mixture_l2()
    query = query[:n_query]
    n, dim = base.shape
    data_sha = sha256_file(os.path.join(prebuilt_dir, "bfsg.data"))
    assert data_sha == __import__("hashlib").sha256(base.tobytes()).hexdigest(), "prebuilt data differs"
    meta = {"case": "c2_1m", "generator": "mixture_l2", "kwargs": {}, "metric": "l2", "n_query": n_query,
            "pqs": list(pqs), "graph_sha256": sha256_file(os.path.join(prebuilt_dir, "bfsg.graph")),
            "data_sha256": data_sha,
            "source": "reference song.cpu (main.cc, g++ -std=c++11 -O3 -march=native)"}
    with tempfile.TemporaryDirectory() as d:
        for f in ("bfsg.data", "bfsg.graph"):
            os.symlink(os.path.join(os.path.abspath(prebuilt_dir), f), os.path.join(d, f))
        datasets.write_libsvm(os.path.join(d, "query.svm"), query)
        for pq in pqs:
            out = subprocess.run([SONG_CPU, "test", "0", "query.svm", str(pq), str(n), str(dim), str(pq), "l2"],
                                 cwd=d, check=True, capture_output=True, text=True).stdout
            ids = np.array([[int(t) for t in line.split()] for line in out.split("\n")[:n_query]], dtype=np.uint32)
            np.save(os.path.join(HERE, f"c2_1m_pq{pq}_ids.npy"), ids)
    with open(os.path.join(HERE, "c2_1m.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("c2_1m", meta["graph_sha256"][:16], flush=True)


if __name__ == "__main__":
    if not os.path.exists(SONG_CPU):
        sys.exit("build the reference binary first: make -C oracle ref")
    if sys.argv[1:2] == ["c2_1m"]:
        run_c2_1m(sys.argv[2])
        sys.exit(0)
    for c in sys.argv[1:] or list(CASES):
        run_case(c)
