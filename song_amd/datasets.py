"""Seeded synthetic workloads for the graph search.

Each generator is deterministic for a given seed (numpy PCG64), so the same
vectors are produced in this container and on any other host.  Vectors are
returned as C-contiguous float32 arrays, the element type of ``bfsg.data``
(reference config.h:11 ``value_t = float``).

Workloads
---------
uniform_cos      15,000 x 16, U[-1,1], cosine               (seed 1)
mixture_l2       1,000,000 x 128 integer-valued Gaussian mixture, L2 (seed 2)
factor_l2        1,000,000 x 960 rank-64 factor model in [0,1], L2 (seed 3)
normalized_ip    2,000,000 x 96 unit Gaussian, inner product  (seed 4)

Sizes are arguments, so tests can draw small instances of the same
distributions.
"""
from __future__ import annotations

import numpy as np

METRIC_L2 = 0
METRIC_IP = 1
METRIC_COS = 2
METRIC_NAMES = {"l2": METRIC_L2, "ip": METRIC_IP, "cos": METRIC_COS}


def uniform_cos(n_base: int = 15000, n_query: int = 5000, dim: int = 16, seed: int = 1):
    """Base and query vectors i.i.d. U[-1,1] (cosine workload)."""
    rng = np.random.default_rng(seed)
    base = rng.uniform(-1.0, 1.0, size=(n_base, dim)).astype(np.float32)
    query = rng.uniform(-1.0, 1.0, size=(n_query, dim)).astype(np.float32)
    return base, query


def _mixture_rows(rng, centers, n, sigma, chunk=1 << 17):
    dim = centers.shape[1]
    out = np.empty((n, dim), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        lab = rng.integers(0, centers.shape[0], size=e - s)
        z = rng.standard_normal(size=(e - s, dim), dtype=np.float32)
        x = centers[lab] + np.float32(sigma) * z
        np.clip(np.rint(x), 0.0, 255.0, out=x)
        x += np.float32(0.0)  # rint(-0.3) is -0.0; store +0.0 as a LIBSVM round trip does
        out[s:e] = x
    return out


def mixture_l2(n_base: int = 1_000_000, n_query: int = 10_000, dim: int = 128,
               n_centers: int = 1000, sigma: float = 20.0, seed: int = 2):
    """Mixture of isotropic Gaussians, centres U[0,255]^dim, clipped and rounded.

    Every coordinate is an integer in [0,255], so every L2 distance is an
    integer below dim*255^2 (< 2^24 for dim <= 258) and is exact in float32.
    """
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0.0, 255.0, size=(n_centers, dim)).astype(np.float32)
    base = _mixture_rows(rng, centers, n_base, sigma)
    query = _mixture_rows(rng, centers, n_query, sigma)
    return base, query


def mixture_l2_queries(n: int, dim: int = 128, n_centers: int = 1000, sigma: float = 20.0, seed: int = 2,
                       stream: int = 1):
    """More queries from the same mixture as mixture_l2(seed=seed) (same centres, an
    independent stream), e.g. training queries for the learned form."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0.0, 255.0, size=(n_centers, dim)).astype(np.float32)
    rq = np.random.default_rng([seed, 1000 + stream])
    return _mixture_rows(rq, centers, n, sigma)


def factor_l2(n_base: int = 1_000_000, n_query: int = 1000, dim: int = 960, rank: int = 64,
              noise: float = 0.05, seed: int = 3, chunk: int = 1 << 16):
    """X = Z W^T + N(0, noise^2), then one global min-max scaling to [0,1]."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(size=(dim, rank), dtype=np.float32)
    n = n_base + n_query
    x = np.empty((n, dim), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        z = rng.standard_normal(size=(e - s, rank), dtype=np.float32)
        x[s:e] = z @ w.T + np.float32(noise) * rng.standard_normal(size=(e - s, dim), dtype=np.float32)
    lo, hi = float(x.min()), float(x.max())
    x -= np.float32(lo)
    x *= np.float32(1.0 / (hi - lo))
    return x[:n_base], x[n_base:].copy()


def normalized_ip(n_base: int = 2_000_000, n_query: int = 100_000, dim: int = 96, seed: int = 4,
                  chunk: int = 1 << 18):
    """Gaussian vectors scaled to unit L2 norm (inner-product workload)."""
    rng = np.random.default_rng(seed)

    def draw(n):
        out = np.empty((n, dim), dtype=np.float32)
        for s in range(0, n, chunk):
            e = min(n, s + chunk)
            z = rng.standard_normal(size=(e - s, dim), dtype=np.float32)
            z /= np.linalg.norm(z, axis=1, keepdims=True)
            out[s:e] = z
        return out

    base = draw(n_base)
    query = draw(n_query)
    return base, query


def hetero_budgets(n_query: int, seed: int = 4, ks=(1, 10, 100), pqs=(32, 64, 128, 256, 512)):
    """Per-query (k, pq_size) pairs drawn uniformly; k is capped at pq_size."""
    rng = np.random.default_rng(seed + 1000)
    k = rng.choice(np.asarray(ks, dtype=np.int32), size=n_query)
    pq = rng.choice(np.asarray(pqs, dtype=np.int32), size=n_query)
    return np.minimum(k, pq).astype(np.int32), pq.astype(np.int32)


def write_libsvm(path: str, x: np.ndarray, one_based: bool = True) -> None:
    """Write rows as LIBSVM text, one row per line, label 0, zeros omitted.

    Values are written with ``repr(float(v))`` (shortest round-trip text for
    the float64 value of each float32), so parsing with ``%lf`` and casting
    to float recovers every coordinate bit-exactly (reference parser.h:36-40).
    """
    off = 1 if one_based else 0
    integer = bool(np.all(np.rint(x) == x))
    with open(path, "w") as f:
        for row in x:
            nz = np.nonzero(row)[0]
            if integer:
                parts = [f"{j + off}:{int(row[j])}" for j in nz]
            else:
                parts = [f"{j + off}:{float(row[j])!r}" for j in nz]
            f.write("0 " + " ".join(parts) + "\n")
