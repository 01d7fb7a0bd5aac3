"""CPU checks of the ground-truth selection: the k smallest (distance, id) pairs of a row,
with ties at the k-th distance spanning more than topk's margin, against a sort of all pairs."""
import numpy as np
import torch

from song_amd import groundtruth


def test_smallest_k_keeps_the_smallest_ids_of_a_tie():
    rng = np.random.default_rng(0)
    for trial in range(60):
        n = int(rng.integers(50, 400))
        kk = int(rng.integers(1, 40))
        d = torch.from_numpy(rng.integers(0, int(rng.integers(2, 30)), size=(7, n)).astype(np.float64))
        ids = torch.arange(n).expand(7, n)[:, torch.from_numpy(rng.permutation(n))].contiguous()
        bd, bi = groundtruth._smallest_k(d, ids, kk, margin=int(rng.integers(0, 5)))
        for r in range(7):
            pairs = sorted(zip(d[r].tolist(), ids[r].tolist()))[:kk]
            assert [p[0] for p in pairs] == bd[r].tolist(), (trial, r)
            assert [p[1] for p in pairs] == bi[r].tolist(), (trial, r)


def test_knn_fp64_on_cpu_tensors_matches_brute_force():
    # the same chunked code path on the CPU device: several base chunks, ties across chunks
    rng = np.random.default_rng(1)
    base = rng.integers(0, 4, size=(3000, 6)).astype(np.float32)
    q = rng.integers(0, 4, size=(40, 6)).astype(np.float32)
    ids, ds = groundtruth.knn_fp64(base, q, 25, "l2", device="cpu", q_chunk=16, b_chunk=700)
    d64 = ((base.astype(np.float64)[None] - q.astype(np.float64)[:, None]) ** 2).sum(-1)
    ref = np.argsort(d64, axis=1, kind="stable")[:, :25]
    assert np.array_equal(ids, ref)
    assert np.array_equal(ds, np.take_along_axis(d64, ref, 1))
