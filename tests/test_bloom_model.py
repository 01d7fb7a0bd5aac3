"""Blocked Bloom visited filter on the CPU: the sizing restates the reference's
auto_tune_bloom (auto_tune_bloom.cc:6-37, 65-86; golden outputs of the reference tool in
tests/golden/auto_tune_bloom.json), and a numpy model of the filter (one 32-bit word per id,
k bits from two murmur3 finalizer mixes, as bloom_insert in the search kernel) has no false
negatives and a false-positive rate at the blocked-filter expectation."""
import json
import math
import os

import numpy as np

import song_amd

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "auto_tune_bloom.json")
M32 = np.uint64(0xFFFFFFFF)


def fmix32(h):
    h = np.asarray(h, dtype=np.uint64) & M32
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x85EBCA6B)) & M32
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0xC2B2AE35)) & M32
    h ^= h >> np.uint64(16)
    return h


def model_masks(ids, wlog, k):
    ids = np.asarray(ids, dtype=np.uint64)
    h1 = fmix32((ids * np.uint64(2) + np.uint64(1)) & M32)
    h2 = fmix32(ids ^ np.uint64(0x9E3779B9))
    mask = np.zeros(len(ids), dtype=np.uint64)
    for i in range(min(k, 6)):
        mask |= np.uint64(1) << ((h2 >> np.uint64(5 * i)) & np.uint64(31))
    for i in range(max(0, k - 6)):
        mask |= np.uint64(1) << ((h1 >> np.uint64(5 * i)) & np.uint64(31))
    return (h1 >> np.uint64(32 - wlog)).astype(np.int64), mask


class BloomModel:
    def __init__(self, wlog, k):
        self.words = np.zeros(1 << wlog, dtype=np.uint64)
        self.wlog, self.k = wlog, k

    def insert(self, ids):
        """sequential inserts; returns the 'new' flags"""
        w, m = model_masks(ids, self.wlog, self.k)
        new = np.empty(len(w), dtype=bool)
        for i, (wi, mi) in enumerate(zip(w, m)):
            old = self.words[wi]
            self.words[wi] = old | mi
            new[i] = (old & mi) != mi
        return new


def test_sizing_equals_the_reference_tool():
    with open(GOLDEN) as f:
        sizes = json.load(f)["sizes"]
    for pq, ref in sizes.items():
        hlog, k = song_amd.bloom_size(int(pq))
        assert hlog == ref["log2_words64"] + 1, pq  # 32-bit words
        assert k == min(9, ref["k"]), pq


def test_model_has_no_false_negatives_and_the_expected_false_positive_rate():
    rng = np.random.default_rng(0)
    for pq in (64, 128, 256):
        wlog, k = song_amd.bloom_size(pq)
        n = int(12.897 * pq)  # the sizing's design load
        ids = rng.choice(1 << 24, size=2 * n, replace=False)
        bf = BloomModel(wlog, k)
        # distinct ids: "already present" during the inserts are false positives too, and rare
        assert bf.insert(ids[:n]).mean() > 0.97
        # every inserted id now tests present
        w, m = model_masks(ids[:n], wlog, k)
        assert ((bf.words[w] & m) == m).all()
        # false positives among fresh ids, against the blocked-filter expectation: a word holding
        # j ids has about 32 (1 - (1 - 1/32)^(k j)) bits set; j ~ Poisson(n / W)
        w2, m2 = model_masks(ids[n:], wlog, k)
        fp = float(((bf.words[w2] & m2) == m2).mean())
        lam = n / (1 << wlog)
        exp = sum(math.exp(-lam) * lam ** j / math.factorial(j) * (1 - (1 - 1 / 32) ** (k * j)) ** k
                  for j in range(0, 60))
        print(f"pq {pq}: 2^{wlog} words, k {k}, false positives {fp:.4f} (expected {exp:.4f})")
        assert abs(fp - exp) < 0.25 * exp + 0.002


def test_bit_positions_use_independent_slices():
    # k bits from 5-bit slices: for k <= 6 all from the second mix; 7..9 add slices of the first
    ids = np.arange(100000, dtype=np.uint64)
    for k in (1, 4, 7, 9):
        _, m = model_masks(ids, 12, k)
        bits = np.array([bin(int(x)).count("1") for x in m[:5000]])
        assert bits.max() <= k and bits.mean() > k * 0.7
