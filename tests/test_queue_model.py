"""Lane-level model of the search kernel's two-level queue, checked against a sorted queue.

The reference keeps an unbounded binary heap of (distance, id) (graph.h:169-216).  The kernel
(song_amd/csrc/search_kernel.hpp) replaces it with a sorted head of at most 64 keys plus an
unsorted bag, refilled by `select64` (bitonic sort per 64-key block, keep-smallest merges) and
truncated by `select_dk` (radix select over 8-bit digits of the 32-bit distance key).  This file
restates those lane programs in numpy, 64 lanes as a vector, and checks:
  * `select64` returns the 64 smallest keys in order (np.sort);
  * `select_dk` returns the r-th smallest distance, ties included (np.sort);
  * the whole queue pops the same keys as a sorted queue under the kernel's drop rules (cut at the
    best list's largest distance, and at the pq-th smallest queued distance, ties kept) and never
    exceeds its capacity.
These are CPU tests of the data structure's design; the GPU tests compare the kernel itself with
the oracle search.
"""
import numpy as np
import pytest

INF = np.uint64(0xFFFFFFFFFFFFFFFF)
LANES = np.arange(64)


def bitonic_steps(v, K, J):
    """Compare-exchange stages J/2, J/4, .., 1 of a bitonic sort of block size K (kernel: bitonic_steps)."""
    up = (LANES & K) == 0
    j = J >> 1
    while j > 0:
        o = v[LANES ^ j]
        lower = (LANES & j) == 0
        keep_min = lower == up
        v = np.where(keep_min, np.minimum(v, o), np.maximum(v, o))
        j >>= 1
    return v


def select64(arr):
    """The 64 smallest keys of arr in ascending order, +inf padded (kernel: select64)."""
    best = np.full(64, INF, dtype=np.uint64)
    for b0 in range(0, len(arr), 64):
        v = np.full(64, INF, dtype=np.uint64)
        blk = arr[b0:b0 + 64]
        v[:len(blk)] = blk
        k = 2
        while k <= 64:
            v = bitonic_steps(v, k, k)
            k <<= 1
        # reversed block against the sorted best: a bitonic sequence whose lower half is the 64 smallest
        best = bitonic_steps(np.minimum(best, v[63 - LANES]), 64, 64)
    return best


def select_dk(dks, r, bits=32):
    """r-th smallest (1-based) of the distance keys, ties counted (kernel: select_dk).

    Four rounds over 8-bit digits, most significant first; each round builds a 256-bin histogram
    of the keys that match the digits chosen so far, and 64 lanes scan it four bins each."""
    prefix, pmask, need = 0, 0, r
    for shift in range(bits - 8, -1, -8):
        hist = np.zeros(256, dtype=np.int64)
        for d in dks:
            if (int(d) & pmask) == prefix:
                hist[(int(d) >> shift) & 255] += 1
        c = hist.reshape(64, 4)
        s = c.sum(1)
        incl = np.cumsum(s)
        excl = incl - s
        L = int(np.nonzero((excl < need) & (incl >= need))[0][0])
        below = excl[L]
        digit = 4 * L
        for t in range(3):
            if below + c[L, t] < need:
                below += c[L, t]
                digit += 1
            else:
                break
        need -= below
        prefix |= digit << shift
        pmask |= 255 << shift
    return prefix


def test_select64_returns_the_64_smallest_in_order():
    rng = np.random.default_rng(1)
    for trial in range(200):
        n = int(rng.integers(1, 1500))
        keys = rng.choice(np.uint64(1) << np.uint64(40), size=n, replace=False).astype(np.uint64)
        sel = select64(keys)
        c = min(64, n)
        assert np.array_equal(sel[:c], np.sort(keys)[:c]), trial
        assert np.all(sel[c:] == INF), trial


def test_select_dk_is_the_rth_smallest_with_ties():
    rng = np.random.default_rng(2)
    for trial in range(200):
        n = int(rng.integers(1, 1500))
        d = rng.integers(0, 1 << 32, size=n, dtype=np.uint64)
        d = d % np.uint64(int(rng.integers(1, 5000)))  # many ties
        r = int(rng.integers(1, n + 1))
        assert select_dk(d, r) == int(np.sort(d)[r - 1]), trial
    # the extremes of the key range
    d = np.array([0, 0xFFFFFFFF, 0x80000000, 0x7FFFFFFF], dtype=np.uint64)
    assert [select_dk(d, r) for r in range(1, 5)] == [0, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF]


class Overflow(Exception):
    """The tie group at the pq-th distance does not fit the slack (the kernel flags the query)."""


def run_queue(pq, slack, seed, steps=1500, start_dist=1 << 20, quantum=1, mmax=32, quiet=0):
    """Drive the two-level queue and a sorted queue with the same candidates.

    Returns (steps run, head refills, bag truncations).

    Each step brings up to mmax - 1 candidates whose distances are the popped distance plus a
    uniform offset in [-50, 400), rounded down to a multiple of `quantum` (larger quanta give larger
    tie groups).  With quiet > 0 every 40 busy steps are followed by `quiet` steps without
    candidates, so the head drains and is refilled from the bag."""
    rng = np.random.default_rng(seed)
    qcap = pq + slack + 128
    head, bag, T = [], [], None  # head sorted (<= 64 after each step), bag unsorted, head <= T < bag
    ref = []                     # sorted queue with the same drop rules, truncated eagerly
    best = []                    # expanded keys, the pq smallest
    ids = iter(range(1, 10 ** 9))
    start = (start_dist << 32) | 0  # the start vertex, id 0
    head, ref = [start], [start]
    refills = truncs = 0
    for step in range(steps):
        if not head and bag:  # refill with the exact 64 smallest bag keys
            refills += 1
            sel = select64(np.array(bag, dtype=np.uint64))
            c = min(64, len(bag))
            head = [int(x) for x in sel[:c]]
            top = head[-1]
            bag = [e for e in bag if e > top]
            T = top if bag else None
        if not head:
            assert not ref
            break
        now = head.pop(0)
        bmax = best[pq - 1] if len(best) >= pq else None
        if bmax is not None and (bmax >> 32) < (now >> 32):  # stop rule, graph.h:188
            # keys the sorted queue dropped may still sit in the head: popping one must stop the
            # search, and the sorted queue must stop at the same point
            assert not ref or (ref[0] >> 32) > (bmax >> 32), step
            break
        assert ref and now == ref.pop(0), step
        best = sorted(best + [now])[:pq]
        bmax = best[pq - 1] if len(best) >= pq else None
        # new candidates at distances around now's, unique ids
        m = int(rng.integers(0, mmax)) if quiet == 0 or step % (40 + quiet) < 40 else 0
        dist = (now >> 32) + rng.integers(-50, 400, size=m)
        cand = [((int(max(dd, 0)) // quantum * quantum) << 32) | next(ids) for dd in dist]
        cut = (bmax >> 32) if bmax is not None else None
        surv = [c for c in cand if cut is None or (c >> 32) <= cut]
        if not surv:
            continue
        ref = sorted(ref + surv)
        c2 = cut
        if len(ref) > pq:
            c2 = ref[pq - 1] >> 32 if c2 is None else min(c2, ref[pq - 1] >> 32)
        if c2 is not None:
            ref = [e for e in ref if (e >> 32) <= c2]
        if len(bag) + 62 > qcap:  # lazy truncation before an append could overflow
            truncs += 1
            dcut = cut
            if pq <= len(head):
                dc = head[pq - 1] >> 32
                dcut = dc if dcut is None else min(dcut, dc)
            elif pq - len(head) <= len(bag):
                dc = select_dk(np.array([e >> 32 for e in bag], dtype=np.uint64), pq - len(head))
                dcut = dc if dcut is None else min(dcut, dc)
            if dcut is not None:
                bag = [e for e in bag if (e >> 32) <= dcut]
            if not bag:
                T = None
            if len(bag) + 62 > qcap:
                raise Overflow(step)
        to_b = [c for c in surv if T is not None and c > T]
        bag += to_b
        head = sorted(head + [c for c in surv if not (T is not None and c > T)])
        if len(head) > 64:
            bag += head[64:]
            head = head[:64]
            T = head[-1]
        assert len(bag) <= qcap and len(head) <= 64
        assert T is None or (max(head, default=0) <= T and all(e > T for e in bag)), step
    return step, refills, truncs


@pytest.mark.parametrize("slack", [0, 8, 32])
def test_small_budgets_pop_like_a_sorted_queue(slack):
    for pq in (20, 64):
        steps, refills, truncs = run_queue(pq, slack, seed=pq * 131 + slack)
        assert steps > 40 and truncs > 0, (pq, steps, truncs)


@pytest.mark.parametrize("pq", [100, 300, 768])
@pytest.mark.parametrize("slack", [0, 8, 32])
def test_two_level_queue_pops_like_a_sorted_queue(pq, slack):
    # busy phases fill the bag and force truncations; quiet phases drain the head so it is refilled
    steps, refills, truncs = run_queue(pq, slack, seed=pq * 131 + slack, quiet=100)
    assert steps > 100 and refills > 0 and truncs > 0, (steps, refills, truncs)


@pytest.mark.parametrize("pq", [100, 300, 768])
def test_two_level_queue_with_ties(pq):
    # distances in steps of 16: tie groups of a few keys at every distance, the pq-th one included
    steps, refills, truncs = run_queue(pq, 32, seed=pq, quantum=16, quiet=100)
    assert steps > 100 and refills > 0 and truncs > 0, (steps, refills, truncs)


def test_tie_pile_beyond_the_slack_is_reported():
    # every candidate at distance 0: the tie group grows without bound and must be reported
    with pytest.raises(Overflow):
        run_queue(20, 0, seed=5, start_dist=0, quantum=1 << 30)
