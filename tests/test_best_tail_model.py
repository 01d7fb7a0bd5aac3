"""Model of the search kernel's pq-best list (song_amd/csrc/search_kernel.hpp, the insert of
`now`), against the reference's semantics: push `now`, then pop the maximum while the list holds
more than pq entries (graph.h:193-195), i.e. the pq smallest keys seen, ascending.

The kernel keeps the list's last 64 positions [tb, tb + 64) in registers (lane j = position
tb + j, +inf past nbest, tb = max(0, nbest - 64)) and the front [0, tb) in LDS.  This model
follows it lane by lane -- the DPP shifts (shl1: lane i takes lane i + 1, shr1: lane i takes
lane i - 1, the edge lane keeps its own), the entry handed to the front while the list grows
past 64, the front's last entry handed back to the tail when a full list takes a key in its
front -- and checks every step against a sorted list.

The head refill's select64 (the 64 smallest keys of an unsorted bag, ascending in lanes 0..63)
is modelled the same way: bitonic sorts and merges as sorting, the one-at-a-time insertion of
the few keys of a block below the running 64th smallest lane by lane.
"""
import numpy as np
import pytest

INF = (np.inf, 2**32)


def shl1(t):
    return t[1:] + t[-1:]


def shr1(t):
    return t[:1] + t[:-1]


class BestList:
    def __init__(self, pq):
        self.pq, self.nbest, self.tb = pq, 0, 0
        self.front = [INF] * pq  # the LDS array (positions [0, tb) live)
        self.bt = [INF] * 64     # the register tail

    def insert(self, now):
        pq, tb, bt = self.pq, self.tb, self.bt
        below = [e < now for e in bt]
        in_tail = tb == 0 or below[0]
        p = tb + sum(below) if in_tail else sum(e < now for e in self.front[:tb])
        if p >= pq:
            return
        j = p - tb
        if 64 <= self.nbest < pq:
            if j >= 0:
                self.front[tb] = now if j == 0 else bt[0]
                up = shl1(bt)
                bt = [up[i] if i < j - 1 else now if i == j - 1 else bt[i] for i in range(64)]
            else:
                self.front[p + 1:tb + 1] = self.front[p:tb]
                self.front[p] = now
            self.tb += 1
        elif j >= 0:
            dn = shr1(bt)
            bt = [bt[i] if i < j else now if i == j else dn[i] for i in range(64)]
        else:
            fl = self.front[tb - 1]
            self.front[p + 1:tb] = self.front[p:tb - 1]
            self.front[p] = now
            dn = shr1(bt)
            bt = [fl] + dn[1:]
        self.nbest = min(self.nbest + 1, pq)
        bt = [INF if i >= self.nbest - self.tb else bt[i] for i in range(64)]
        self.bt = bt
        if self.nbest == pq:
            self.bmax = bt[pq - 1 - self.tb]

    def items(self):
        return self.front[:self.tb] + self.bt[:self.nbest - self.tb]


@pytest.mark.parametrize("pq", [1, 2, 5, 63, 64, 65, 66, 100, 128, 129, 200, 300])
@pytest.mark.parametrize("order", ["ascending_noise", "random"])
def test_register_tail_keeps_the_pq_smallest(pq, order):
    rng = np.random.default_rng(pq * 7 + (order == "random"))
    n = 4 * pq + 150
    if order == "random":
        d = rng.random(n)
    else:
        # expansions come out nearly ascending, with occasional keys far below the tail
        d = np.sort(rng.random(n)) + rng.normal(0, 0.02, n)
        far = rng.random(n) < 0.05
        d[far] -= rng.random(far.sum())
    d = np.round(d, 3)  # equal distances: ties broken by id, as the kernel's keys
    keys = [(float(x), i) for i, x in enumerate(d)]
    bl, ref = BestList(pq), []
    for key in keys:
        bl.insert(key)
        ref = sorted(ref + [key])[:pq]
        assert bl.items() == ref
        assert bl.tb == max(0, bl.nbest - 64)
        if bl.nbest == pq:
            assert bl.bmax == ref[-1]


def zero_halves(x):
    return ((x - 0x00010001) & ~x & 0x80008000) & 0xFFFFFFFF


def test_swar_group_test_matches_direct_compares():
    # group_has / the first-empty choice of vis16_insert against per-slot compares, with the
    # values that stress the borrow (0, 1, 0x8000, 0xFFFF and entries with displacement bits)
    rng = np.random.default_rng(5)
    special = np.array([0, 1, 2, 0x7FFF, 0x8000, 0x8001, 0xFFFF, 0x4001, 0xC000], dtype=np.int64)
    for _ in range(20000):
        halves = np.where(rng.random(8) < 0.5, rng.choice(special, 8), rng.integers(0, 1 << 16, 8))
        words = [int(halves[2 * w]) | (int(halves[2 * w + 1]) << 16) for w in range(4)]
        e = int(rng.choice([int(rng.choice(halves)), int(rng.integers(1, 1 << 16))]))
        if e == 0:
            e = 1
        ee = e * 0x00010001
        has = any(zero_halves(w ^ ee) for w in words)
        assert has == (e in [int(h) for h in halves])
        slot = None
        for wi, w in enumerate(words):
            z = zero_halves(w)
            if z:
                slot = 2 * wi + (0 if z & 0x8000 else 1)
                break
        first_empty = next((t for t in range(8) if halves[t] == 0), None)
        assert slot == first_empty


SEL_INSERT = 16


def select64(arr):
    """Lane-level model of the kernel's select64 over arr (list of distinct keys)."""
    n = len(arr)
    blk = lambda b0: [arr[b0 + i] if b0 + i < n else INF for i in range(64)]  # noqa: E731
    best = sorted(blk(0))
    paths = {"skip": 0, "insert": 0, "merge": 0}
    for b0 in range(64, n, 64):
        v = blk(b0)
        top = best[63]
        m = [i for i in range(64) if v[i] < top]
        if len(m) > SEL_INSERT:
            best = sorted(best + sorted(v))[:64]  # bitonic sort of v, then the bitonic merge
            paths["merge"] += 1
            continue
        paths["insert" if m else "skip"] += 1
        for j in m:
            x = v[j]
            r = sum(e < x for e in best)
            dn = shr1(best)
            best = [best[i] if i < r else x if i == r else dn[i] for i in range(64)]
    return best, paths


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 127, 128, 129, 500, 1100])
@pytest.mark.parametrize("order", ["random", "ascending_noise", "descending", "ties"])
def test_select64_returns_the_64_smallest_in_order(n, order):
    rng = np.random.default_rng(n * 3 + len(order))
    if order == "random":
        d = rng.random(n)
    elif order == "ascending_noise":
        d = np.sort(rng.random(n)) + rng.normal(0, 0.05, n)
    elif order == "descending":
        d = -np.arange(n, dtype=np.float64)
    else:
        d = rng.integers(0, 5, n).astype(np.float64)  # few distances: order by id
    ids = rng.permutation(n)
    keys = [(float(x), int(i)) for x, i in zip(d, ids)]
    best, paths = select64(keys)
    ref = sorted(keys)[:64]
    assert best == ref + [INF] * (64 - len(ref))
    if order == "ascending_noise" and n >= 500:
        assert paths["merge"] <= 2  # nearly ascending bags rarely need a full merge
    if order == "descending":
        # every key beats the running 64th: blocks with more than SEL_INSERT keys are merged
        assert paths["merge"] == sum(min(64, n - b0) > SEL_INSERT for b0 in range(64, n, 64))
        assert paths["skip"] == 0


def test_select64_takes_keys_just_below_the_running_64th():
    # every later block holds keys between the running 63rd and 64th smallest (and above), so a
    # threshold one rank off, or an insertion at the wrong rank, changes the result
    keys = [(float(2 * i), i) for i in range(64)]
    nid = 64
    for b in range(6):
        top = sorted(keys)[63][0]
        blk = [(top - 0.5 - 0.001 * j, nid + j) for j in range(3)] + [(top + 1 + j, nid + 3 + j) for j in range(61)]
        nid += 64
        keys += blk
    best, paths = select64(keys)
    assert best == sorted(keys)[:64]
    assert paths["insert"] == 6
