"""Model of the search kernel's compact visited set (16-bit slots), against a Python set.

The kernel (song_amd/csrc/search_kernel.hpp, vis16_insert / vis16_test) stores ids below
2^idbits through the bijection h = id * 0x9E3779B1 mod 2^idbits: the top gbits bits of h pick
the home group of 8 slots (one 16-byte LDS read), and a slot holds the other lob = idbits - gbits
<= 13 bits as lo + 1 (never 0) with the displacement d <= 3 of its group in bits 14..15.  An id
goes to the first empty slot of the first group, at displacement 0..3, that has one; a lookup
stops at the first group holding the id's entry for that displacement or an empty slot.  Equal
(group, entry) means equal h, hence equal id, so the set is exact: these tests check it against a
Python set, with collisions, dense tables and the overflow report (no free slot in 4 groups).
"""
import numpy as np
import pytest

MULT = 0x9E3779B1


def key(idv, idbits, gbits):
    h = (idv * MULT) & ((1 << idbits) - 1)
    lob = idbits - gbits
    return h >> lob, (h & ((1 << lob) - 1)) + 1


class Vis16:
    def __init__(self, hlog, idbits):
        self.gbits = hlog - 3
        assert idbits - self.gbits <= 13
        self.idbits = idbits
        self.slots = np.zeros(1 << hlog, dtype=np.uint16)  # 8 slots per group

    def insert(self, idv):
        """0 = present, 1 = inserted, 2 = overflow (as the kernel's return codes)."""
        home, lo1 = key(idv, self.idbits, self.gbits)
        gmask = (1 << self.gbits) - 1
        for d in range(4):
            g = (home + d) & gmask
            e = lo1 | (d << 14)
            grp = self.slots[g * 8:(g + 1) * 8]
            if (grp == e).any():
                return 0
            empty = np.nonzero(grp == 0)[0]
            if len(empty):
                grp[empty[0]] = e
                return 1
        return 2

    def test(self, idv):
        home, lo1 = key(idv, self.idbits, self.gbits)
        gmask = (1 << self.gbits) - 1
        for d in range(4):
            g = (home + d) & gmask
            e = lo1 | (d << 14)
            grp = self.slots[g * 8:(g + 1) * 8]
            if (grp == e).any():
                return True
            if (grp == 0).any():
                return False
        return False


def test_hash_is_a_bijection_on_the_id_range():
    for idbits in (10, 14):
        hs = {(i * MULT) & ((1 << idbits) - 1) for i in range(1 << idbits)}
        assert len(hs) == 1 << idbits


@pytest.mark.parametrize("hlog,idbits", [(10, 20), (12, 20), (13, 20), (12, 14), (11, 16), (13, 10)])
def test_compact_set_is_exact(hlog, idbits):
    rng = np.random.default_rng(hlog * 100 + idbits)
    vs = Vis16(hlog, idbits)
    ref = set()
    n_ids = 1 << idbits
    limit = min((1 << hlog) * 3 // 4, n_ids)  # the kernel's spill threshold, or every id
    overflow = 0
    while len(ref) < limit:
        # repeats of known ids mixed with new ones, as a search rediscovers neighbours
        idv = int(rng.choice(list(ref))) if ref and rng.random() < 0.5 else int(rng.integers(0, n_ids))
        r = vs.insert(idv)
        if r == 2:
            overflow += 1
            assert idv not in ref  # only an absent id can fail to find room
            continue
        assert (r == 1) == (idv not in ref), idv
        ref.add(idv)
    # every member is found and no other id is (all ids for small ranges, a sample otherwise)
    probe = range(n_ids) if n_ids <= 1 << 16 else rng.integers(0, n_ids, 200_000)
    for idv in probe:
        assert vs.test(int(idv)) == (int(idv) in ref), int(idv)
    for idv in ref:
        assert vs.test(idv)
    assert overflow <= len(ref) // 100  # overflows are rare at the 3/4 load limit


def test_overflow_reported_when_four_groups_are_full():
    # ids sharing one home group: the 33rd cannot find room within 4 groups of 8
    idbits, hlog = 14, 10
    vs = Vis16(hlog, idbits)
    gbits = hlog - 3
    same = [i for i in range(1 << idbits) if key(i, idbits, gbits)[0] == 5]
    results = [vs.insert(i) for i in same[:40]]
    assert results[:32] == [1] * 32
    assert results[32:] == [2] * 8
    assert all(vs.test(i) for i in same[:32]) and not any(vs.test(i) for i in same[32:40])
