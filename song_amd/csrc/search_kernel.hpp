// Best-first graph search on gfx950: one query per 64-lane wavefront.
//
// Semantics follow the reference CPU search FixedDegreeGraph::astar_multi_start_search
// (graph.h:169-216): start at the entry vertex, repeatedly expand the (dist, id)-smallest
// unexpanded candidate, stop when the k-best expanded list is full and its largest distance
// is strictly below the next candidate's (graph.h:188), mark every discovered neighbour
// visited (graph.h:198-205), return the expanded list ascending by (dist, id).
//
// The CPU's unbounded queue is replaced by a bounded one that only ever drops entries that
// can never be expanded: an entry c is dropped when the pq-best expanded list is full with
// largest distance < c.d, or when the queue holds pq entries with distance < c.d.  Entries
// tied with the pq-th distance are kept (up to `slack`), so the expansion order -- and
// therefore the output -- equals the CPU's.  The queue is two-level: a sorted head of the
// (up to) 64 smallest entries, from which the minimum is popped, and an unsorted bag of the
// rest (every head entry <= T < every bag entry).  New candidates at most T are merged into
// the head, the others appended to the bag; the head's overflow moves to the bag; an empty
// head is refilled with the bag's 64 smallest (exact selection); the bag is bounded lazily
// (the drops above are applied when it nears its capacity).
//
// Per-wave LDS layout (bytes, in order):
//   visited  : H slots, no deletion: compact 16-bit entries in groups of 8 (vis16_insert) or
//              32-bit (id + 1) in groups of 4 (vis_insert), one 16-byte read per group; or,
//              approximate, a blocked Bloom filter of 2^hlog 32-bit words (bloom_insert)
//   scratch  : 64 keys    merge staging of the queue head (the head itself is one key per
//                          lane in registers, ascending over lanes [h, h + nq), nq <= 64)
//   best     : pq keys, sorted ascending, first nbest live (pq-best expanded list); absent
//              with bglob, where the list's front lives in the wave slot's global region
//   cid      : 256 u32   candidate ids of the current expansion / radix-select bins
//   qv       : qld f32   the query, zero-padded to a multiple of 16
// The bag (qcap = pq + slack + 128 keys, unsorted, first bn live) is in global memory, one
// region per wave slot: it is written by plain stores as candidates arrive and read only when
// the head runs empty or the bag nears its capacity, so LDS holds only what every step
// touches and more waves fit on a CU.  The best list's last 64 positions are one key per lane
// in registers; its front (the first nbest - 64) is written once per entry while the list
// grows and read only when an expansion lands below the register tail, which is rare, and at
// the end -- so with bglob it lives after the bag in the slot's global region (pq keys) and a
// large pq does not cost LDS.
//
// Each wave owns its LDS slice (one wave per workgroup in search_kernel; 1-4 waves, each
// on its own slice, in search_persistent).  LDS instructions of a wave execute in issue
// order, so lanes exchange data through LDS with only a compiler fence between the write
// and the read (lds_fence), and global loads stay in flight across those points.
//
// Per expansion the dependent global round trips are: adjacency row of the expanded
// vertex (prefetched one step ahead for the predicted next vertex, the queue head after
// the pop) and one gather of all new candidates' rows (all rows issued at once).
//
// Keys order by (distance, id) lexicographically:
//   KeyF (fp32 mode):  one u64 = (orderable float bits << 32) | id
//   KeyD (exact mode): orderable double bits + u32 id
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace song {

enum { M_L2 = 0, M_IP = 1, M_COS = 2 };

enum : uint32_t { F_VIS_OVF = 1u, F_Q_OVF = 2u, F_STEP_CAP = 4u, F_BAD_LAUNCH = 0x80000000u };

struct SearchParams {
    const float *data;      // n x ld
    const uint32_t *graph;  // n x gstride, [deg, ids...]
    const float *norms;     // cosine: per-vertex squared norm (fmaf chain), else null
    const float *queries;   // nq x dim
    const uint32_t *entry;  // nq x n_entry start vertices or null (vertex 0)
    int n_entry;            // entries per query (1..64)
    const int *steps_arr;   // per-query expansion cap or null
    const int *pq_arr;      // per-query pq or null
    const int *k_arr;       // per-query k or null
    const uint32_t *qidx;   // query index remap (retry pass) or null
    uint32_t *out_ids;      // nq x out_ld
    float *out_dist;        // nq x out_ld or null
    uint32_t *out_flags;    // nq or null
    uint32_t *out_stats;    // nq x 8 or null: expansions, distances, sum(1+deg), prefetch hits,
                            // settle step (last expansion that changed the first k), head
                            // refills from the bag, bag truncations, 0
    uint64_t n;
    int dim, ld, gstride;
    int qld;                // query floats in LDS: dim padded to 16
    int nq, pq, k, out_ld;
    int hlog, slack, max_steps, qcap;
    uint32_t *gvis;         // per-launch-slot spill tables (2^ghlog u32 each) or null
    int ghlog;
    int qbase;              // first query (or qidx entry) of this launch
    const uint8_t *data_lp; // low-precision rows (BF16 / FP8 kernels), lp_stride bytes each
    const float *scale;     // per-row scale of data_lp
    size_t lp_stride;
    int rerank;             // low-precision kernels: re-rank the final list in exact FP32
    // persistent kernel: waves take list positions t = (*work_next)++ while t < *work_count
    // and search query qidx[t] (t itself without a list)
    uint32_t *work_next;
    const uint32_t *work_count;
    uint32_t wave_lds;      // LDS bytes of one wave's slice (every launch; the layout must fit)
    int vis16;              // 1: compact visited table (16-bit slots, needs idbits - (hlog - 3) <= 13)
    int idbits;             // ids are below 2^idbits
    unsigned char *gbag;    // queue bags, one per launch slot / persistent wave, gbag_stride bytes
    size_t gbag_stride;     // >= qcap keys
    // queries whose visited set or queue overflowed are appended here for a re-run
    uint32_t *retry_list;
    uint32_t *retry_count;
    int wave_prio;          // persistent kernel: 1 = raise the waves' issue priority (s_setprio 2)
    int bloom_k;            // 0: exact visited table; 1..9: blocked Bloom filter of 2^hlog 32-bit words,
                            // bloom_k bits per id in one word (approximate: false positives skip vertices)
    int multiprobe;         // several-queries-per-wave kernel: expansions per step (1, or 2: the next
                            // queue head too; approximate)
};

// bytes of a slot's bag before the best-list front (bglob): qcap keys rounded to 256
template <class Key> __device__ __forceinline__ size_t bag_bytes(int qcap) {
    return ((size_t)qcap * sizeof(Key) + 255) & ~(size_t)255;
}

// ---------------------------------------------------------------- keys

__device__ __forceinline__ uint32_t ord32(float f) {
    uint32_t u = __float_as_uint(f);
    if (f == 0.0f) u = 0u;            // -0 and +0 compare equal on the CPU
    if (f != f) u = 0x7f800000u;      // NaN (zero-norm cosine) sorts as +inf
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord32(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(u);
}
__device__ __forceinline__ uint64_t ord64(double d) {
    uint64_t u = (uint64_t)__double_as_longlong(d);
    if (d == 0.0) u = 0ull;
    if (d != d) u = 0x7ff0000000000000ull;
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double unord64(uint64_t o) {
    uint64_t u = (o & 0x8000000000000000ull) ? (o & 0x7fffffffffffffffull) : ~o;
    return __longlong_as_double((long long)u);
}

__device__ __forceinline__ uint32_t rl32(uint32_t v, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ __forceinline__ uint64_t rl64(uint64_t v, int lane) {
    return ((uint64_t)rl32((uint32_t)(v >> 32), lane) << 32) | rl32((uint32_t)v, lane);
}

struct KeyF {
    static constexpr int DKBITS = 32;  // significant bits of dk()
    uint64_t v;
    __device__ static KeyF make_f(float d, uint32_t id) { return KeyF{((uint64_t)ord32(d) << 32) | id}; }
    __device__ static KeyF inf() { return KeyF{~0ull}; }
    __device__ uint32_t id() const { return (uint32_t)v; }
    __device__ uint64_t dk() const { return v >> 32; }
    __device__ float dist() const { return unord32((uint32_t)(v >> 32)); }
    __device__ KeyF rl(int lane) const { return KeyF{rl64(v, lane)}; }
};
__device__ __forceinline__ bool lt(const KeyF &a, const KeyF &b) { return a.v < b.v; }

struct KeyD {
    static constexpr int DKBITS = 64;
    uint64_t d;
    uint32_t i, pad;
    __device__ static KeyD make_d(double dist, uint32_t id) { return KeyD{ord64(dist), id, 0u}; }
    __device__ static KeyD inf() { return KeyD{~0ull, ~0u, 0u}; }
    __device__ uint32_t id() const { return i; }
    __device__ uint64_t dk() const { return d; }
    __device__ float dist() const { return (float)unord64(d); }
    __device__ KeyD rl(int lane) const { return KeyD{rl64(d, lane), rl32(i, lane), 0u}; }
};
__device__ __forceinline__ bool lt(const KeyD &a, const KeyD &b) {
    return a.d < b.d || (a.d == b.d && a.i < b.i);
}

__device__ __forceinline__ KeyF rl_key(const KeyF &k, int lane) { return k.rl(lane); }
__device__ __forceinline__ KeyD rl_key(const KeyD &k, int lane) { return k.rl(lane); }

// whole-wave lane shifts by one (DPP wave_shl:1 / wave_shr:1): shl1 gives lane i the value of
// lane i + 1 (lane 63 keeps its own), shr1 the value of lane i - 1 (lane 0 keeps its own)
template <int CTRL> __device__ __forceinline__ uint32_t dpp32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, CTRL, 0xF, 0xF, false);
}
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;
template <int CTRL> __device__ __forceinline__ uint64_t dpp64(uint64_t x) {
    return ((uint64_t)dpp32<CTRL>((uint32_t)(x >> 32)) << 32) | dpp32<CTRL>((uint32_t)x);
}
// c ? a : b field by field (a conditional over two key lvalues selects between their addresses,
// which puts the keys in scratch memory)
__device__ __forceinline__ KeyF ksel(bool c, const KeyF &a, const KeyF &b) { return KeyF{c ? a.v : b.v}; }
__device__ __forceinline__ KeyD ksel(bool c, const KeyD &a, const KeyD &b) {
    return KeyD{c ? a.d : b.d, c ? a.i : b.i, 0u};
}
__device__ __forceinline__ KeyF shl1(const KeyF &k) { return KeyF{dpp64<DPP_WAVE_SHL1>(k.v)}; }
__device__ __forceinline__ KeyF shr1(const KeyF &k) { return KeyF{dpp64<DPP_WAVE_SHR1>(k.v)}; }
__device__ __forceinline__ KeyD shl1(const KeyD &k) {
    return KeyD{dpp64<DPP_WAVE_SHL1>(k.d), dpp32<DPP_WAVE_SHL1>(k.i), 0u};
}
__device__ __forceinline__ KeyD shr1(const KeyD &k) {
    return KeyD{dpp64<DPP_WAVE_SHR1>(k.d), dpp32<DPP_WAVE_SHR1>(k.i), 0u};
}

template <bool EXACT> struct KeySel { using T = KeyF; };
template <> struct KeySel<true> { using T = KeyD; };

// ---------------------------------------------------------------- wave helpers

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int lanes_below(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// compiler-only fence for LDS traffic inside one wave (see the layout note above)
__device__ __forceinline__ void lds_fence() { __atomic_signal_fence(__ATOMIC_SEQ_CST); }
// wait for every vector-memory operation of this wave (s_waitcnt vmcnt(0)).  Placed at the end
// of the rare bag paths (refill, truncation): their loads land in registers that the common path
// reuses, and without a wait on the rare path itself the compiler has to wait for all memory
// operations -- the adjacency prefetch included -- at the top of every expansion.
__device__ __forceinline__ void vm_drain() { __builtin_amdgcn_s_waitcnt(0x0F70); }
// before lanes read bag entries that other lanes of the wave stored to global memory: the
// stores complete first (workgroup scope: the wave's own CU)
__device__ __forceinline__ void bag_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup"); }

// lowest set bit exponent of a nonzero float (value = odd * 2^e); INT_MAX for 0
__device__ __forceinline__ int lsb_exp(float p) {
    uint32_t u = __float_as_uint(p) & 0x7fffffffu;
    if (u == 0u) return 0x3fffffff;
    int e = (int)(u >> 23);
    uint32_t m = u & 0x7fffffu;
    if (e) m |= 0x800000u;
    else e = 1;
    return e - 150 + __builtin_ctz(m);
}

// visited set: (id + 1) in an open-addressing table probed in aligned groups of 4 slots,
// one 16-byte LDS read per group.  An id goes to the first empty slot of the first group
// on its probe sequence that has one; with no deletions a lookup can therefore stop at the
// first group holding an empty slot.  Lanes inserting the same id race through
// compare-and-swap; exactly one of them sees the slot empty.  Returns true if id was absent.
__device__ __forceinline__ bool vis_insert(uint32_t *vis, uint32_t mask, int hlog, uint32_t id) {
    const uint32_t key = id + 1u;
    const uint32_t gmask = mask >> 2;
    uint32_t g = (id * 0x9E3779B1u) >> (34 - hlog);
    for (;;) {
        const uint4 v = *reinterpret_cast<const uint4 *>(vis + (g << 2));
        if (v.x == key || v.y == key || v.z == key || v.w == key) return false;
        const uint32_t e = v.x == 0u ? 0u : v.y == 0u ? 1u : v.z == 0u ? 2u : v.w == 0u ? 3u : 4u;
        if (e == 4u) {
            g = (g + 1u) & gmask;
            continue;
        }
        const uint32_t old = atomicCAS(&vis[(g << 2) + e], 0u, key);
        if (old == 0u) return true;
        if (old == key) return false;
        lds_fence();  // another id took the slot: read the group again
    }
}

// read-only lookup in the LDS table
__device__ __forceinline__ bool vis_test(const uint32_t *vis, uint32_t mask, int hlog, uint32_t id) {
    const uint32_t key = id + 1u;
    const uint32_t gmask = mask >> 2;
    uint32_t g = (id * 0x9E3779B1u) >> (34 - hlog);
    for (;;) {
        const uint4 v = *reinterpret_cast<const uint4 *>(vis + (g << 2));
        if (v.x == key || v.y == key || v.z == key || v.w == key) return true;
        if (v.x == 0u || v.y == 0u || v.z == 0u || v.w == 0u) return false;
        g = (g + 1u) & gmask;
    }
}

// Compact visited set (16-bit slots, half the LDS for the same slot count).  For ids below
// 2^idbits, h = id * 0x9E3779B1 mod 2^idbits is a bijection: its top gbits = hlog - 3 bits pick
// the home group of 8 slots (one 16-byte read), and a slot holds the other lob = idbits - gbits
// <= 13 bits as lo + 1 (never 0) with bits 14..15 = the displacement d <= 3 of the group that
// holds it.  Equal entries at equal displacement from equal homes mean equal h, hence equal ids,
// so the set is exact.  An id takes the first empty slot of the first group at d = 0..3 that has
// one (lookups stop there as in the 32-bit table); with none free in those 4 groups the insert
// reports an overflow and the caller continues in the global table.  Lanes inserting the same id
// read the same group state, choose the same slot and race through compare-and-swap on its word.
__device__ __forceinline__ void vis16_key(uint32_t id, int idbits, int gbits, uint32_t &home, uint32_t &lo1) {
    const uint32_t hm = idbits >= 32 ? 0xFFFFFFFFu : ((1u << idbits) - 1u);
    const uint32_t h = (id * 0x9E3779B1u) & hm;
    const int lob = idbits - gbits;
    home = h >> lob;
    lo1 = (h & ((1u << lob) - 1u)) + 1u;
}
// the zero 16-bit halves of x, flagged in bits 15 and 31: (x - 0x00010001) & ~x & 0x80008000.
// A zero low half is always flagged and a nonzero one never is; the high half's flag is exact
// unless the low half is zero (its borrow can flag a high half of 1).  So "some half is zero" and
// "the lowest zero half" read exactly from the flags.
__device__ __forceinline__ uint32_t zero_halves(uint32_t x) { return (x - 0x00010001u) & ~x & 0x80008000u; }
// some half of the group equals e (ee = e in both halves)
__device__ __forceinline__ bool group_has(const uint4 &v, uint32_t ee) {
    return (zero_halves(v.x ^ ee) | zero_halves(v.y ^ ee) | zero_halves(v.z ^ ee) | zero_halves(v.w ^ ee)) != 0u;
}
// 0: present, 1: inserted, 2: no room within 4 groups
__device__ __forceinline__ int vis16_insert(uint32_t *vis, int gbits, int idbits, uint32_t id) {
    uint32_t home, lo1;
    vis16_key(id, idbits, gbits, home, lo1);
    const uint32_t gmask = (1u << gbits) - 1u;
    for (uint32_t d = 0; d < 4u; ++d) {
        const uint32_t g = (home + d) & gmask;
        const uint32_t e = lo1 | (d << 14);
        for (;;) {
            const uint4 v = *reinterpret_cast<const uint4 *>(vis + (g << 2));
            if (group_has(v, e * 0x00010001u)) return 0;
            // the first empty slot: the first word with a zero half, then its lower zero half
            uint32_t old = v.x, z = zero_halves(v.x), wi = 0u;
            if (z == 0u) { old = v.y; z = zero_halves(v.y); wi = 1u; }
            if (z == 0u) { old = v.z; z = zero_halves(v.z); wi = 2u; }
            if (z == 0u) { old = v.w; z = zero_halves(v.w); wi = 3u; }
            if (z == 0u) break;  // group full: next displacement
            const uint32_t nw = old | (e << ((z & 0x8000u) ? 0 : 16));
            if (atomicCAS(&vis[(g << 2) + wi], old, nw) == old) return 1;
            lds_fence();  // the word changed: read the group again
        }
    }
    return 2;
}
__device__ __forceinline__ bool vis16_test(const uint32_t *vis, int gbits, int idbits, uint32_t id) {
    uint32_t home, lo1;
    vis16_key(id, idbits, gbits, home, lo1);
    const uint32_t gmask = (1u << gbits) - 1u;
    for (uint32_t d = 0; d < 4u; ++d) {
        const uint32_t g = (home + d) & gmask;
        const uint32_t e = lo1 | (d << 14);
        const uint4 v = *reinterpret_cast<const uint4 *>(vis + (g << 2));
        if (group_has(v, e * 0x00010001u)) return true;
        if ((zero_halves(v.x) | zero_halves(v.y) | zero_halves(v.z) | zero_halves(v.w)) != 0u) return false;
    }
    return false;
}

// insert into this query's global spill table (same scheme, device-scope CAS)
__device__ __forceinline__ bool gvis_insert(uint32_t *g, uint32_t mask, int hlog, uint32_t id) {
    const uint32_t key = id + 1u;
    uint32_t s = (id * 0x9E3779B1u) >> (32 - hlog);
    for (;;) {
        const uint32_t old = atomicCAS(&g[s], 0u, key);
        if (old == 0u) return true;
        if (old == key) return false;
        s = (s + 1u) & mask;
    }
}

// Blocked Bloom filter (the reference's BlockedBloomFilter, blocked_bloomfilter.h:9-82, with
// one 32-bit word per block): an id sets k bits of one word, all chosen by two murmur3
// finalizer mixes of the id -- the word from the top wlog bits of the first, the bit positions
// from 5-bit slices of the second (positions 7..9 from the first's low bits, below the word
// index for wlog <= 17).  Insert = one LDS atomic OR; the id counts as new unless all k bits
// were already set.  Lanes inserting the same id see exactly one "new" (the ORs serialise).
// No overflow and no deletion; a false positive skips a vertex that was never visited.
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
__device__ __forceinline__ uint32_t bloom_mask(uint32_t h1, uint32_t h2, int k) {
    uint32_t m = 0u;
#pragma unroll
    for (int i = 0; i < 6; ++i)
        if (i < k) m |= 1u << ((h2 >> (5 * i)) & 31u);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (6 + i < k) m |= 1u << ((h1 >> (5 * i)) & 31u);
    return m;
}
__device__ __forceinline__ bool bloom_insert(uint32_t *bf, int wlog, int k, uint32_t id) {
    const uint32_t h1 = fmix32(id * 2u + 1u), h2 = fmix32(id ^ 0x9E3779B9u);
    const uint32_t mask = bloom_mask(h1, h2, k);
    const uint32_t old = atomicOr(&bf[h1 >> (32 - wlog)], mask);
    return (old & mask) != mask;
}

// ---------------------------------------------------------------- distances

// Final distance from the accumulated dot/L2 sum, per reference data.h:28-62.
template <int METRIC>
__device__ __forceinline__ double finish_exact(double s, float lena, float lenv) {
    if (METRIC == M_L2) return s;
    if (METRIC == M_IP) return -s;
    int sign = s < 0 ? 1 : -1;
    return (double)sign * (s * s / (double)lena / (double)lenv);
}
template <int METRIC>
__device__ __forceinline__ float finish_f32(float s, float lena, float lenv) {
    if (METRIC == M_L2) return s;
    if (METRIC == M_IP) return -s;
    float sign = s < 0 ? 1.0f : -1.0f;
    return sign * (s * s / lena / lenv);
}

// One lane, sequential over the dimensions in the CPU's order (exact fallback).
template <int METRIC>
__device__ double seq_sum(const float *row, const float *qv, int dim) {
    double s = 0.0;
    for (int i = 0; i < dim; ++i) {
        float x = row[i], q = qv[i];
        float p;
        if (METRIC == M_L2) {
            float d = x - q;
            p = d * d;
        } else {
            p = x * q;
        }
        s += (double)p;
    }
    return s;
}

// ---------------------------------------------------------------- row storage

// Base rows are read in 16-byte units.  ST_F32: the float32 rows (bfsg.data values,
// zero-padded); ST_BF16 / ST_FP8: a low-precision copy, value = code * scale[row], with a
// power-of-two scale for BF16 (lossless for integers up to 256) and max|x|/448 for FP8-E4M3.
enum { ST_F32 = 0, ST_BF16 = 1, ST_FP8 = 2 };
template <int STORE> struct Store;
template <> struct Store<ST_F32> { static constexpr int EPU = 4; };
template <> struct Store<ST_BF16> { static constexpr int EPU = 8; };
template <> struct Store<ST_FP8> { static constexpr int EPU = 16; };

template <int STORE>
__device__ __forceinline__ void decode_unit(const uint4 u, float (&x)[Store<STORE>::EPU]) {
    if constexpr (STORE == ST_F32) {
        x[0] = __uint_as_float(u.x); x[1] = __uint_as_float(u.y);
        x[2] = __uint_as_float(u.z); x[3] = __uint_as_float(u.w);
    } else if constexpr (STORE == ST_BF16) {
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            x[2 * t] = __uint_as_float(w[t] << 16);
            x[2 * t + 1] = __uint_as_float(w[t] & 0xffff0000u);
        }
    } else {
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const auto lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)w[t], false);
            const auto hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)w[t], true);
            x[4 * t] = lo[0]; x[4 * t + 1] = lo[1]; x[4 * t + 2] = hi[0]; x[4 * t + 3] = hi[1];
        }
    }
}

// Distances of the m candidates cid[0..m): returns lane c's key (c < m).
// SUB lanes share one row; each lane keeps its share of all rows of a pass in registers
// (up to NVU 16-byte units) so every row load of the pass is issued before the first use.
template <int METRIC, bool EXACT, int STORE, int SUB, int NVU, class Key>
__device__ __forceinline__ Key eval_candidates(const SearchParams &P, int m, uint32_t cand, const uint32_t *cid,
                                               const float *qv, float lenv) {
    constexpr int EPU = Store<STORE>::EPU;
    const int lane = lane_id();
    constexpr int G = 64 / SUB;
    const int g = lane / SUB, sl = lane % SUB;
    const int nu = (STORE == ST_F32) ? (P.ld >> 2) : (int)(P.lp_stride >> 4);
    Key mine = Key::inf();
    bool redo = false;
    for (int base = 0; base < m; base += G) {
        const int c = base + g;
        const bool valid = c < m;
        const uint32_t id = valid ? cid[c] : 0u;
        const uint4 *row = (STORE == ST_F32)
                               ? reinterpret_cast<const uint4 *>(P.data + (size_t)id * (size_t)P.ld)
                               : reinterpret_cast<const uint4 *>(P.data_lp + (size_t)id * P.lp_stride);
        const float scale = (STORE != ST_F32 && valid) ? P.scale[id] : 1.0f;
        float facc = 0.0f;
        double dacc = 0.0, sab = 0.0;
        int mn = 0x3fffffff;
        for (int j0 = 0; j0 < nu; j0 += NVU * SUB) {
            uint4 u[NVU];
            if (valid) {
                // indices past the row are clamped (and skipped below): one predicate for all loads
#pragma unroll
                for (int v = 0; v < NVU; ++v) u[v] = row[min(j0 + sl + v * SUB, nu - 1)];
            }
#pragma unroll
            for (int v = 0; v < NVU; ++v) {
                const int j = j0 + sl + v * SUB;
                if (valid && j < nu) {
                    float x[EPU];
                    decode_unit<STORE>(u[v], x);
                    const float *q = qv + j * EPU;
                    if constexpr (!EXACT) {
#pragma unroll
                        for (int t = 0; t < EPU; ++t) {
                            const float xv = (STORE == ST_F32) ? x[t] : x[t] * scale;
                            if (METRIC == M_L2) {
                                const float d = xv - q[t];
                                facc += d * d;
                            } else {
                                facc += xv * q[t];
                            }
                        }
                    } else {
                        // float products (rounded), double accumulation; exact in any order when
                        // all products are multiples of 2^mn and sum |p| < 2^(mn+53)
#pragma unroll
                        for (int t = 0; t < EPU; ++t) {
                            float pr;
                            if (METRIC == M_L2) {
                                const float d = x[t] - q[t];
                                pr = d * d;
                            } else {
                                pr = x[t] * q[t];
                            }
                            dacc += (double)pr;
                            sab += (double)fabsf(pr);
                            mn = min(mn, lsb_exp(pr));
                        }
                    }
                }
            }
        }
        Key k;
        bool seq = false;
        if constexpr (!EXACT) {
#pragma unroll
            for (int o = SUB / 2; o > 0; o >>= 1) facc += __shfl_xor(facc, o);
            const float lena = (METRIC == M_COS && valid) ? P.norms[id] : 0.0f;
            k = Key::make_f(finish_f32<METRIC>(facc, lena, lenv), id);
        } else {
#pragma unroll
            for (int o = SUB / 2; o > 0; o >>= 1) {
                dacc += __shfl_xor(dacc, o);
                sab += __shfl_xor(sab, o);
                mn = min(mn, __shfl_xor(mn, o));
            }
            seq = !((sab == 0.0) || (mn < 0x3fffffff && sab < ldexp(1.0, mn + 53)));
            const float lena = (METRIC == M_COS && valid) ? P.norms[id] : 0.0f;
            k = Key::make_d(finish_exact<METRIC>(dacc, lena, lenv), id);
        }
        // lane c (base <= c < base + G) takes the key of group c - base
        const int src = ((lane - base) & (G - 1)) * SUB;
        Key got;
        if constexpr (!EXACT) {
            got.v = ((uint64_t)(uint32_t)__shfl((int)(k.v >> 32), src) << 32) | (uint32_t)__shfl((int)k.v, src);
        } else {
            got.d = ((uint64_t)(uint32_t)__shfl((int)(k.d >> 32), src) << 32) | (uint32_t)__shfl((int)k.d, src);
            got.i = (uint32_t)__shfl((int)k.i, src);
            got.pad = 0u;
        }
        const bool seq_got = __shfl((int)seq, src) != 0;
        if (lane >= base && lane < base + G && lane < m) {
            mine = got;
            redo = seq_got;
        }
    }
    if constexpr (EXACT) {
        // sequential re-evaluation in the CPU's order, one lane per flagged candidate
        if (redo) {
            const double s = seq_sum<METRIC>(P.data + (size_t)cand * (size_t)P.ld, qv, P.dim);
            const float lena = (METRIC == M_COS) ? P.norms[cand] : 0.0f;
            mine = Key::make_d(finish_exact<METRIC>(s, lena, lenv), cand);
        }
    }
    return mine;
}

// wave-wide minimum of a KeyF (all lanes get it)
__device__ __forceinline__ KeyF wave_min(KeyF k) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t v = ((uint64_t)(uint32_t)__shfl_xor((int)(k.v >> 32), o) << 32) |
                           (uint32_t)__shfl_xor((int)k.v, o);
        if (v < k.v) k.v = v;
    }
    return k;
}

// Count of the leading entries of the sorted run arr[0..n) that satisfy `below`, a
// predicate that holds on a prefix.  64-ary search: each round samples the last entry of
// 64 equal chunks (one LDS read per lane) and keeps the chunk holding the boundary.
template <class Key, class Pred>
__device__ __forceinline__ int prefix_count(const Key *arr, int n, Pred below) {
    const int lane = lane_id();
    int lo = 0, len = n;
    while (len > 64) {
        const int stride = (len + 63) >> 6;
        const int end = lo + len;
        const int idx = min(lo + (lane + 1) * stride, end) - 1;
        const bool v = lo + lane * stride < end;
        const Key e = arr[v ? idx : lo];
        const int c = __popcll(ballot(v && below(e)));
        lo = min(lo + c * stride, end);  // the last chunk may be partial
        len = min(stride, end - lo);
    }
    const bool v = lane < len;
    const Key e = arr[v ? lo + lane : 0];
    return lo + __popcll(ballot(v && below(e)));
}

// arr[p, e) moves to arr[p + 1, e + 1) and arr[p] = x (blocks of 64, last block first, so each
// block is read before the block below it writes into its first slot)
template <class Key> __device__ __forceinline__ void shift_insert(Key *arr, int p, int e, const Key &x) {
    const int lane = lane_id();
    for (int blk = (e - 1) >> 6; blk >= (p >> 6) && e > p; --blk) {
        const int i = blk * 64 + lane;
        const bool live = i >= p && i < e;
        const Key v = arr[live ? i : 0];
        if (live) arr[i + 1] = v;
    }
    if (lane == 0) arr[p] = x;
}

// ---------------------------------------------------------------- wave-wide key sets

__device__ __forceinline__ KeyF shfl_key(const KeyF &k, int src) {
    return KeyF{((uint64_t)(uint32_t)__shfl((int)(k.v >> 32), src) << 32) | (uint32_t)__shfl((int)k.v, src)};
}
__device__ __forceinline__ KeyD shfl_key(const KeyD &k, int src) {
    return KeyD{((uint64_t)(uint32_t)__shfl((int)(k.d >> 32), src) << 32) | (uint32_t)__shfl((int)k.d, src),
                (uint32_t)__shfl((int)k.i, src), 0u};
}
template <class Key> __device__ __forceinline__ Key kmin(const Key &a, const Key &b) { return lt(b, a) ? b : a; }
template <class Key> __device__ __forceinline__ Key kmax(const Key &a, const Key &b) { return lt(b, a) ? a : b; }

// Bitonic merge steps j = J/2 .. 1 over one key per lane: within blocks of K lanes the
// sequence ascends where (lane & K) == 0 and descends otherwise (K = 64: ascending).
template <class Key>
__device__ __forceinline__ Key bitonic_steps(Key v, int K, int J) {
    const int lane = lane_id();
    const bool up = (lane & K) == 0;
    for (int j = J >> 1; j > 0; j >>= 1) {
        const Key o = shfl_key(v, lane ^ j);
        v = (((lane & j) == 0) == up) ? kmin(v, o) : kmax(v, o);
    }
    return v;
}

// The c = min(64, n) smallest keys of arr[0..n) (any order), ascending in lanes 0..c-1 with
// +inf above.  The first 64-key block is sorted (bitonic); each further block is compared with the
// running 64th smallest.  A block with no key below it is skipped, up to SEL_INSERT keys below it
// are inserted one at a time (rank by ballot, the lanes above shift up one), more are sorted and
// merged: min(best[i], block[63 - i]) holds the 64 smallest of both as a bitonic sequence.
// Bags are appended in nearly ascending order, so most blocks after the first are skipped or
// take a few insertions.
constexpr int SEL_INSERT = 16;
template <class Key>
__device__ __forceinline__ Key select64(const Key *arr, int n) {
    const int lane = lane_id();
    Key best = (lane < n) ? arr[lane] : Key::inf();
    // two blocks' loads are in flight while one is sorted or scanned (two register slots, the
    // loop unrolled by two, so no slot is copied while its load is in flight)
    Key na = (64 + lane < n) ? arr[64 + lane] : Key::inf();
    Key nb = (128 + lane < n) ? arr[128 + lane] : Key::inf();
    for (int k = 2; k <= 64; k <<= 1) best = bitonic_steps(best, k, k);
    const auto take = [&](Key v) {
        uint64_t m = ballot(lt(v, best.rl(63)));
        if (__popcll(m) > SEL_INSERT) {
            for (int k = 2; k <= 64; k <<= 1) v = bitonic_steps(v, k, k);
            best = bitonic_steps(kmin(best, shfl_key(v, 63 - lane)), 64, 64);
        } else {
            for (; m != 0ull; m &= m - 1ull) {
                const Key x = v.rl(__builtin_ctzll(m));
                // rank 64 (x above the current 64th, lowered by earlier insertions): no change
                const int r = __popcll(ballot(lt(best, x)));
                const Key dn = shr1(best);
                best = ksel(lane < r, best, ksel(lane == r, x, dn));
            }
        }
    };
    for (int b0 = 64; b0 < n; b0 += 128) {
        const Key va = na;
        if (b0 + 128 < n) na = (b0 + 128 + lane < n) ? arr[b0 + 128 + lane] : Key::inf();
        take(va);
        if (b0 + 64 < n) {
            const Key vb = nb;
            if (b0 + 192 < n) nb = (b0 + 192 + lane < n) ? arr[b0 + 192 + lane] : Key::inf();
            take(vb);
        }
    }
    return best;
}

// The r-th smallest (1-based, r <= n) distance key of arr[0..n), radix select on 8-bit digits
// from the top; hist: 256 u32 of this wave's LDS.
template <class Key>
__device__ uint64_t select_dk(const Key *arr, int n, int r, uint32_t *hist) {
    const int lane = lane_id();
    uint64_t prefix = 0, pmask = 0;
    uint32_t need = (uint32_t)r;
    for (int shift = Key::DKBITS - 8; shift >= 0; shift -= 8) {
        for (int b = lane; b < 256; b += 64) hist[b] = 0u;
        lds_fence();
        for (int i = lane; i < n; i += 64) {
            const uint64_t d = arr[i].dk();
            if ((d & pmask) == prefix) atomicAdd(&hist[(uint32_t)(d >> shift) & 255u], 1u);
        }
        lds_fence();
        // lane l holds digits 4l..4l+3; an inclusive scan over the lanes finds the digit
        // where the count of smaller entries reaches `need`
        const uint32_t c0 = hist[4 * lane], c1 = hist[4 * lane + 1], c2 = hist[4 * lane + 2],
                       c3 = hist[4 * lane + 3];
        const uint32_t sum = c0 + c1 + c2 + c3;
        uint32_t incl = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = (uint32_t)__shfl_up((int)incl, o);
            if (lane >= o) incl += t;
        }
        const uint32_t excl = incl - sum;
        const int L = __builtin_ctzll(ballot(excl < need && incl >= need));
        uint32_t below = rl32(excl, L);
        int digit = 4 * L;
        const uint32_t d0 = rl32(c0, L), d1 = rl32(c1, L), d2 = rl32(c2, L);
        if (below + d0 < need) {
            below += d0;
            ++digit;
            if (below + d1 < need) {
                below += d1;
                ++digit;
                if (below + d2 < need) {
                    below += d2;
                    ++digit;
                }
            }
        }
        need -= below;
        prefix |= (uint64_t)digit << shift;
        pmask |= 255ull << shift;
        lds_fence();
    }
    return prefix;
}

// In-place compactions of arr[0..n) (order not kept); return the new count.  Entries only
// move left: the writes of block j land below the start of block j + 1, so they never touch
// the two blocks whose loads are in flight (issued two blocks ahead into two register slots,
// the loop unrolled by two), and block j's own slots are written only after its load is used.
template <class Key, class Keep>
__device__ __forceinline__ int compact(Key *arr, int n, Keep keep_fn) {
    const int lane = lane_id();
    int out = 0;
    Key na = (lane < n) ? arr[lane] : Key::inf();
    Key nb = (64 + lane < n) ? arr[64 + lane] : Key::inf();
    const auto put = [&](int b0, const Key &e) {
        const bool keep = b0 + lane < n && keep_fn(e);
        const uint64_t km = ballot(keep);
        if (keep) arr[out + lanes_below(km)] = e;
        out += __popcll(km);
    };
    for (int b0 = 0; b0 < n; b0 += 128) {
        const Key ea = na;
        if (b0 + 128 < n) na = (b0 + 128 + lane < n) ? arr[b0 + 128 + lane] : Key::inf();
        put(b0, ea);
        if (b0 + 64 < n) {
            const Key eb = nb;
            if (b0 + 192 < n) nb = (b0 + 192 + lane < n) ? arr[b0 + 192 + lane] : Key::inf();
            put(b0 + 64, eb);
        }
    }
    return out;
}
template <class Key>
__device__ __forceinline__ int keep_dk_upto(Key *arr, int n, uint64_t dmax) {
    return compact(arr, n, [&](const Key &e) { return e.dk() <= dmax; });
}
template <class Key>
__device__ __forceinline__ int keep_above(Key *arr, int n, const Key &kx) {
    return compact(arr, n, [&](const Key &e) { return lt(kx, e); });
}

// shapes: (lanes per candidate row, 16-byte units per lane) by units per row
template <int SHAPE> struct Shape;
template <> struct Shape<0> { static constexpr int SUB = 4, NV = 2; };    // <= 8 units
template <> struct Shape<1> { static constexpr int SUB = 4, NV = 8; };    // <= 32 units
template <> struct Shape<2> { static constexpr int SUB = 4, NV = 16; };   // <= 64 units
template <> struct Shape<3> { static constexpr int SUB = 16, NV = 16; };  // <= 256 units
template <> struct Shape<4> { static constexpr int SUB = 64, NV = 16; };  // wider rows, chunked

// ---------------------------------------------------------------- two waves per query

// Two wavefronts per query (search_duo): the search wave runs search_query; the helper wave
// reads the search wave's predicted next expansion (the queue head it prefetches), and for that
// vertex's neighbours that are not yet visited computes the distances into a small LDS cache
// keyed by id, so that when the search wave expands that vertex its candidates' distances are
// usually there and the row gather leaves its dependent chain.  The helper only reads the
// search wave's visited table and never changes its state; a cache entry is the search key
// itself ((distance bits << 32) | id, one 64-bit LDS store), computed by the same code as the
// search wave's own, so results are unchanged whether or not an entry is used.
//
// Queries are separated by an epoch: at a new query the search wave bumps the epoch, waits
// (bounded) for the helper's acknowledgement -- given between two of its work units, after
// every store of the previous unit -- and only then clears and uses the cache; without the
// acknowledgement in time it searches that query without the cache.  The helper leaves when the
// search wave sets `done` after its last query.
struct DuoMail {
    uint32_t epoch, ack, tseq, tepoch, target, done, lenv_bits, pad[9];
};
constexpr int DUO_CACHE = 256;                 // cache entries (8 bytes each)
constexpr int DUO_ACK_SPINS = 4096;            // the search wave's bounded wait for the acknowledgement
constexpr int DUO_BYTES = (int)sizeof(DuoMail) + DUO_CACHE * 8 + 64 * 4;  // mailbox, cache, helper ids
__device__ __forceinline__ uint32_t duo_slot(uint32_t id) { return (id * 0x9E3779B1u) >> 24; }
__device__ __forceinline__ uint32_t lds_ld(const uint32_t *p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
__device__ __forceinline__ void lds_st(uint32_t *p, uint32_t v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// ---------------------------------------------------------------- the search kernel

// One query on one wave: qi is the query, slot_q the wave's spill table, smem its LDS.
// BAGL: the queue's bag sits in the wave's LDS slice after the query (qcap keys) instead of in
// global memory -- for latency-bound launches with few waves per CU, where a refill or a
// truncation then scans LDS instead of waiting on global-memory round trips.
// HELP: the search wave of search_duo (float32 keys); duo = the mailbox and cache in LDS.
// BG: the best list's front in the slot's global region after the bag (float32 rows, bag in
// global memory); the launch sizes each slot's global region for it (bag_bytes + pq keys).
template <int METRIC, bool EXACT, int STORE, int SHAPE, bool BAGL = false, bool HELP = false, bool BG = false>
__device__ __forceinline__ void search_query(const SearchParams &P, const int qi, const int slot_q,
                                             unsigned char *smem, unsigned char *duo = nullptr) {
    using Key = typename KeySel<EXACT>::T;
    constexpr int SUB = Shape<SHAPE>::SUB, NV = Shape<SHAPE>::NV;
    constexpr int QH = 64;  // head entries: one per lane, ascending from lane h
    const int lane = lane_id();
    lds_fence();
    // per-query budgets are clamped to the launch's LDS sizing (pq) and output rows (out_ld)
    const int pq = P.pq_arr ? max(1, min(P.pq_arr[qi], P.pq)) : P.pq;
    const int k = max(0, min(min(P.k_arr ? P.k_arr[qi] : P.k, pq), P.out_ld));
    const uint32_t H = 1u << P.hlog;
    const uint32_t hmask = H - 1u;
    const uint32_t vis_limit = (H >> 2) * 3u;

    uint32_t *vis = reinterpret_cast<uint32_t *>(smem);
    const bool bloom = P.bloom_k > 0;
    const uint32_t vis_bytes = bloom ? (4u << P.hlog) : H * (P.vis16 ? 2u : 4u);
    const int gbits = P.hlog - 3;                               // compact table: groups of 8
    Key *S = reinterpret_cast<Key *>(smem + vis_bytes);        // merge scratch, QH keys
    // the best list's front: LDS after the scratch, or (bglob: float32 rows, global bag) the
    // slot's global region after its bag
    static_assert(!BG || (STORE == ST_F32 && !BAGL), "the global front needs float32 rows and a global bag");
    constexpr bool bglob = BG;
    Key *B = S + QH;
    // the front (computed where used, in rare paths, so that no registers hold a global
    // pointer through the loop)
    const auto Bf = [&]() -> Key * {
        if constexpr (BG)
            return reinterpret_cast<Key *>(P.gbag + (size_t)slot_q * P.gbag_stride + bag_bytes<Key>(P.qcap));
        else
            return B;
    };
    // global front: its stores are complete before it is read
    const auto front_fence = [&]() {
        if constexpr (BG) bag_fence();
    };
    uint32_t *cid = reinterpret_cast<uint32_t *>(B + (bglob ? 0 : ((P.pq + 1) & ~1)));  // 256 u32
    float *qv = reinterpret_cast<float *>(cid + 256);
    // bag, qcap keys: global memory (one region per slot), or this wave's LDS after the query
    Key *Bg = BAGL ? reinterpret_cast<Key *>(qv + P.qld)
                   : reinterpret_cast<Key *>(P.gbag + (size_t)slot_q * P.gbag_stride);
    // the layout must fit the slice and the bag the launch allocated: otherwise flag the query
    // and touch neither (a host / kernel mismatch becomes an error, not out-of-range accesses)
    const unsigned char *lds_end = BAGL ? reinterpret_cast<unsigned char *>(Bg + P.qcap)
                                        : reinterpret_cast<unsigned char *>(qv + P.qld);
    const size_t gneed = bglob ? bag_bytes<Key>(P.qcap) + (size_t)P.pq * sizeof(Key) : (size_t)P.qcap * sizeof(Key);
    if (lds_end - smem > (ptrdiff_t)P.wave_lds ||
        (!BAGL && (!P.gbag || gneed > P.gbag_stride))) {
        if (lane == 0 && P.out_flags) P.out_flags[qi] = F_BAD_LAUNCH;
        return;
    }

    for (uint32_t i = lane; i < (vis_bytes >> 4); i += 64) reinterpret_cast<uint4 *>(vis)[i] = make_uint4(0u, 0u, 0u, 0u);
    for (int i = lane; i < P.qld; i += 64) qv[i] = (i < P.dim) ? P.queries[(size_t)qi * P.dim + i] : 0.0f;
    lds_fence();

    float lenv = 0.0f;
    if (METRIC == M_COS) {
        // query norm in the CPU's order: lenv = fmaf(v[i], v[i], lenv) (data.h:55 under FMA contraction)
        if (lane == 0)
            for (int i = 0; i < P.dim; ++i) lenv = fmaf(qv[i], qv[i], lenv);
        lenv = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(lenv)));
    }

    bool use_cache = false;
    uint32_t epoch = 0u, cache_hits = 0u;
    DuoMail *mail = nullptr;
    uint64_t *dcache = nullptr;
    if constexpr (HELP) {
        static_assert(!EXACT, "the helper's cache holds float32 keys");
        mail = reinterpret_cast<DuoMail *>(duo);
        dcache = reinterpret_cast<uint64_t *>(duo + sizeof(DuoMail));
        // new query (its vector and norm are in place): bump the epoch, wait for the helper to
        // acknowledge it, then clear the cache
        epoch = uni(lds_ld(&mail->epoch)) + 1u;
        if (lane == 0) {
            lds_st(&mail->lenv_bits, __float_as_uint(lenv));
            lds_st(&mail->epoch, epoch);
        }
        for (int s = 0; s < DUO_ACK_SPINS && !use_cache; ++s) {
            use_cache = uni(lds_ld(&mail->ack)) == epoch;
            if (!use_cache) __builtin_amdgcn_s_sleep(1);
        }
        if (use_cache)
            for (int i = lane; i < DUO_CACHE; i += 64) dcache[i] = ~0ull;
        lds_fence();
    }
    // the predicted next expansion, for the helper
    const auto publish = [&](uint32_t v) {
        if constexpr (HELP) {
            if (use_cache && lane == 0) {
                mail->target = v;
                mail->tepoch = epoch;
                lds_st(&mail->tseq, mail->tseq + 1u);
            }
        }
    };

    uint32_t flags = 0u;
    int steps = 0;
    uint32_t nvis = 0u;
    // statistics counters live in lane 0 of a vector register (read once at the end), which
    // keeps them out of the scalar registers the loop is short of
    uint32_t adj_words = 0u, pf_hits = 0u, refills = 0u, truncs = 0u;
    const uint32_t l0 = lane == 0 ? 1u : 0u;
    // visited spill: once the LDS table nears its load limit, new ids go to a global table
    bool spill = false;
    uint32_t nglob = 0u;
    uint32_t *gv = nullptr;
    const uint32_t GH = 1u << P.ghlog, gmask = GH - 1u, glimit = (GH >> 2) * 3u;
    int h = 0, nq = 0, nbest = 0;
    uint64_t bmax = ~0ull;  // distance key of best[pq-1] once the list is full
    // the best list's last positions [tb, tb + 64) live in registers: lane j holds position
    // tb + j (+inf past nbest); the front [0, tb) is in LDS.  tb = max(0, nbest - 64).
    Key bt = Key::inf();
    int tb = 0;
    // the queue head lives in registers: lanes [h, h + nq) hold its keys in ascending order
    Key hq = Key::inf();

    // start vertices (graph.h:175-181: vertex 0; the learned form passes its own), each
    // entered once through the visited set and placed in the queue in (dist, id) order
    {
        // entries >= n (0xFFFFFFFF pads shorter lists) are skipped; none left means vertex 0
        const int ne = P.entry ? P.n_entry : 1;
        uint32_t sid = (lane < ne) ? (P.entry ? P.entry[(size_t)qi * ne + lane] : 0u) : 0xFFFFFFFFu;
        bool use = lane < ne && (uint64_t)sid < P.n;
        if (ballot(use) == 0ull) {
            use = lane == 0;
            sid = 0u;
        }
        int ins = 0;
        if (use)
            ins = bloom ? (bloom_insert(vis, P.hlog, P.bloom_k, sid) ? 1 : 0)
                        : P.vis16 ? vis16_insert(vis, gbits, P.idbits, sid) : (vis_insert(vis, hmask, P.hlog, sid) ? 1 : 0);
        // no room for a start vertex (up to 64 in an empty table): the query is flagged and not
        // searched; the epilogue reports it for a re-run with a larger table
        if (ballot(ins == 2) != 0ull) flags |= F_VIS_OVF;
        const bool fresh = ins == 1;
        const uint64_t fm = ballot(fresh);
        const int m0 = __popcll(fm);
        if (fresh) cid[lanes_below(fm)] = sid;
        nvis = (uint32_t)m0;
        lds_fence();
        const uint32_t cand0 = (lane < m0) ? cid[lane] : 0u;
        const Key k0 = eval_candidates<METRIC, EXACT, STORE, SUB, NV, Key>(P, m0, cand0, cid, qv, lenv);
        int rank = 0;
        for (int j = 0; j < m0; ++j) rank += lt(k0.rl(j), k0) ? 1 : 0;
        if (lane < m0) S[rank] = k0;
        lds_fence();
        hq = (lane < m0) ? S[lane] : Key::inf();
        nq = m0;
        lds_fence();
    }
    const int max_steps = P.steps_arr ? min(P.steps_arr[qi], P.max_steps) : P.max_steps;
    int settle = 0;  // lane 0 holds it

    // adjacency rows loaded one expansion ahead, in two slots used in turn: an expansion of
    // parity X reads its row from slot X where it lies and loads the next head's row into slot
    // Y, whose row the previous expansion has consumed.  No row is copied between registers, so
    // nothing waits for a load before the row is first used (a copy of a register whose load is
    // in flight waits for that load).  The loop body is instantiated once per parity.
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    uint32_t pid[2] = {NONE, NONE}, prow[2] = {0u, 0u};
    // all lanes load (lanes past the row read its last word and are never used: a row's
    // degree is below gstride), so the load needs no lane mask and no zero fill
    const auto adj_row = [&](uint32_t v) -> uint32_t {
        return P.graph[(size_t)v * (size_t)P.gstride + (uint32_t)min(lane, P.gstride - 1)];
    };

    int bn = 0;          // bag entries: unsorted, all above T
    Key T = Key::inf();  // head entries <= T < bag entries (+inf while the bag is empty)
    // one expansion; returns true to leave the loop
    const auto expand = [&](auto parity) -> bool {
        constexpr int X = decltype(parity)::value, Y = 1 - X;
        if (nq == 0 && bn == 0) return true;
        if (steps >= max_steps) {
            flags |= F_STEP_CAP;
            return true;
        }
        if (nq == 0) {
            // refill the head with the (up to) 64 smallest bag entries
            const int c = min(QH, bn);
            refills += l0;
            if constexpr (BAGL) lds_fence(); else bag_fence();
            const Key sel = select64(Bg, bn);
            hq = sel;  // +inf in lanes >= c
            const Key top = sel.rl(c - 1);
            h = 0;
            nq = c;
            bn = keep_above(Bg, bn, top);
            T = bn > 0 ? top : Key::inf();
            pid[X] = NONE;
            if constexpr (!BAGL) vm_drain();
        }
        const Key now = hq.rl(h);
        // stop rule, graph.h:188: |best| == pq and best.max.d < now.d (distance only, strict)
        if (nbest == pq && bmax < now.dk()) return true;
        ++steps;
        ++h;
        --nq;

        // adjacency row of `now`: slot X holds it when the prefetch predicted this vertex;
        // otherwise it is loaded into slot X now
        if (pid[X] == now.id()) {
            pf_hits += l0;
        } else {
            pid[X] = now.id();
            prow[X] = adj_row(now.id());
        }
        const uint32_t r = prow[X];
        // prefetch the row of the new queue head, the likely next expansion, into slot Y.  The
        // load is issued on every expansion (vertex 0's row when the head is empty) so that a
        // load always follows this expansion's row: the wait before the row's first use then
        // leaves the prefetch in flight.
        pid[Y] = nq > 0 ? hq.rl(h).id() : NONE;
        prow[Y] = adj_row(nq > 0 ? pid[Y] : 0u);
        if (nq > 0) publish(pid[Y]);

        // insert `now` into the sorted pq-best list (push + pop-max of graph.h:193-195).
        // Expansions come out in nearly ascending order, so `now` almost always lands in the
        // register tail: one compare and a lane shift, no LDS round trip.
        {
            const uint64_t below = ballot(lt(bt, now));
            const bool in_tail = tb == 0 || (below & 1ull);
            const int p = in_tail ? tb + __popcll(below)
                                  : (front_fence(), prefix_count(Bf(), tb, [&](const Key &e) { return lt(e, now); }));
            if (p < k && lane == 0) settle = steps;
            if (p < pq) {
                const int j = p - tb;  // tail lane of the new entry; negative: it lands in the front
                if (nbest >= 64 && nbest < pq) {
                    // growing past 64 entries: tb advances by one and position tb joins the front
                    if (j >= 0) {
                        // the 65 entries [tb, nbest] with `now` at j: the first goes to the front,
                        // lanes below j - 1 take their upper neighbour's entry
                        if (lane == 0) Bf()[tb] = ksel(j == 0, now, bt);
                        const Key up = shl1(bt);
                        bt = ksel(lane < j - 1, up, ksel(lane == j - 1, now, bt));
                    } else {
                        // front [p, tb) moves to [p + 1, tb + 1); the tail keeps its entries
                        shift_insert(Bf(), p, tb, now);  // (a front in global memory: after the fence above)
                    }
                    ++tb;
                } else if (j >= 0) {
                    // below 64 entries (tb = 0), or full (the entry at pq - 1 leaves)
                    const Key dn = shr1(bt);
                    bt = ksel(lane < j, bt, ksel(lane == j, now, dn));
                } else {
                    // full, `now` in the front: the front's last entry moves into the tail
                    const Key fl = Bf()[tb - 1];
                    shift_insert(Bf(), p, tb - 1, now);
                    const Key dn = shr1(bt);
                    bt = ksel(lane == 0, fl, dn);
                }
                nbest = min(nbest + 1, pq);
                bt = ksel(lane >= nbest - tb, Key::inf(), bt);
                if (nbest == pq) bmax = bt.rl(pq - 1 - tb).dk();
                lds_fence();
            }
        }

        // neighbours of `now` in row order (graph.h:196-205)
        const int deg = (int)rl32(r, 0);
        adj_words += l0 * (1u + (uint32_t)deg);
        if (!bloom && !spill && nvis + 64u > vis_limit) {
            if (!P.gvis) {
                flags |= F_VIS_OVF;
                return true;
            }
            // switch to the global table for the rest of this query: clear it first
            spill = true;
            gv = P.gvis + (size_t)slot_q * GH;
            for (uint32_t i = lane; i < GH; i += 64) gv[i] = 0u;
            __threadfence_block();
        }
        bool fresh = false, vovf = false;
        if (lane >= 1 && lane <= deg) {
            if (bloom) {
                fresh = bloom_insert(vis, P.hlog, P.bloom_k, r);
            } else if (!spill) {
                if (P.vis16) {
                    const int ins = vis16_insert(vis, gbits, P.idbits, r);
                    fresh = ins == 1;
                    vovf = ins == 2;
                } else {
                    fresh = vis_insert(vis, hmask, P.hlog, r);
                }
            } else {
                const bool in_lds = P.vis16 ? vis16_test(vis, gbits, P.idbits, r) : vis_test(vis, hmask, P.hlog, r);
                fresh = !in_lds && gvis_insert(gv, gmask, P.ghlog, r);
            }
        }
        if (!bloom && P.vis16 && ballot(vovf) != 0ull) {
            // the compact table had no room near these ids: continue in the global table
            if (!P.gvis) {
                flags |= F_VIS_OVF;
                return true;
            }
            if (!spill) {
                spill = true;
                gv = P.gvis + (size_t)slot_q * GH;
                for (uint32_t i = lane; i < GH; i += 64) gv[i] = 0u;
                __threadfence_block();
            }
            if (vovf) fresh = gvis_insert(gv, gmask, P.ghlog, r);
        }
        const uint64_t fm = ballot(fresh);
        const int m = __popcll(fm);
        if (fresh) cid[lanes_below(fm)] = r;
        nvis += (uint32_t)m;
        if (spill) {
            nglob += (uint32_t)m;
            if (nglob + 64u > glimit) {
                flags |= F_VIS_OVF;
                return true;
            }
        }
        if (m == 0) return false;
        lds_fence();
        const uint32_t cand = (lane < m) ? cid[lane] : 0u;
        Key mine;
        bool cached = false;
        if constexpr (HELP) {
            if (use_cache) {
                // every candidate's distance in the helper's cache: no gather
                const uint64_t ce = (lane < m) ? dcache[duo_slot(cand)] : ~0ull;
                const bool hit = lane < m && (uint32_t)ce == cand;
                if (ballot(lane < m && !hit) == 0ull) {
                    mine = KeyF{(lane < m) ? ce : ~0ull};
                    cached = true;
                    cache_hits += l0;
                }
            }
        }
        if (!cached) mine = eval_candidates<METRIC, EXACT, STORE, SUB, NV, Key>(P, m, cand, cid, qv, lenv);

        // ---- the candidates: at most T into the sorted head, above T into the bag
        const uint64_t cut = (nbest == pq) ? bmax : ~0ull;  // drop distance keys above cut
        const bool surv = lane < m && mine.dk() <= cut;
        if (ballot(surv) == 0ull) return false;
        if (bn + 62 > P.qcap) {
            // make room: drop bag entries beyond the pq-th smallest distance of head + bag
            // (ties kept) or beyond bmax; none of them can be expanded before the stop rule
            uint64_t dcut = (nbest == pq) ? bmax : ~0ull;
            truncs += l0;
            if constexpr (BAGL) lds_fence(); else bag_fence();
            if (pq <= nq) dcut = min(dcut, hq.rl(h + pq - 1).dk());
            else if (pq - nq <= bn) dcut = min(dcut, select_dk(Bg, bn, pq - nq, cid));
            bn = keep_dk_upto(Bg, bn, dcut);
            if constexpr (!BAGL) vm_drain();
            if (bn == 0) T = Key::inf();
            if (bn + 62 > P.qcap) {  // the tie group at the pq-th distance exceeds the slack
                flags |= F_Q_OVF;
                return true;
            }
        }
        const bool to_b = surv && lt(T, mine);
        const bool to_q = surv && !to_b;
        const uint64_t bm = ballot(to_b);
        if (to_b) Bg[bn + lanes_below(bm)] = mine;
        bn += __popcll(bm);
        const uint64_t sm = ballot(to_q);
        const int ms = __popcll(sm);
        if (ms > 0) {
            // merge into the register head.  A head key moves down by the h consumed lanes and
            // up by the survivors below it; a survivor lands after the head keys and survivors
            // below it.  Keys are distinct, so the destinations are a permutation of
            // [0, nq + ms); those past the last lane go to the bag (they exceed every key that
            // stays), the rest pass through the LDS scratch.
            const bool live = lane >= h && lane < h + nq;
            int cnt = 0, sdst = 0;
            for (uint64_t bits = sm; bits; bits &= bits - 1) {
                const int j = __builtin_ctzll(bits);
                const Key o = mine.rl(j);
                cnt += (live && lt(o, hq)) ? 1 : 0;
                const int d = __popcll(ballot(live && lt(hq, o))) + __popcll(ballot(to_q && lt(mine, o)));
                if (lane == j) sdst = d;
            }
            const int hdst = lane - h + cnt;
            const int nn = nq + ms;
            if (nn > QH) {
                if (live && hdst >= QH) Bg[bn + hdst - QH] = hq;
                if (to_q && sdst >= QH) Bg[bn + sdst - QH] = mine;
            }
            if (live && hdst < QH) S[hdst] = hq;
            if (to_q && sdst < QH) S[sdst] = mine;
            lds_fence();
            hq = (lane < nn) ? S[lane] : Key::inf();
            h = 0;
            if (nn > QH) {
                // the head keeps its QH smallest; T drops to its new largest key
                bn += nn - QH;
                nq = QH;
                T = hq.rl(QH - 1);
            } else {
                nq = nn;
            }
            lds_fence();
        }
        // the queue head may have changed: refresh the prefetch target only if it did
        if (nq > 0 && hq.rl(h).id() != pid[Y]) {
            pid[Y] = hq.rl(h).id();
            prow[Y] = adj_row(pid[Y]);
            publish(pid[Y]);
        }
        return false;
    };
    for (;;) {
        if (flags != 0u || expand(std::integral_constant<int, 0>{})) break;
        if (expand(std::integral_constant<int, 1>{})) break;
    }
    // the whole best list in LDS (or, bglob, in the slot's global region) from here on
    if constexpr (BG) B = Bf();
    if (lane < nbest - tb) B[tb + lane] = bt;
    if constexpr (BG) bag_fence();
    else lds_fence();

    const int cnt = min(k, nbest);
    const size_t ob = (size_t)qi * (size_t)P.out_ld;
    bool done = false;
    if constexpr (STORE != ST_F32 && !EXACT) {
        if (P.rerank) {
            // exact FP32 distances for the whole pq-best list, then the k smallest in order
            for (int base = 0; base < nbest; base += 64) {
                const int c = min(64, nbest - base);
                if (lane < c) cid[lane] = B[base + lane].id();
                lds_fence();
                const uint32_t cand = (lane < c) ? cid[lane] : 0u;
                const Key ex = eval_candidates<METRIC, false, ST_F32, 4, 8, Key>(P, c, cand, cid, qv, lenv);
                if (lane < c) B[base + lane] = ex;
                lds_fence();
            }
            for (int r = 0; r < k; ++r) {
                Key best = Key::inf();
                int bi = -1;
                for (int i = lane; i < nbest; i += 64) {
                    const Key e = B[i];
                    if (lt(e, best)) {
                        best = e;
                        bi = i;
                    }
                }
                const Key w = wave_min(best);
                if (bi >= 0 && best.v == w.v) B[bi] = Key::inf();
                lds_fence();
                if (lane == 0) {
                    const bool ok = r < cnt;
                    P.out_ids[ob + r] = ok ? w.id() : 0xFFFFFFFFu;
                    if (P.out_dist) P.out_dist[ob + r] = ok ? w.dist() : __int_as_float(0x7f800000);
                }
            }
            done = true;
        }
    }
    // output: first k of the ascending list (graph.h:209-215)
    for (int i = lane; i < k && !done; i += 64) {
        const bool ok = i < cnt;
        const Key e = ok ? B[i] : Key::inf();
        P.out_ids[ob + i] = ok ? e.id() : 0xFFFFFFFFu;
        if (P.out_dist) P.out_dist[ob + i] = ok ? e.dist() : __int_as_float(0x7f800000);
    }
    if (lane == 0) {
        if (P.retry_list && (flags & (F_VIS_OVF | F_Q_OVF))) P.retry_list[atomicAdd(P.retry_count, 1u)] = (uint32_t)qi;
        if (P.out_flags) P.out_flags[qi] = flags;
        if (P.out_stats) {
            uint32_t *o = P.out_stats + 8 * (size_t)qi;
            o[0] = (uint32_t)steps;
            o[1] = nvis;
            o[2] = adj_words;
            o[3] = pf_hits;
            o[4] = (uint32_t)settle;
            o[5] = refills;
            o[6] = truncs;
            o[7] = cache_hits;
        }
    }
}

// One workgroup (one wave) per query of the launch.
template <int METRIC, bool EXACT, int STORE, int SHAPE, bool BG = false>
__global__ __launch_bounds__(64) void search_kernel(SearchParams P) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int slot_q = (int)blockIdx.x;
    const int qi = P.qidx ? (int)P.qidx[P.qbase + slot_q] : P.qbase + slot_q;
    search_query<METRIC, EXACT, STORE, SHAPE, false, false, BG>(P, qi, slot_q, smem);
}

// Persistent grid: every wave takes queries from the shared list until it is exhausted
// (each wave reaches the exit once *work_next passes *work_count).  A workgroup holds
// blockDim.x / 64 waves, each on its own wave_lds-byte slice of LDS and its own spill table.
template <int METRIC, bool EXACT, int STORE, int SHAPE, bool BAGL = false, bool BG = false>
__global__ __launch_bounds__(256) void search_persistent(SearchParams P) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int w = (int)(threadIdx.x >> 6);
    const int slot = (int)blockIdx.x * (int)(blockDim.x >> 6) + w;
    unsigned char *mine = smem + (size_t)w * P.wave_lds;
    const uint32_t count = *P.work_count;
    // a tier whose searches are the longest runs at raised issue priority, so that waves of
    // concurrently running tiers on the same SIMDs do not stretch its critical path
    if (P.wave_prio) __builtin_amdgcn_s_setprio(2);
    for (;;) {
        // Every lane issues the add (lane 0 adds 1, the others 0) and lane 0's old value is
        // broadcast.  No lane-0 branch here: the query's lane-0 epilogue ends the previous
        // iteration, and a second branch on the same predicate lets the compiler thread the
        // back edge so that lanes 1..63 start the next iteration without lane 0.
        const uint32_t v = atomicAdd(P.work_next, lane_id() == 0 ? 1u : 0u);
        const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
        if (t >= count) break;
        const int qi = P.qidx ? (int)P.qidx[t] : (int)t;
        search_query<METRIC, EXACT, STORE, SHAPE, BAGL, false, BG>(P, qi, slot, mine);
    }
}

// The helper wave of search_duo (see DuoMail): slice = the search wave's LDS slice.
template <int METRIC, int STORE, int SHAPE, bool BG>
__device__ void duo_helper(const SearchParams &P, unsigned char *slice, unsigned char *duo) {
    constexpr int SUB = Shape<SHAPE>::SUB, NV = Shape<SHAPE>::NV;
    const int lane = lane_id();
    DuoMail *mail = reinterpret_cast<DuoMail *>(duo);
    uint64_t *dcache = reinterpret_cast<uint64_t *>(duo + sizeof(DuoMail));
    uint32_t *hid = reinterpret_cast<uint32_t *>(dcache + DUO_CACHE);
    // the search wave's layout (search_query): visited | 64 keys | pq keys (not with bglob) | 256 u32 | qv
    const uint32_t H = 1u << P.hlog;
    const uint32_t vis_bytes = H * (P.vis16 ? 2u : 4u);
    const uint32_t *vis = reinterpret_cast<const uint32_t *>(slice);
    const KeyF *S = reinterpret_cast<const KeyF *>(slice + vis_bytes);
    const uint32_t *cid = reinterpret_cast<const uint32_t *>(S + 64 + (BG ? 0 : ((P.pq + 1) & ~1)));
    const float *qv = reinterpret_cast<const float *>(cid + 256);
    uint32_t my_epoch = 0u, last_t = 0u;
    float lenv = 0.0f;
    for (;;) {
        uint32_t t = 0u;
        for (;;) {
            if (uni(lds_ld(&mail->done))) return;
            const uint32_t e = uni(lds_ld(&mail->epoch));
            if (e != my_epoch) {
                // everything this wave stored for the previous epoch is written: acknowledge
                my_epoch = e;
                lenv = __uint_as_float(uni(lds_ld(&mail->lenv_bits)));
                lds_fence();
                if (lane == 0) lds_st(&mail->ack, e);
            }
            t = uni(lds_ld(&mail->tseq));
            if (t != last_t) break;
            __builtin_amdgcn_s_sleep(2);
        }
        last_t = t;
        const uint32_t v = uni(mail->target), te = uni(mail->tepoch);
        if (te != my_epoch || (uint64_t)v >= P.n) continue;
        const uint32_t r = P.graph[(size_t)v * (size_t)P.gstride + (uint32_t)min(lane, P.gstride - 1)];
        const int deg = (int)rl32(r, 0);
        bool want = false;
        if (lane >= 1 && lane <= deg)
            want = !(P.vis16 ? vis16_test(vis, P.hlog - 3, P.idbits, r) : vis_test(vis, H - 1u, P.hlog, r));
        const uint64_t fm = ballot(want);
        const int m = __popcll(fm);
        if (m == 0) continue;
        if (want) hid[lanes_below(fm)] = r;
        lds_fence();
        const uint32_t cand = (lane < m) ? hid[lane] : 0u;
        const KeyF k = eval_candidates<METRIC, false, STORE, SUB, NV, KeyF>(P, m, cand, hid, qv, lenv);
        if (lane < m) dcache[duo_slot(cand)] = k.v;
        lds_fence();
    }
}

// Two waves per query, one query per workgroup at a time: wave 0 searches (the persistent work
// loop of search_persistent), wave 1 helps (duo_helper).  LDS: the search wave's wave_lds slice,
// then DUO_BYTES of mailbox, cache and helper ids.
template <int METRIC, int STORE, int SHAPE, bool BG = false>
__global__ __launch_bounds__(128) void search_duo(SearchParams P) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned char *duo = smem + P.wave_lds;
    if (threadIdx.x < sizeof(DuoMail) / 4) reinterpret_cast<uint32_t *>(duo)[threadIdx.x] = 0u;
    __syncthreads();
    // both waves of a prioritised tier: the helper's distances are on the search wave's path
    if (P.wave_prio) __builtin_amdgcn_s_setprio(2);
    if (threadIdx.x >= 64) {
        duo_helper<METRIC, STORE, SHAPE, BG>(P, smem, duo);
        return;
    }
    const uint32_t count = *P.work_count;
    for (;;) {
        // every lane issues the add (see search_persistent)
        const uint32_t v = atomicAdd(P.work_next, lane_id() == 0 ? 1u : 0u);
        const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
        if (t >= count) break;
        const int qi = P.qidx ? (int)P.qidx[t] : (int)t;
        search_query<METRIC, false, STORE, SHAPE, false, true, BG>(P, qi, (int)blockIdx.x, smem, duo);
    }
    lds_fence();
    if (lane_id() == 0) lds_st(&reinterpret_cast<DuoMail *>(duo)->done, 1u);
}

}  // namespace song
