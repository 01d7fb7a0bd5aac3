// Several queries per 64-lane wavefront: the reference's N_MULTIQUERY
// (warp_astar_accelerator.h:27-30, 48-49: query cid = tid / CRITICAL_STEP), here with
// sub-groups of L = 16 or 32 lanes, each running one query's search on its own, with
// the same semantics as search_query (graph.h:169-216: (dist, id) order, strict-< stop rule
// on distance, every discovered neighbour visited, ascending output), so the ids equal the
// one-query-per-wave kernel's.
//
// For small rows (few 16-byte units) most of a 64-lane wave has nothing to do in the
// one-query kernel; here the wave's instructions serve G = 64 / L queries at once.  The
// structures are built for sub-groups: the queue is one sorted array in LDS (ascending,
// popped from a head index, new keys placed by a 16/32-ary rank search and a block shift,
// bounded by the tie-safe cut at the pq-th distance), the best list a sorted LDS array,
// the visited set the same exact LDS table with its global spill, the adjacency row read
// as L-word slices and prefetched one expansion ahead.  All control flow is uniform within
// a sub-group; sub-groups diverge freely (a query that ends takes the next from the work
// counter while the others continue).
//
// Per-query LDS slice (bytes, in order; group_lds_bytes on the host must match):
//   visited : H slots (compact 16-bit or 32-bit, as search_query)
//   queue   : QC = pq + slack + 2L keys, live entries [h, h + nq)
//   best    : pq keys (rounded to even), first nbest live
//   stage   : 64 keys   candidate keys of one expansion / the start vertices
//   cid     : 64 u32    candidate ids
//   qv      : qld f32   the query, zero-padded to a multiple of 16
#pragma once

#include "search_kernel.hpp"

namespace song {

template <int L> struct Sub {
    static constexpr uint32_t MASK = L >= 32 ? 0xFFFFFFFFu : ((1u << L) - 1u);
    int lane, sl, sbase;
    __device__ Sub() : lane(lane_id()), sl(lane_id() % L), sbase(lane_id() - lane_id() % L) {}
    // this sub-group's predicate bits, lane sbase at bit 0 (uniform within the sub-group)
    __device__ uint32_t ballot_(bool p) const { return (uint32_t)(__ballot(p) >> sbase) & MASK; }
    __device__ int below(uint32_t m) const { return __popc(m & ((1u << sl) - 1u)); }
    __device__ uint32_t bcast(uint32_t v, int j) const { return (uint32_t)__shfl((int)v, sbase + j); }
    __device__ float bcastf(float v, int j) const { return __shfl(v, sbase + j); }
};

template <class Key, int L, class Pred>
__device__ __forceinline__ int sub_prefix_count(const Sub<L> &S, const Key *arr, int n, Pred below) {
    int lo = 0, len = n;
    while (len > L) {
        const int stride = (len + L - 1) / L;
        const int end = lo + len;
        const int idx = min(lo + (S.sl + 1) * stride, end) - 1;
        const bool v = lo + S.sl * stride < end;
        const Key e = arr[v ? idx : lo];
        const int c = __popc(S.ballot_(v && below(e)));
        lo = min(lo + c * stride, end);
        len = min(stride, end - lo);
    }
    const bool v = S.sl < len;
    const Key e = arr[v ? lo + S.sl : lo];
    return lo + __popc(S.ballot_(v && below(e)));
}

// arr[p, e) -> arr[p + 1, e + 1), blocks of L from the end
template <class Key, int L>
__device__ __forceinline__ void sub_shift_right(const Sub<L> &S, Key *arr, int p, int e) {
    for (int blk = (e - 1) / L; e > p && blk >= p / L; --blk) {
        const int i = blk * L + S.sl;
        const bool live = i >= p && i < e;
        const Key v = arr[live ? i : p];
        lds_fence();
        if (live) arr[i + 1] = v;
        lds_fence();
    }
}

// distances of the m candidates cid[0..m) into stage[0..m): SUB lanes per row, L / SUB rows
// per pass; exact mode flags (stage[c].pad) the sums a lane must redo in the CPU's order
template <int METRIC, bool EXACT, int STORE, int SUB, int NVU, int L, class Key>
__device__ __forceinline__ void sub_eval(const Sub<L> &S, const SearchParams &P, int m, const uint32_t *cid,
                                         Key *stage, const float *qv, float lenv) {
    constexpr int EPU = Store<STORE>::EPU;
    constexpr int G2 = L / SUB;
    const int g2 = S.sl / SUB, s2 = S.sl % SUB;
    const int nu = (STORE == ST_F32) ? (P.ld >> 2) : (int)(P.lp_stride >> 4);
    for (int base = 0; base < m; base += G2) {
        const int c = base + g2;
        const bool valid = c < m;
        const uint32_t id = valid ? cid[c] : 0u;
        const uint4 *row = (STORE == ST_F32) ? reinterpret_cast<const uint4 *>(P.data + (size_t)id * (size_t)P.ld)
                                             : reinterpret_cast<const uint4 *>(P.data_lp + (size_t)id * P.lp_stride);
        const float scale = (STORE != ST_F32 && valid) ? P.scale[id] : 1.0f;
        float facc = 0.0f;
        double dacc = 0.0, sab = 0.0;
        int mn = 0x3fffffff;
        for (int j0 = 0; j0 < nu; j0 += NVU * SUB) {
            uint4 u[NVU];
            if (valid) {
#pragma unroll
                for (int v = 0; v < NVU; ++v) u[v] = row[min(j0 + s2 + v * SUB, nu - 1)];
            }
#pragma unroll
            for (int v = 0; v < NVU; ++v) {
                const int j = j0 + s2 + v * SUB;
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
        if constexpr (!EXACT) {
#pragma unroll
            for (int o = SUB / 2; o > 0; o >>= 1) facc += __shfl_xor(facc, o);
            if (valid && s2 == 0) {
                const float lena = (METRIC == M_COS) ? P.norms[id] : 0.0f;
                stage[c] = Key::make_f(finish_f32<METRIC>(facc, lena, lenv), id);
            }
        } else {
#pragma unroll
            for (int o = SUB / 2; o > 0; o >>= 1) {
                dacc += __shfl_xor(dacc, o);
                sab += __shfl_xor(sab, o);
                mn = min(mn, __shfl_xor(mn, o));
            }
            if (valid && s2 == 0) {
                const bool seq = !((sab == 0.0) || (mn < 0x3fffffff && sab < ldexp(1.0, mn + 53)));
                const float lena = (METRIC == M_COS) ? P.norms[id] : 0.0f;
                Key kk = Key::make_d(finish_exact<METRIC>(dacc, lena, lenv), id);
                kk.pad = seq ? 1u : 0u;
                stage[c] = kk;
            }
        }
    }
    lds_fence();
    if constexpr (EXACT) {
        // flagged sums again, one lane per candidate, sequential over the dimensions
        for (int c = S.sl; c < m; c += L) {
            const Key e = stage[c];
            if (e.pad) {
                const uint32_t id = e.id();
                const double s = seq_sum<METRIC>(P.data + (size_t)id * (size_t)P.ld, qv, P.dim);
                const float lena = (METRIC == M_COS) ? P.norms[id] : 0.0f;
                stage[c] = Key::make_d(finish_exact<METRIC>(s, lena, lenv), id);
            }
        }
        lds_fence();
    }
}

// One persistent grid; a workgroup holds blockDim.x / 64 waves of 64 / L queries each, every
// query on its own wave_lds-byte LDS slice (wave_lds = one query's slice here) and spill table.
template <int METRIC, bool EXACT, int STORE, int SHAPE, int L>
__global__ __launch_bounds__(256) void search_group_persistent(SearchParams P) {
    using Key = typename KeySel<EXACT>::T;
    constexpr int SUB = Shape<SHAPE>::SUB, NV = Shape<SHAPE>::NV;
    static_assert(SUB <= L, "a row's lanes must lie in one sub-group");
    constexpr int G = 64 / L;
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    extern __shared__ __align__(16) unsigned char smem[];
    const Sub<L> S;
    const int w = (int)(threadIdx.x >> 6);
    const int g = S.lane / L;
    const int slot = ((int)blockIdx.x * (int)(blockDim.x >> 6) + w) * G + g;
    unsigned char *mine = smem + (size_t)(w * G + g) * P.wave_lds;
    const uint32_t count = *P.work_count;

    // layout of this query's slice
    const uint32_t H = 1u << P.hlog;
    const uint32_t hmask = H - 1u;
    const uint32_t vis_limit = (H >> 2) * 3u;
    const int gbits = P.hlog - 3;
    const uint32_t vis_bytes = H * (P.vis16 ? 2u : 4u);
    const int QC = P.pq + P.slack + 2 * L;
    uint32_t *vis = reinterpret_cast<uint32_t *>(mine);
    Key *Q = reinterpret_cast<Key *>(mine + vis_bytes);
    Key *B = Q + QC;
    Key *stage = B + ((P.pq + 1) & ~1);
    uint32_t *cid = reinterpret_cast<uint32_t *>(stage + 64);
    float *qv = reinterpret_cast<float *>(cid + 64);
    // the layout must fit the slice, and a row must fit the sub-group's two L-word slices
    const bool layout_ok = reinterpret_cast<unsigned char *>(qv + P.qld) - mine <= (ptrdiff_t)P.wave_lds &&
                           (P.entry == nullptr || (P.n_entry >= 1 && P.n_entry <= 64)) && P.gstride == 32;
    const uint32_t GH = 1u << P.ghlog, gmask = GH - 1u, glimit = (GH >> 2) * 3u;
    uint32_t *gv = P.gvis ? P.gvis + (size_t)slot * GH : nullptr;

    // per-query state, uniform within the sub-group
    bool active = false, done = false;
    int qi = 0, pq = 0, k = 0, max_steps = 0, steps = 0, settle = 0;
    int h = 0, nq = 0, nbest = 0;
    uint64_t bmax = ~0ull;
    uint32_t flags = 0u, nvis = 0u, nglob = 0u, adj_words = 0u, pf_hits = 0u, compactions = 0u, truncs = 0u;
    bool spill = false;
    float lenv = 0.0f;
    uint32_t pid = NONE, prow0 = 0u, prow1 = 0u;
    const auto load_row = [&](uint32_t v, uint32_t &r0, uint32_t &r1) {
        const uint32_t *gr = P.graph + (size_t)v * (size_t)P.gstride;
        r0 = gr[min(S.sl, P.gstride - 1)];
        r1 = (L < 32) ? gr[min(S.sl + L, P.gstride - 1)] : 0u;
    };

    // visited insert of this lane's id; 0 present, 1 new, 2 no room in the LDS table
    const auto vis_add = [&](uint32_t id) -> int {
        if (!spill) {
            if (P.vis16) return vis16_insert(vis, gbits, P.idbits, id);
            return vis_insert(vis, hmask, P.hlog, id) ? 1 : 0;
        }
        const bool in_lds = P.vis16 ? vis16_test(vis, gbits, P.idbits, id) : vis_test(vis, hmask, P.hlog, id);
        return (!in_lds && gvis_insert(gv, gmask, P.ghlog, id)) ? 1 : 0;
    };
    const auto start_spill = [&]() -> bool {
        if (!gv) {
            flags |= F_VIS_OVF;
            return false;
        }
        if (!spill) {
            spill = true;
            for (uint32_t i = S.sl; i < GH; i += L) gv[i] = 0u;
            __threadfence_block();
        }
        return true;
    };
    // sorted insert of key e into the queue (nothing if it can never be expanded); false on
    // an overflow of the tie group at the pq-th distance
    const auto q_insert = [&](const Key &e) -> bool {
        if (nbest == pq && e.dk() > bmax) return true;
        if (h + nq + 1 > QC) {
            // move the live entries to the front
            for (int i0 = 0; i0 < nq; i0 += L) {
                const int i = i0 + S.sl;
                const Key v = Q[i < nq ? h + i : h];
                lds_fence();
                if (i < nq) Q[i] = v;
                lds_fence();
            }
            h = 0;
            compactions += S.sl == 0 ? 1u : 0u;
        }
        const int r = (nq > 0 && lt(Q[h + nq - 1], e)) ? nq
                                                        : sub_prefix_count(S, Q + h, nq, [&](const Key &x) { return lt(x, e); });
        sub_shift_right(S, Q + h, r, nq);
        if (S.sl == 0) Q[h + r] = e;
        lds_fence();
        ++nq;
        if (nq > pq + P.slack) {
            // drop the entries beyond the pq-th smallest distance (ties kept): nothing beyond it
            // can be expanded before the stop rule
            truncs += S.sl == 0 ? 1u : 0u;
            const uint64_t dq = Q[h + pq - 1].dk();
            const uint64_t dcut = (nbest == pq && bmax < dq) ? bmax : dq;
            nq = sub_prefix_count(S, Q + h, nq, [&](const Key &x) { return x.dk() <= dcut; });
            if (nq > pq + P.slack) {
                flags |= F_Q_OVF;
                return false;
            }
        }
        return true;
    };

    const auto begin = [&](int q) {
        qi = q;
        flags = 0u;
        steps = settle = 0;
        h = nq = nbest = 0;
        bmax = ~0ull;
        nvis = nglob = adj_words = pf_hits = compactions = truncs = 0u;
        spill = false;
        pid = NONE;
        pq = P.pq_arr ? max(1, min(P.pq_arr[qi], P.pq)) : P.pq;
        k = max(0, min(min(P.k_arr ? P.k_arr[qi] : P.k, pq), P.out_ld));
        max_steps = P.steps_arr ? min(P.steps_arr[qi], P.max_steps) : P.max_steps;
        for (uint32_t i = S.sl; i < (vis_bytes >> 4); i += L) reinterpret_cast<uint4 *>(vis)[i] = make_uint4(0u, 0u, 0u, 0u);
        for (int i = S.sl; i < P.qld; i += L) qv[i] = (i < P.dim) ? P.queries[(size_t)qi * P.dim + i] : 0.0f;
        lds_fence();
        lenv = 0.0f;
        if (METRIC == M_COS) {
            if (S.sl == 0)
                for (int i = 0; i < P.dim; ++i) lenv = fmaf(qv[i], qv[i], lenv);
            lenv = S.bcastf(lenv, 0);
        }
        // start vertices, each entered once through the visited set (entries >= n skipped;
        // none left means vertex 0), then placed in the queue in (dist, id) order
        const int ne = P.entry ? P.n_entry : 1;
        int m0 = 0;
        bool any = false;
        for (int e0 = 0; e0 < ne; e0 += L) {
            const int j = e0 + S.sl;
            const uint32_t sid = (j < ne) ? (P.entry ? P.entry[(size_t)qi * ne + j] : 0u) : NONE;
            const bool use = j < ne && (uint64_t)sid < P.n;
            any |= S.ballot_(use) != 0u;
            const int ins = use ? vis_add(sid) : 0;
            if (S.ballot_(ins == 2) != 0u) flags |= F_VIS_OVF;
            const uint32_t fm = S.ballot_(ins == 1);
            if (ins == 1) cid[m0 + S.below(fm)] = sid;
            m0 += __popc(fm);
        }
        if (!any) {
            const int ins = S.sl == 0 ? vis_add(0u) : 0;
            if (S.ballot_(ins == 2) != 0u) flags |= F_VIS_OVF;
            if (S.sl == 0 && ins == 1) cid[0] = 0u;
            m0 = __popc(S.ballot_(ins == 1));
        }
        lds_fence();
        nvis = (uint32_t)m0;
        if (flags) return;
        sub_eval<METRIC, EXACT, STORE, SUB, NV, L, Key>(S, P, m0, cid, stage, qv, lenv);
        for (int c = 0; c < m0; ++c) q_insert(stage[c]);
    };

    const auto finish = [&]() {
        const int cnt = min(k, nbest);
        const size_t ob = (size_t)qi * (size_t)P.out_ld;
        bool written = false;
        if constexpr (STORE != ST_F32 && !EXACT) {
            if (P.rerank) {
                // exact FP32 distances for the whole pq-best list, then its k smallest in order
                for (int base = 0; base < nbest; base += 64) {
                    const int c = min(64, nbest - base);
                    for (int i = S.sl; i < c; i += L) cid[i] = B[base + i].id();
                    lds_fence();
                    sub_eval<METRIC, false, ST_F32, 4, 8, L, Key>(S, P, c, cid, stage, qv, lenv);
                    for (int i = S.sl; i < c; i += L) B[base + i] = stage[i];
                    lds_fence();
                }
                for (int r = 0; r < k; ++r) {
                    uint64_t best = ~0ull;
                    int bi = -1;
                    for (int i = S.sl; i < nbest; i += L) {
                        const uint64_t e = B[i].v;
                        if (e < best) {
                            best = e;
                            bi = i;
                        }
                    }
                    uint64_t m = best;
#pragma unroll
                    for (int o = L / 2; o > 0; o >>= 1) {
                        const uint64_t x = ((uint64_t)(uint32_t)__shfl_xor((int)(m >> 32), o) << 32) |
                                           (uint32_t)__shfl_xor((int)m, o);
                        if (x < m) m = x;
                    }
                    if (bi >= 0 && best == m) B[bi] = Key::inf();
                    lds_fence();
                    if (S.sl == 0) {
                        const bool ok = r < cnt;
                        const Key w = Key{m};
                        P.out_ids[ob + r] = ok ? w.id() : 0xFFFFFFFFu;
                        if (P.out_dist) P.out_dist[ob + r] = ok ? w.dist() : __int_as_float(0x7f800000);
                    }
                }
                written = true;
            }
        }
        for (int i = S.sl; i < k && !written; i += L) {
            const bool ok = i < cnt;
            const Key e = ok ? B[i] : Key::inf();
            P.out_ids[ob + i] = ok ? e.id() : 0xFFFFFFFFu;
            if (P.out_dist) P.out_dist[ob + i] = ok ? e.dist() : __int_as_float(0x7f800000);
        }
        if (S.sl == 0) {
            if (P.retry_list && (flags & (F_VIS_OVF | F_Q_OVF))) P.retry_list[atomicAdd(P.retry_count, 1u)] = (uint32_t)qi;
            if (P.out_flags) P.out_flags[qi] = flags;
            if (P.out_stats) {
                uint32_t *o = P.out_stats + 8 * (size_t)qi;
                o[0] = (uint32_t)steps;
                o[1] = nvis;
                o[2] = adj_words;
                o[3] = pf_hits;
                o[4] = (uint32_t)settle;
                o[5] = compactions;
                o[6] = truncs;
                o[7] = 0u;
            }
        }
    };

    // `now` into the best list (push + pop-max of graph.h:193-195), then its neighbours in row
    // order (graph.h:196-205) through the visited set: the new ones are appended to cid[m..).
    // Lane sl holds row words sl and sl + L.  False when the search must end (visited overflow).
    const auto visit = [&](const Key &now, uint32_t r0, uint32_t r1, int &m) -> bool {
        {
            const int p = (nbest == 0 || lt(B[nbest - 1], now))
                              ? nbest
                              : sub_prefix_count(S, B, nbest, [&](const Key &x) { return lt(x, now); });
            if (p < k) settle = steps;
            if (p < pq) {
                sub_shift_right(S, B, p, min(nbest, pq - 1));
                if (S.sl == 0) B[p] = now;
                lds_fence();
                nbest = min(nbest + 1, pq);
                if (nbest == pq) bmax = B[pq - 1].dk();
            }
        }
        const int deg = (int)S.bcast(r0, 0);
        adj_words += S.sl == 0 ? 1u + (uint32_t)deg : 0u;
        if (!spill && nvis + (uint32_t)m + 32u > vis_limit && !start_spill()) return false;
#pragma unroll
        for (int half = 0; half < (L < 32 ? 2 : 1); ++half) {
            const int word = S.sl + half * L;
            const uint32_t id = half ? r1 : r0;
            const bool mine = word >= 1 && word <= deg;
            int ins = mine ? vis_add(id) : 0;
            if (S.ballot_(ins == 2) != 0u) {
                // the compact table has no room near these ids: continue in the global table
                if (!start_spill()) return false;
                if (ins == 2) ins = gvis_insert(gv, gmask, P.ghlog, id) ? 1 : 0;
            }
            const uint32_t fm = S.ballot_(ins == 1);
            if (ins == 1) cid[m + S.below(fm)] = id;
            m += __popc(fm);
        }
        return true;
    };

    // one step: the queue head (stop rule first) and, with multi-probe, the next head too
    // (the reference's N_MULTIPROBE, warp_astar_accelerator.h:184-197: the stop rule is
    // tested on the step's first pop only); true when the search has ended
    const auto expand = [&]() -> bool {
        if (flags) return true;
        if (nq == 0) return true;
        if (steps >= max_steps) {
            flags |= F_STEP_CAP;
            return true;
        }
        const Key now = Q[h];
        // stop rule, graph.h:188
        if (nbest == pq && bmax < now.dk()) return true;
        ++steps;
        ++h;
        --nq;
        uint32_t r0, r1;
        if (pid == now.id()) {
            r0 = prow0;
            r1 = prow1;
            pf_hits += S.sl == 0 ? 1u : 0u;
        } else {
            load_row(now.id(), r0, r1);
        }
        // the likely next expansion's row (the new queue head), one expansion ahead
        pid = nq > 0 ? Q[h].id() : NONE;
        load_row(nq > 0 ? pid : 0u, prow0, prow1);
        int m = 0;
        if (!visit(now, r0, r1, m)) return true;
        if (P.multiprobe > 1 && nq > 0 && steps < max_steps) {
            const Key now2 = Q[h];
            ++steps;
            ++h;
            --nq;
            const uint32_t s0 = prow0, s1 = prow1;  // its row: the prefetch above
            pf_hits += S.sl == 0 ? 1u : 0u;
            pid = nq > 0 ? Q[h].id() : NONE;
            load_row(nq > 0 ? pid : 0u, prow0, prow1);
            if (!visit(now2, s0, s1, m)) return true;
        }
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
        sub_eval<METRIC, EXACT, STORE, SUB, NV, L, Key>(S, P, m, cid, stage, qv, lenv);
        for (int c = 0; c < m; ++c)
            if (!q_insert(stage[c])) return true;
        // the queue head may have changed: refresh the prefetch target only if it did
        if (nq > 0 && Q[h].id() != pid) {
            pid = Q[h].id();
            load_row(pid, prow0, prow1);
        }
        return false;
    };

    if (P.wave_prio) __builtin_amdgcn_s_setprio(2);
    for (;;) {
        // sub-groups without a query take the next list position (every lane issues the add,
        // the first lane of a needy sub-group adds 1; the branch is uniform over the wave)
        const bool need = !active && !done;
        if (__ballot(need) != 0ull) {
            const uint32_t v = atomicAdd(P.work_next, (need && S.sl == 0) ? 1u : 0u);
            const uint32_t t = S.bcast(v, 0);
            if (need) {
                if (t >= count) {
                    done = true;
                } else {
                    const int q = P.qidx ? (int)P.qidx[t] : (int)t;
                    if (!layout_ok) {
                        if (S.sl == 0 && P.out_flags) P.out_flags[q] = F_BAD_LAUNCH;
                    } else {
                        begin(q);
                        active = true;
                    }
                }
            }
        }
        if (__ballot(!done) == 0ull) break;
        if (active && expand()) {
            finish();
            active = false;
        }
    }
}

}  // namespace song
