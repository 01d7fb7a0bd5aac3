// Batching form of the search (include/song_amd.h, song_batch_*): queries with uneven
// k and pq in one run, persistent per-tier kernels, capture-safe enqueue, hipGraph replay.
//
// One run enqueues, on the caller's stream:
//   1. batch_reset_kernel   zero the run's counters (tier sizes, work counters, retry count)
//   2. batch_bucket_kernel  query i -> the first tier with pq_i <= pq_max, appended to that
//                           tier's list (wave-aggregated atomics); no tier -> BAD_PARAMS
//   3. search_persistent    one launch per tier, largest pq first: its grid fills the GPU at
//                           the tier's LDS size, each wave takes list positions from the
//                           tier's counter.  With parallel tiers the launches after the first
//                           go to forked streams (event fork / join), so a tier's workgroups
//                           take the slots that another tier's finished waves free
//   4. search_persistent    the retry configuration over the queries step 3 flagged, after
//                           the join
// Nothing here allocates, clears memory with a memset or waits on the host, so the steps
// can be captured into a hipGraph as they are (the fork / join becomes graph edges).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "index_internal.hpp"

using namespace song;

namespace {

struct TierBounds {
    int pq_max[SONG_MAX_TIERS];
    int n;
};

__global__ void batch_reset_kernel(uint32_t *ctr, int n) {
    const int i = (int)threadIdx.x;
    if (i < n) ctr[i] = 0u;
}

__global__ __launch_bounds__(256) void batch_bucket_kernel(const int *d_pq, int pq, uint32_t nq, TierBounds tb,
                                                           uint32_t *count, uint32_t *lists, uint32_t list_ld,
                                                           uint32_t *flags, uint32_t *ids, int k_ld) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < nq;
    const int p = valid ? (d_pq ? d_pq[i] : pq) : 0;
    int t = -1;
    if (valid && p >= 1)
        for (int j = tb.n - 1; j >= 0; --j)
            if (p <= tb.pq_max[j]) t = j;
    if (valid && t < 0) {
        flags[i] = SONG_FLAG_BAD_PARAMS;
        for (int j = 0; j < k_ld; ++j) ids[(size_t)i * k_ld + j] = 0xFFFFFFFFu;
    }
    // every lane adds 0 or 1 and gets its own slot (the compiler folds the wave's adds
    // into one atomic); no lane-divergent atomic feeding a cross-lane read
    for (int j = 0; j < tb.n; ++j) {
        const bool mine = valid && t == j;
        const uint32_t pos = atomicAdd(&count[j], mine ? 1u : 0u);
        if (mine) lists[(size_t)j * list_ld + pos] = i;
    }
}

struct Launch {
    Plan plan;
    SearchFn fn = nullptr;
    int groups = 0, waves = 1;
    int qpw = 1;  // queries per wave
    int multiprobe = 1;
    int duo = 0;  // two waves per query
    size_t wave_lds = 0;
    uint32_t *gvis = nullptr;
    int ghlog = 13;  // spill table per wave: 2^ghlog u32
    unsigned char *gbag = nullptr;  // queue bag per wave: bag_stride(plan) bytes
    size_t bag_stride = 0;
    int pq_max = 0;
};

struct RunArgs {
    const float *q;
    uint64_t nq;
    int pq, k;
    song_query_params qp;
    uint32_t *ids;
    float *dist;
    int k_ld;
    uint32_t *flags, *stats;
    bool operator==(const RunArgs &o) const {
        return q == o.q && nq == o.nq && pq == o.pq && k == o.k && qp.pq == o.qp.pq && qp.k == o.qp.k &&
               qp.entry == o.qp.entry && qp.n_entry == o.qp.n_entry && qp.steps == o.qp.steps && ids == o.ids &&
               dist == o.dist && k_ld == o.k_ld && flags == o.flags && stats == o.stats;
    }
};

}  // namespace

struct song_batch {
    song_index *ix = nullptr;
    song_batch_config cfg;
    uint64_t nq_max = 0;
    int n_cu = 0;
    std::vector<Launch> tiers;  // n_tiers, then the retry
    uint32_t *d_ctr = nullptr;  // count[T], next[T], retry count, retry next
    uint32_t *d_lists = nullptr;  // T x nq_max
    uint32_t *d_retry = nullptr;  // nq_max
    std::vector<void *> allocs;
    // parallel tiers: one stream per tier after the first and the fork / join events
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> ev;
    struct Cached {
        RunArgs a;
        void *stream;
        hipGraphExec_t exec;
    };
    std::vector<Cached> graphs;
};

static int default_tier_hlog(int pq) { return default_hlog(pq); }

extern "C" void song_batch_config_default(song_batch_config *c, int pq_max) {
    memset(c, 0, sizeof(*c));
    c->mode = SONG_MODE_EXACT;
    c->storage = SONG_STORE_F32;
    c->rerank = 1;
    c->max_steps = 0;
    c->spill_log2 = 14;
    c->n_tiers = 1;
    song_tier t;
    t.pq_max = pq_max;
    t.hash_log2 = 0;
    t.slack = 32;
    t.shape = -1;
    t.waves_per_group = 1;
    t.groups_per_cu = 0;
    t.visited_bits = 0;
    t.storage = -1;
    t.bloom_k = 0;
    t.queries_per_wave = 1;
    t.multiprobe = 1;
    t.bag_in_lds = 0;
    t.waves_per_query = 1;
    t.best_in_global = -1;  // = 0: a tier's share of the batch is not known when it is created
    c->tier[0] = t;
    c->retry = t;
    c->retry.hash_log2 = 0;  // the largest tier's filter x 4
    c->retry.slack = 4 * 32 + 64;
    c->retry.groups_per_cu = 0;
    c->parallel_tiers = 1;
    c->priority_first = 1;
}

extern "C" int song_resolve_mode(const song_index *ix, const float *d_queries, uint64_t nq, int *mode, void *stream) {
    if (!ix || !mode) return fail(SONG_E_ARG, "index/mode is null");
    *mode = SONG_MODE_EXACT;
    if (!ix->integer_valued || ix->metric == SONG_COS || nq == 0) return SONG_OK;
    if (!d_queries) return fail(SONG_E_ARG, "queries are null");
    int qint = 0;
    float qmin = 0.f, qmax = 0.f;
    int rc = device_range(d_queries, nq * (uint64_t)ix->dim, &qint, &qmin, &qmax, (hipStream_t)stream);
    if (rc != SONG_OK) return rc;
    *mode = fp32_exact(ix, qint, qmin, qmax) ? SONG_MODE_FP32 : SONG_MODE_EXACT;
    return SONG_OK;
}

static int check_tier(const song_tier &t, int idx) {
    if (t.pq_max < 1 || t.pq_max > 4096) return fail(SONG_E_ARG, "tier %d: pq_max=%d not in [1,4096]", idx, t.pq_max);
    if (t.bloom_k < -1 || t.bloom_k > 9) return fail(SONG_E_ARG, "tier %d: bloom_k=%d not in [-1,9]", idx, t.bloom_k);
    if (t.bloom_k == 0 && t.hash_log2 != 0 && (t.hash_log2 < 8 || t.hash_log2 > 16))
        return fail(SONG_E_ARG, "tier %d: hash_log2=%d not in [8,16]", idx, t.hash_log2);
    if (t.bloom_k != 0 && (t.hash_log2 < 0 || t.hash_log2 > 15))
        return fail(SONG_E_ARG, "tier %d: Bloom filter of 2^%d words not in [2^1,2^15]", idx, t.hash_log2);
    if (t.slack < 0 || t.slack > 4096) return fail(SONG_E_ARG, "tier %d: slack=%d", idx, t.slack);
    if (t.shape < -1 || t.shape > 4) return fail(SONG_E_ARG, "tier %d: shape=%d not in [-1,4]", idx, t.shape);
    if (t.waves_per_group != 1 && t.waves_per_group != 2 && t.waves_per_group != 4)
        return fail(SONG_E_ARG, "tier %d: waves_per_group=%d not 1, 2 or 4", idx, t.waves_per_group);
    if (t.groups_per_cu < 0 || t.groups_per_cu > 64)
        return fail(SONG_E_ARG, "tier %d: groups_per_cu=%d", idx, t.groups_per_cu);
    if (t.visited_bits != 0 && t.visited_bits != 16 && t.visited_bits != 32)
        return fail(SONG_E_ARG, "tier %d: visited_bits=%d not 0, 16 or 32", idx, t.visited_bits);
    if (t.queries_per_wave != 0 && t.queries_per_wave != 1 && t.queries_per_wave != 2 && t.queries_per_wave != 4)
        return fail(SONG_E_ARG, "tier %d: queries_per_wave=%d not 1, 2 or 4", idx, t.queries_per_wave);
    if (t.multiprobe < 0 || t.multiprobe > 2) return fail(SONG_E_ARG, "tier %d: multiprobe=%d not 1 or 2", idx, t.multiprobe);
    if (t.multiprobe == 2 && t.queries_per_wave < 2)
        return fail(SONG_E_ARG, "tier %d: multi-probe runs with several queries per wave", idx);
    if (t.bag_in_lds != 0 && t.bag_in_lds != 1) return fail(SONG_E_ARG, "tier %d: bag_in_lds=%d", idx, t.bag_in_lds);
    if (t.waves_per_query < 0 || t.waves_per_query > 2)
        return fail(SONG_E_ARG, "tier %d: waves_per_query=%d not 1 or 2", idx, t.waves_per_query);
    if (t.storage < -1 || t.storage > SONG_STORE_FP8)
        return fail(SONG_E_ARG, "tier %d: storage=%d not -1 or a SONG_STORE_* value", idx, t.storage);
    return SONG_OK;
}

extern "C" int song_batch_destroy(song_batch *b) {
    if (!b) return SONG_OK;
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.exec);
    for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
    for (hipStream_t s : b->side) (void)hipStreamDestroy(s);
    for (void *p : b->allocs) (void)hipFree(p);
    delete b;
    return SONG_OK;
}

extern "C" int song_batch_create(song_index *ix, const song_batch_config *cfg, uint64_t nq_max, song_batch **out) {
    if (!out) return fail(SONG_E_ARG, "out is null");
    *out = nullptr;
    if (!ix || !cfg) return fail(SONG_E_ARG, "index/config is null");
    if (nq_max == 0 || nq_max > 0x7fffffffull) return fail(SONG_E_ARG, "nq_max=%llu", (unsigned long long)nq_max);
    const song_batch_config &c = *cfg;
    if (c.mode != SONG_MODE_EXACT && c.mode != SONG_MODE_FP32)
        return fail(SONG_E_ARG, "batch mode must be EXACT or FP32 (song_resolve_mode picks one)");
    if (c.storage < SONG_STORE_F32 || c.storage > SONG_STORE_FP8) return fail(SONG_E_ARG, "storage=%d", c.storage);
    if (c.n_tiers < 1 || c.n_tiers > SONG_MAX_TIERS) return fail(SONG_E_ARG, "n_tiers=%d", c.n_tiers);
    const int spill = c.spill_log2 ? c.spill_log2 : 14;
    if (spill < 10 || spill > 18) return fail(SONG_E_ARG, "spill_log2=%d not in [10,18]", c.spill_log2);
    for (int t = 0; t < c.n_tiers; ++t) {
        int rc = check_tier(c.tier[t], t);
        if (rc != SONG_OK) return rc;
        if (t > 0 && c.tier[t].pq_max <= c.tier[t - 1].pq_max) return fail(SONG_E_ARG, "tier pq_max not ascending");
    }
    song_tier rt = c.retry;
    const int pq_top = c.tier[c.n_tiers - 1].pq_max;
    rt.pq_max = pq_top;
    rt.bloom_k = 0;  // re-runs use the exact table
    {
        int rc = check_tier(rt, c.n_tiers);
        if (rc != SONG_OK) return rc;
    }
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));

    song_batch *b = new song_batch();
    b->ix = ix;
    b->cfg = c;
    b->cfg.spill_log2 = spill;
    b->nq_max = nq_max;
    b->n_cu = prop.multiProcessorCount;
    auto cleanup = [&](int code) {
        song_batch_destroy(b);
        return code;
    };
    const bool exact = c.mode == SONG_MODE_EXACT;
    // the re-run's default table is 4x the largest exact tier's (or the default for the top pq
    // when every tier uses a Bloom filter)
    int top_hlog = 0;
    for (int t = 0; t < c.n_tiers; ++t)
        if (c.tier[t].bloom_k == 0)
            top_hlog = std::max(top_hlog, c.tier[t].hash_log2 ? c.tier[t].hash_log2 : default_tier_hlog(c.tier[t].pq_max));
    if (top_hlog == 0) top_hlog = default_tier_hlog(pq_top);
    for (int t = 0; t <= c.n_tiers; ++t) {
        const bool is_retry = t == c.n_tiers;
        const song_tier &ti = is_retry ? rt : c.tier[t];
        int hlog = ti.hash_log2 ? ti.hash_log2 : (ti.bloom_k ? 0 : default_tier_hlog(ti.pq_max));
        if (is_retry && !ti.hash_log2) hlog = std::min(16, top_hlog + 2);
        Launch L;
        L.pq_max = ti.pq_max;
        const int store = ti.storage >= 0 ? ti.storage : c.storage;
        const int qpw = is_retry ? 1 : std::max(1, ti.queries_per_wave);
        const int bagl = is_retry ? 0 : ti.bag_in_lds;
        const int wpq = is_retry ? 1 : std::max(1, ti.waves_per_query);
        int rc = make_plan(ix, exact, store, ti.pq_max, ti.slack, hlog, &L.plan, ti.shape, ti.visited_bits, ti.bloom_k,
                           qpw, bagl, wpq, (is_retry || ti.best_in_global < 0) ? 0 : ti.best_in_global);
        // the re-run's default table (the largest tier's, 4x) may not fit one wave's LDS: take the
        // largest that does (the re-run also has a 4x larger global spill table)
        while (rc == SONG_E_OVERFLOW && is_retry && !ti.hash_log2 && hlog > top_hlog) {
            --hlog;
            rc = make_plan(ix, exact, store, ti.pq_max, ti.slack, hlog, &L.plan, ti.shape, ti.visited_bits, 0);
        }
        if (rc != SONG_OK) return cleanup(rc);
        L.fn = search_fn(ix->metric, L.plan, true);
        L.duo = L.plan.duo;
        // two waves per query: one query (search wave + helper wave) per workgroup
        L.waves = L.duo ? 1 : ti.waves_per_group;
        L.qpw = 64 / L.plan.lanes;
        L.multiprobe = is_retry ? 1 : std::max(1, ti.multiprobe);
        // one wave's LDS (lanes 64), or one query's slice (several queries per wave)
        L.wave_lds = (L.plan.lds + 15) & ~(size_t)15;
        const size_t lds = L.wave_lds * (size_t)L.waves * (size_t)L.qpw + (L.duo ? (size_t)DUO_BYTES : 0);
        if (lds > ix->max_lds)
            return cleanup(fail(SONG_E_OVERFLOW, "tier %d: %d waves x %zu bytes of LDS exceed %zu", t, L.waves,
                                L.wave_lds, ix->max_lds));
        if (lds > 65536 &&
            hipFuncSetAttribute((const void *)L.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return cleanup(fail(SONG_E_HIP, "tier %d: LDS attribute %zu", t, lds));
        int per_cu = ti.groups_per_cu;
        if (per_cu == 0) {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)L.fn, 64 * L.waves * (L.duo ? 2 : 1),
                                                             lds) != hipSuccess ||
                nb < 1)
                return cleanup(fail(SONG_E_HIP, "tier %d: occupancy query failed", t));
            per_cu = nb;
        }
        L.groups = per_cu * b->n_cu;
        const size_t slots = (size_t)L.groups * L.waves * L.qpw;
        // the re-run gets spill tables 4x larger than the tiers'
        L.ghlog = is_retry ? std::min(18, spill + 2) : spill;
        if (hipMalloc(&L.gvis, slots * ((size_t)4 << L.ghlog)) != hipSuccess)
            return cleanup(fail(SONG_E_HIP, "tier %d: spill tables (%zu slots)", t, slots));
        b->allocs.push_back(L.gvis);
        // queue bags: the one-query-per-wave kernel's (the sub-group kernel keeps its queue in LDS)
        L.bag_stride = (L.qpw > 1 || L.plan.bag_lds) ? 0 : song::bag_stride(L.plan);
        if (L.bag_stride) {
            if (hipMalloc(&L.gbag, slots * L.bag_stride) != hipSuccess)
                return cleanup(fail(SONG_E_HIP, "tier %d: queue bags (%zu slots)", t, slots));
            b->allocs.push_back(L.gbag);
        }
        b->tiers.push_back(L);
    }
    const int T = c.n_tiers;
    if (hipMalloc(&b->d_ctr, sizeof(uint32_t) * (2 * T + 2)) != hipSuccess ||
        (b->allocs.push_back(b->d_ctr), hipMalloc(&b->d_lists, sizeof(uint32_t) * nq_max * T)) != hipSuccess ||
        (b->allocs.push_back(b->d_lists), hipMalloc(&b->d_retry, sizeof(uint32_t) * nq_max)) != hipSuccess)
        return cleanup(fail(SONG_E_HIP, "batch work lists"));
    b->allocs.push_back(b->d_retry);
    if (c.parallel_tiers && T > 1) {
        // default-priority, non-blocking streams; event 0 forks, events 1.. join
        for (int t = 1; t < T; ++t) {
            hipStream_t s = nullptr;
            if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
                return cleanup(fail(SONG_E_HIP, "tier stream"));
            b->side.push_back(s);
        }
        for (int t = 0; t < T; ++t) {
            hipEvent_t e = nullptr;
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                return cleanup(fail(SONG_E_HIP, "tier event"));
            b->ev.push_back(e);
        }
    }
    if (hipMemset(b->d_ctr, 0, sizeof(uint32_t) * (2 * T + 2)) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
        return cleanup(fail(SONG_E_HIP, "batch counters"));
    *out = b;
    return SONG_OK;
}

static int check_args(const song_batch *b, const RunArgs &a) {
    const song_index *ix = b->ix;
    if (a.nq > b->nq_max) return fail(SONG_E_ARG, "nq=%llu > nq_max", (unsigned long long)a.nq);
    if (!a.q || !a.ids || !a.flags) return fail(SONG_E_ARG, "queries/ids/flags are null");
    if (a.pq < 1 || a.k < 1 || a.k_ld < a.k) return fail(SONG_E_ARG, "pq=%d k=%d k_ld=%d", a.pq, a.k, a.k_ld);
    if (!a.qp.pq && a.pq > b->tiers[b->cfg.n_tiers - 1].pq_max)
        return fail(SONG_E_ARG, "pq=%d exceeds the largest tier", a.pq);
    if (a.qp.subset) return fail(SONG_E_ARG, "query subsets are not supported by batches");
    if (a.qp.entry && (a.qp.n_entry < 1 || a.qp.n_entry > 64))
        return fail(SONG_E_ARG, "n_entry=%d not in [1,64]", a.qp.n_entry);
    for (const Launch &L : b->tiers)
        if (L.plan.store != SONG_STORE_F32 && !ix->d_lp[L.plan.store])
            return fail(SONG_E_ARG, "storage %d not built", L.plan.store);
    return SONG_OK;
}

static int enqueue(song_batch *b, const RunArgs &a, hipStream_t st) {
    const song_index *ix = b->ix;
    const song_batch_config &c = b->cfg;
    const int T = c.n_tiers;
    uint32_t *count = b->d_ctr, *next = b->d_ctr + T, *rcount = b->d_ctr + 2 * T, *rnext = rcount + 1;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, st, b->d_ctr, 2 * T + 2);
    HIP_TRY(hipGetLastError());
    TierBounds tb{};
    tb.n = T;
    for (int t = 0; t < T; ++t) tb.pq_max[t] = c.tier[t].pq_max;
    hipLaunchKernelGGL(batch_bucket_kernel, dim3((unsigned)((a.nq + 255) / 256)), dim3(256), 0, st, a.qp.pq, a.pq,
                       (uint32_t)a.nq, tb, count, b->d_lists, (uint32_t)b->nq_max, a.flags, a.ids, a.k_ld);
    HIP_TRY(hipGetLastError());

    SearchParams sp{};
    sp.data = ix->d_data;
    sp.graph = ix->d_graph;
    sp.norms = ix->d_norms;
    sp.queries = a.q;
    sp.entry = a.qp.entry;
    sp.n_entry = a.qp.entry ? a.qp.n_entry : 1;
    sp.steps_arr = a.qp.steps;
    sp.pq_arr = a.qp.pq;
    sp.k_arr = a.qp.k;
    sp.out_ids = a.ids;
    sp.out_dist = a.dist;
    sp.out_flags = a.flags;
    sp.out_stats = a.stats;
    sp.n = ix->n;
    sp.dim = ix->dim;
    sp.ld = ix->ld;
    sp.gstride = ix->gstride;
    sp.nq = (int)a.nq;
    sp.k = a.k;
    sp.out_ld = a.k_ld;
    sp.max_steps = c.max_steps > 0 ? c.max_steps : 1000000;
    sp.rerank = c.rerank;
    sp.ghlog = c.spill_log2;
    sp.qbase = 0;
    // launch order: tiers by descending pq (longest searches first), then the retry; with
    // parallel tiers every tier but the first launched runs on its own forked stream
    const bool par = !b->side.empty();
    if (par) HIP_TRY(hipEventRecord(b->ev[0], st));
    for (int o = 0; o <= T; ++o) {
        const int t = o < T ? T - 1 - o : T;
        const Launch &L = b->tiers[t];
        const bool is_retry = t == T;
        hipStream_t ls = st;
        if (par && o > 0 && o < T) {
            ls = b->side[o - 1];
            HIP_TRY(hipStreamWaitEvent(ls, b->ev[0], 0));
        }
        if (par && is_retry)
            for (int j = 1; j < T; ++j) HIP_TRY(hipStreamWaitEvent(st, b->ev[j], 0));
        SearchParams p = sp;
        // without per-query budgets every query has the call's pq (the LDS slice is sized
        // for the tier's pq_max >= pq)
        p.pq = a.qp.pq ? L.pq_max : a.pq;
        p.hlog = L.plan.hlog;
        p.slack = L.plan.slack;
        p.qcap = L.plan.qcap;
        p.qld = L.plan.qld;
        p.data_lp = ix->d_lp[L.plan.store];
        p.scale = ix->d_scale[L.plan.store];
        p.lp_stride = ix->lp_stride[L.plan.store];
        p.gvis = L.gvis;
        p.ghlog = L.ghlog;
        p.gbag = L.gbag;
        p.gbag_stride = L.bag_stride;
        p.wave_lds = (uint32_t)L.wave_lds;
        p.vis16 = L.plan.vis16;
        p.bloom_k = L.plan.bloom_k;
        p.multiprobe = L.multiprobe;
        p.idbits = ix->idbits;
        // with tiers running side by side, the largest-pq tier (launched first, the longest
        // searches) gets issue priority
        p.wave_prio = (par && o == 0 && c.priority_first) ? 1 : 0;
        if (!is_retry) {
            p.qidx = b->d_lists + (size_t)t * b->nq_max;
            p.work_count = count + t;
            p.work_next = next + t;
            p.retry_list = b->d_retry;
            p.retry_count = rcount;
        } else {
            p.qidx = b->d_retry;
            p.work_count = rcount;
            p.work_next = rnext;
            p.retry_list = nullptr;
            p.retry_count = nullptr;
        }
        hipLaunchKernelGGL(L.fn, dim3((unsigned)L.groups), dim3(64 * L.waves * (L.duo ? 2 : 1)),
                           L.wave_lds * L.waves * L.qpw + (L.duo ? (size_t)DUO_BYTES : 0), ls, p);
        HIP_TRY(hipGetLastError());
        if (ls != st) HIP_TRY(hipEventRecord(b->ev[o], ls));
    }
    return SONG_OK;
}

static RunArgs make_args(const float *q, uint64_t nq, int pq, int k, const song_query_params *qp, uint32_t *ids,
                         float *dist, int k_ld, uint32_t *flags, uint32_t *stats) {
    RunArgs a{};
    a.q = q;
    a.nq = nq;
    a.pq = pq;
    a.k = k;
    if (qp) a.qp = *qp;
    if (!a.qp.entry) a.qp.n_entry = 1;
    a.ids = ids;
    a.dist = dist;
    a.k_ld = k_ld;
    a.flags = flags;
    a.stats = stats;
    return a;
}

extern "C" int song_batch_run(song_batch *b, const float *d_queries, uint64_t nq, int pq, int k,
                              const song_query_params *qp, uint32_t *d_ids, float *d_dist, int k_ld, uint32_t *d_flags,
                              uint32_t *d_stats, void *stream) {
    if (!b) return fail(SONG_E_ARG, "batch is null");
    if (nq == 0) return SONG_OK;
    RunArgs a = make_args(d_queries, nq, pq, k, qp, d_ids, d_dist, k_ld, d_flags, d_stats);
    int rc = check_args(b, a);
    if (rc != SONG_OK) return rc;
    return enqueue(b, a, (hipStream_t)stream);
}

extern "C" int song_batch_run_graph(song_batch *b, const float *d_queries, uint64_t nq, int pq, int k,
                                    const song_query_params *qp, uint32_t *d_ids, float *d_dist, int k_ld,
                                    uint32_t *d_flags, uint32_t *d_stats, void *stream) {
    if (!b) return fail(SONG_E_ARG, "batch is null");
    if (!stream) return fail(SONG_E_ARG, "graph runs need a stream other than the default one");
    if (nq == 0) return SONG_OK;
    RunArgs a = make_args(d_queries, nq, pq, k, qp, d_ids, d_dist, k_ld, d_flags, d_stats);
    int rc = check_args(b, a);
    if (rc != SONG_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    for (auto &g : b->graphs)
        if (g.stream == stream && g.a == a) {
            HIP_TRY(hipGraphLaunch(g.exec, st));
            return SONG_OK;
        }
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    rc = enqueue(b, a, st);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc != SONG_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(SONG_E_HIP, "capture: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(SONG_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
    if (b->graphs.size() >= 8) {
        (void)hipGraphExecDestroy(b->graphs.front().exec);
        b->graphs.erase(b->graphs.begin());
    }
    b->graphs.push_back({a, stream, exec});
    HIP_TRY(hipGraphLaunch(exec, st));
    return SONG_OK;
}

extern "C" int song_batch_counts(song_batch *b, uint32_t *out, void *stream) {
    if (!b || !out) return fail(SONG_E_ARG, "batch/out is null");
    const int T = b->cfg.n_tiers;
    std::vector<uint32_t> h(2 * T + 2);
    HIP_TRY(hipMemcpyAsync(h.data(), b->d_ctr, h.size() * 4, hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    for (int t = 0; t < T; ++t) out[t] = h[t];
    out[T] = h[2 * T];
    return SONG_OK;
}

extern "C" int song_batch_grid(const song_batch *b, int t, int *groups, int *waves_per_group, int *lds_per_wave) {
    if (!b || t < 0 || t > b->cfg.n_tiers) return fail(SONG_E_ARG, "tier %d", t);
    const Launch &L = b->tiers[t];
    if (groups) *groups = L.groups;
    if (waves_per_group) *waves_per_group = L.waves;
    if (lds_per_wave) *lds_per_wave = (int)(L.wave_lds * L.qpw);
    return SONG_OK;
}
