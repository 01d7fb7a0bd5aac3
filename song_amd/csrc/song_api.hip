// C ABI of the graph search (include/song_amd.h): index upload, launch planning,
// overflow re-runs.  Everything that touches the search runs on the GPU; there is no
// CPU search path in this library.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "index_internal.hpp"

using namespace song;

namespace song {

static thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace song

// visited spill tables: 2^13 u32 (32 KB) per launch slot, launches of at most kChunk queries
static constexpr int kGhlog = 13;
static constexpr unsigned kChunk = 32768;

static int ensure_gvis(song_index *ix, void *stream, size_t slots, uint32_t **out) {
    std::lock_guard<std::mutex> g(ix->spill_mu);
    song_index::Spill &sp = ix->spill[stream];
    if (sp.slots < slots) {
        if (sp.d) {
            // the previous buffer may still be read by work queued on this stream
            HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
            HIP_TRY(hipFree(sp.d));
        }
        sp.d = nullptr;
        sp.slots = 0;
        HIP_TRY(hipMalloc(&sp.d, slots * ((size_t)4 << kGhlog)));
        sp.slots = slots;
    }
    *out = sp.d;
    return SONG_OK;
}

// queue bags: `stride` bytes (qcap keys) per launch slot, grown on demand like the tables
static int ensure_gbag(song_index *ix, void *stream, size_t slots, size_t stride, unsigned char **out) {
    std::lock_guard<std::mutex> g(ix->spill_mu);
    song_index::Spill &sp = ix->spill[stream];
    const size_t need = slots * stride;
    if (sp.bag_bytes < need) {
        if (sp.bag) {
            HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
            HIP_TRY(hipFree(sp.bag));
        }
        sp.bag = nullptr;
        sp.bag_bytes = 0;
        HIP_TRY(hipMalloc(&sp.bag, need));
        sp.bag_bytes = need;
    }
    *out = sp.bag;
    return SONG_OK;
}

extern "C" const char *song_last_error(void) { return song::g_err.c_str(); }

extern "C" void song_search_opts_default(song_search_opts *o) {
    o->mode = SONG_MODE_AUTO;
    o->slack = 128;
    o->hash_log2 = 0;
    o->max_steps = 0;
    o->retry = 1;
    o->storage = SONG_STORE_F32;
    o->rerank = 1;
    o->visited_bits = 0;
    o->bloom_k = 0;
    o->best_in_global = -1;
}

// Low-precision copy of one row per thread: BF16 with a power-of-two scale (round to
// nearest even), or FP8-E4M3 with scale max|x|/448 (hardware conversion of the clamped
// value).  The scale of an all-zero row is 1.
__global__ void quantize_kernel(const float *data, uint64_t n, int dim, int ld, int store, uint8_t *out,
                                size_t stride, float *scale) {
    uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const float *r = data + v * (uint64_t)ld;
    float amax = 0.0f;
    for (int i = 0; i < dim; ++i) amax = fmaxf(amax, fabsf(r[i]));
    float s = 1.0f;
    if (amax > 0.0f) {
        if (store == SONG_STORE_BF16) {
            int e;
            frexpf(amax, &e);
            s = ldexpf(1.0f, e);
        } else {
            s = amax / 448.0f;
        }
    }
    scale[v] = s;
    uint32_t *o = reinterpret_cast<uint32_t *>(out + v * stride);
    const int words = (int)(stride / 4);
    for (int w = 0; w < words; ++w) {
        uint32_t word = 0u;
        if (store == SONG_STORE_BF16) {
            for (int t = 0; t < 2; ++t) {
                const int i = 2 * w + t;
                const float x = i < dim ? r[i] / s : 0.0f;
                const uint32_t u = __float_as_uint(x);
                const uint32_t code = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
                word |= (code & 0xffffu) << (16 * t);
            }
        } else {
            float y[4];
            for (int t = 0; t < 4; ++t) {
                const int i = 4 * w + t;
                y[t] = i < dim ? fminf(fmaxf(r[i] / s, -448.0f), 448.0f) : 0.0f;
            }
            int lo = __builtin_amdgcn_cvt_pk_fp8_f32(y[0], y[1], 0, false);
            word = (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(y[2], y[3], lo, true);
        }
        o[w] = word;
    }
}

// squared norm per base row with one sequential fmaf chain (data.h:54 as the reference
// build contracts it), the `lena` of the cosine distance
__global__ void norms_kernel(const float *data, uint64_t n, int dim, int ld, float *out) {
    uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const float *r = data + v * (uint64_t)ld;
    float s = 0.0f;
    for (int i = 0; i < dim; ++i) s = fmaf(r[i], r[i], s);
    out[v] = s;
}

// integer-valued test and range of a float array: out[0] = 1 if any non-integer,
// out[1] / out[2] = min / max as orderable bits
__global__ void range_kernel(const float *x, uint64_t count, unsigned *out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned nonint = 0, mn = 0xffffffffu, mx = 0u;
    for (; i < count; i += stride) {
        float v = x[i];
        if (v != rintf(v)) nonint = 1;
        unsigned o = ord32(v);
        mn = min(mn, o);
        mx = max(mx, o);
    }
    if (nonint) atomicOr(&out[0], 1u);
    atomicMin(&out[1], mn);
    atomicMax(&out[2], mx);
}

static float unord_host(unsigned o) {
    unsigned u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

int song::device_range(const float *d, uint64_t count, int *integer, float *vmin, float *vmax, hipStream_t st) {
    unsigned init[3] = {0u, 0xffffffffu, 0u}, res[3];
    unsigned *d_out = nullptr;
    HIP_TRY(hipMalloc(&d_out, sizeof(init)));
    HIP_TRY(hipMemcpyAsync(d_out, init, sizeof(init), hipMemcpyHostToDevice, st));
    int blocks = (int)std::min<uint64_t>(4096, (count + 255) / 256 + 1);
    hipLaunchKernelGGL(range_kernel, dim3(blocks), dim3(256), 0, st, d, count, d_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(res, d_out, sizeof(res), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipFree(d_out));
    *integer = res[0] == 0u;
    *vmin = unord_host(res[1]);
    *vmax = unord_host(res[2]);
    return SONG_OK;
}

extern "C" int song_index_create(const float *data, uint64_t n, int dim, const uint64_t *graph, int shift,
                                 int metric, song_index **out) {
    if (!out) return fail(SONG_E_ARG, "out is null");
    *out = nullptr;
    if (!data || !graph) return fail(SONG_E_ARG, "data/graph is null");
    if (n == 0 || n >= 0xFFFFFFFFull) return fail(SONG_E_ARG, "n=%llu out of range", (unsigned long long)n);
    if (dim < 1 || dim > 16384) return fail(SONG_E_ARG, "dim=%d out of range", dim);
    if (shift < 1 || shift > 6) return fail(SONG_E_ARG, "shift=%d not in [1,6]", shift);
    if (metric < SONG_L2 || metric > SONG_COS) return fail(SONG_E_ARG, "metric=%d", metric);
    const int slots = 1 << shift;
    const int gstride = slots <= 32 ? 32 : 64;
    std::vector<uint32_t> g32((size_t)n * gstride, 0u);
    for (uint64_t v = 0; v < n; ++v) {
        const uint64_t *r = graph + v * (uint64_t)slots;
        uint64_t deg = r[0];
        if (deg >= (uint64_t)slots)
            return fail(SONG_E_ARG, "vertex %llu: degree %llu >= %d", (unsigned long long)v,
                        (unsigned long long)deg, slots);
        uint32_t *o = &g32[(size_t)v * gstride];
        o[0] = (uint32_t)deg;
        for (uint64_t j = 1; j <= deg; ++j) {
            if (r[j] >= n)
                return fail(SONG_E_ARG, "vertex %llu: neighbour id %llu >= n", (unsigned long long)v,
                            (unsigned long long)r[j]);
            o[j] = (uint32_t)r[j];
        }
    }

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SONG_E_NODEV, "no GPU visible");

    song_index *ix = new song_index();
    ix->n = n;
    ix->dim = dim;
    ix->ld = (dim + 3) & ~3;
    ix->metric = metric;
    ix->shift = shift;
    ix->gstride = gstride;
    ix->idbits = n <= 1 ? 1 : 64 - __builtin_clzll((unsigned long long)(n - 1));
    auto cleanup = [&](int code) {
        song_index_destroy(ix);
        return code;
    };
    if (hipGetDevice(&ix->device) != hipSuccess) return cleanup(fail(SONG_E_HIP, "hipGetDevice failed"));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ix->device) == hipSuccess) {
        ix->max_lds = std::max<size_t>(65536, prop.sharedMemPerBlock);
        ix->n_cu = std::max(1, prop.multiProcessorCount);
    }

    size_t dbytes = (size_t)n * ix->ld * sizeof(float);
    if (hipMalloc(&ix->d_data, dbytes) != hipSuccess) return cleanup(fail(SONG_E_HIP, "hipMalloc data %zu", dbytes));
    if (hipMemset(ix->d_data, 0, dbytes) != hipSuccess) return cleanup(fail(SONG_E_HIP, "hipMemset data"));
    if (hipMemcpy2D(ix->d_data, ix->ld * sizeof(float), data, dim * sizeof(float), dim * sizeof(float), n,
                    hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(SONG_E_HIP, "upload data"));
    size_t gbytes = g32.size() * sizeof(uint32_t);
    if (hipMalloc(&ix->d_graph, gbytes) != hipSuccess) return cleanup(fail(SONG_E_HIP, "hipMalloc graph"));
    if (hipMemcpy(ix->d_graph, g32.data(), gbytes, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(SONG_E_HIP, "upload graph"));
    if (metric == SONG_COS) {
        if (hipMalloc(&ix->d_norms, n * sizeof(float)) != hipSuccess) return cleanup(fail(SONG_E_HIP, "hipMalloc norms"));
        hipLaunchKernelGGL(norms_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, ix->d_data, n, dim,
                           ix->ld, ix->d_norms);
        if (hipGetLastError() != hipSuccess) return cleanup(fail(SONG_E_HIP, "norms kernel"));
    }
    int rc = device_range(ix->d_data, (uint64_t)n * ix->ld, &ix->integer_valued, &ix->vmin, &ix->vmax, 0);
    if (rc != SONG_OK) return cleanup(rc);
    if (hipDeviceSynchronize() != hipSuccess) return cleanup(fail(SONG_E_HIP, "index upload sync"));
    *out = ix;
    return SONG_OK;
}

extern "C" int song_index_add_storage(song_index *ix, int store) {
    if (!ix) return fail(SONG_E_ARG, "index is null");
    if (store != SONG_STORE_BF16 && store != SONG_STORE_FP8) return fail(SONG_E_ARG, "storage=%d", store);
    if (ix->d_lp[store]) return SONG_OK;
    const int epu = store == SONG_STORE_BF16 ? 8 : 16;
    const size_t esz = store == SONG_STORE_BF16 ? 2 : 1;
    const size_t stride = (size_t)((ix->dim + epu - 1) / epu) * epu * esz;
    HIP_TRY(hipMalloc(&ix->d_lp[store], stride * ix->n));
    HIP_TRY(hipMalloc(&ix->d_scale[store], sizeof(float) * ix->n));
    hipLaunchKernelGGL(quantize_kernel, dim3((unsigned)((ix->n + 255) / 256)), dim3(256), 0, 0, ix->d_data, ix->n,
                       ix->dim, ix->ld, store, ix->d_lp[store], stride, ix->d_scale[store]);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    ix->lp_stride[store] = stride;
    return SONG_OK;
}

extern "C" int song_index_get_storage(const song_index *ix, int store, void *codes, float *scales,
                                      uint64_t *row_bytes) {
    if (!ix || store < SONG_STORE_BF16 || store > SONG_STORE_FP8 || !ix->d_lp[store])
        return fail(SONG_E_ARG, "storage %d not built", store);
    if (row_bytes) *row_bytes = ix->lp_stride[store];
    if (codes) HIP_TRY(hipMemcpy(codes, ix->d_lp[store], ix->lp_stride[store] * ix->n, hipMemcpyDeviceToHost));
    if (scales) HIP_TRY(hipMemcpy(scales, ix->d_scale[store], sizeof(float) * ix->n, hipMemcpyDeviceToHost));
    return SONG_OK;
}

extern "C" int song_index_destroy(song_index *ix) {
    if (!ix) return SONG_OK;
    for (int i = 0; i < 3; ++i) {
        if (ix->d_lp[i]) (void)hipFree(ix->d_lp[i]);
        if (ix->d_scale[i]) (void)hipFree(ix->d_scale[i]);
    }
    if (ix->d_data) (void)hipFree(ix->d_data);
    if (ix->d_graph) (void)hipFree(ix->d_graph);
    if (ix->d_norms) (void)hipFree(ix->d_norms);
    for (auto &kv : ix->spill) {
        if (kv.second.d) (void)hipFree(kv.second.d);
        if (kv.second.bag) (void)hipFree(kv.second.bag);
    }
    delete ix;
    return SONG_OK;
}

extern "C" int song_index_integer_valued(const song_index *ix) { return ix ? ix->integer_valued : 0; }

// ---------------------------------------------------------------- launch planning

// must match the per-wave layout of search_query: visited (2^hlog slots of 4 or, compact, 2
// bytes) | head merge scratch (64 keys) | best (pq rounded to even; not with bglob) | 256
// candidate ids / radix bins | query (the head is in registers, the bag in global memory)
size_t song::lds_bytes(bool exact, int pq, int qcap, int hlog, int qld, int vis16, int bloom_k, int bag_lds,
                       int bglob) {
    size_t key = exact ? 16 : 8;
    const size_t vis = bloom_k > 0 ? ((size_t)4 << hlog) : ((size_t)(vis16 ? 2 : 4) << hlog);
    return vis + (size_t)64 * key + (bglob ? 0 : (size_t)((pq + 1) & ~1) * key) + 256 * 4 + (size_t)qld * 4 +
           (bag_lds ? (size_t)qcap * key : 0);
}

// auto_tune_bloom.cc:65-86 restated: n = 12.897 pq expected insertions; the smallest filter of
// m = 64 * 2^i bits (i = 1..15) whose false-positive rate p = (1 - e^(-k/r))^k, r = m / n,
// k = round(ln 2 * r) (auto_tune_bloom.cc:6-37) is below 0.03, else i = 15.  Returned as
// 2^(i + 1) 32-bit words and k (at most 9, the kernel's limit).
void song::bloom_size(int pq, int *hlog32, int *k) {
    const double n = (double)pq * 12.897;
    int bi = 15;
    double bk = 0;
    for (int i = 1; i <= 15; ++i) {
        const double m = 64.0 * (double)(1 << i);
        const double r = m / n;
        const double kk = std::round(std::log(2.0) * r);
        const double p = std::pow(1.0 - std::exp(-kk / r), kk);
        bi = i;
        bk = kk;
        if (p < 0.03) break;
    }
    *hlog32 = bi + 1;
    *k = (int)std::max(1.0, std::min(9.0, bk));
}

extern "C" int song_bloom_size(int pq, int *hash_log2, int *bloom_k) {
    if (pq < 1 || !hash_log2 || !bloom_k) return fail(SONG_E_ARG, "pq=%d", pq);
    bloom_size(pq, hash_log2, bloom_k);
    return SONG_OK;
}

size_t song::bag_stride(const Plan &p) {
    const size_t key = p.exact ? 16 : 8;
    const size_t bag = ((size_t)p.qcap * key + 255) & ~(size_t)255;  // as bag_bytes in the kernel
    return p.bglob ? bag + (((size_t)p.pq * key + 255) & ~(size_t)255) : bag;
}

size_t song::group_lds_bytes(bool exact, int pq, int slack, int hlog, int qld, int vis16, int lanes) {
    const size_t key = exact ? 16 : 8;
    const size_t qc = (size_t)pq + (size_t)slack + 2 * (size_t)lanes;
    return ((size_t)(vis16 ? 2 : 4) << hlog) + qc * key + (size_t)((pq + 1) & ~1) * key + 64 * key + 64 * 4 +
           (size_t)qld * 4;
}

int song::make_plan(const song_index *ix, bool exact, int store, int pq, int slack, int hlog, Plan *p, int shape,
                    int visited_bits, int bloom_k, int queries_per_wave, int bag_lds, int waves_per_query,
                    int best_in_global, uint64_t nq_hint) {
    if (pq > 4096) return fail(SONG_E_ARG, "pq=%d exceeds 4096", pq);
    if (queries_per_wave != 1 && queries_per_wave != 2 && queries_per_wave != 4)
        return fail(SONG_E_ARG, "queries_per_wave=%d not 1, 2 or 4", queries_per_wave);
    if (queries_per_wave > 1 && bloom_k != 0)
        return fail(SONG_E_ARG, "several queries per wave use the exact visited table (bloom_k must be 0)");
    if (queries_per_wave > 1 && ix->gstride != 32)
        return fail(SONG_E_ARG, "several queries per wave need graph rows of 32 slots (shift <= 5)");
    p->lanes = 64 / queries_per_wave;
    if (bag_lds && (queries_per_wave > 1 || store != SONG_STORE_F32))
        return fail(SONG_E_ARG, "an LDS bag needs one query per wave and float32 rows");
    p->bag_lds = bag_lds ? 1 : 0;
    if (waves_per_query != 1 && waves_per_query != 2)
        return fail(SONG_E_ARG, "waves_per_query=%d not 1 or 2", waves_per_query);
    if (waves_per_query == 2 && (exact || store != SONG_STORE_F32 || bloom_k != 0 || queries_per_wave > 1 || bag_lds))
        return fail(SONG_E_ARG, "two waves per query need FP32 mode on float32 rows, the exact visited table, one "
                                "query per wave and the bag in global memory");
    p->duo = waves_per_query == 2 ? 1 : 0;
    if (best_in_global < -1 || best_in_global > 1) return fail(SONG_E_ARG, "best_in_global=%d", best_in_global);
    const bool bglob_ok = queries_per_wave == 1 && !bag_lds && store == SONG_STORE_F32;
    if (best_in_global == 1 && !bglob_ok)
        return fail(SONG_E_ARG, "the best list in global memory needs one query per wave, float32 rows and the bag "
                                "in global memory");
    p->pq = pq;
    if (bloom_k < -1 || bloom_k > 9) return fail(SONG_E_ARG, "bloom_k=%d not in [-1,9]", bloom_k);
    if (bloom_k != 0) {
        int bh = 0, bk = 0;
        bloom_size(pq, &bh, &bk);
        if (hlog <= 0) hlog = bh;
        if (bloom_k < 0) bloom_k = bk;
        if (hlog < 1 || hlog > 15) return fail(SONG_E_ARG, "Bloom filter of 2^%d words", hlog);
    }
    p->bloom_k = bloom_k;
    if (store != SONG_STORE_F32 && !ix->d_lp[store]) return fail(SONG_E_ARG, "storage %d not built", store);
    p->exact = exact && store == SONG_STORE_F32;
    p->store = store;
    const int units = store == SONG_STORE_F32 ? ix->ld / 4 : (int)(ix->lp_stride[store] / 16);
    p->shape = shape >= 0 ? std::min(shape, 4) : shape_for_units(units);
    p->hlog = hlog;
    p->slack = slack;
    // queue buffer: pq + slack live entries, 64 of merge room, 64 of head room
    p->qcap = pq + slack + 128;
    p->qld = (ix->dim + 15) & ~15;
    // compact visited slots when an id's bits beyond the group index (groups of 8) number 0..13
    const int lob = ix->idbits - (hlog - 3);
    p->vis16 = (bloom_k == 0 && visited_bits != 32 && hlog >= 4 && lob >= 0 && lob <= 13) ? 1 : 0;
    p->bglob = best_in_global == 1 ? 1 : 0;
    if (best_in_global < 0 && bglob_ok) {
        const size_t lds_front = (lds_bytes(p->exact, pq, p->qcap, hlog, p->qld, p->vis16, bloom_k, 0, 0) + 15) & ~15;
        const uint64_t per_cu = std::max<uint64_t>(1, ix->max_lds / lds_front);
        p->bglob = (nq_hint == 0 || nq_hint > per_cu * (uint64_t)ix->n_cu) ? 1 : 0;
    }
    if (p->lanes < 64) {
        p->shape = std::min(p->shape, 3);  // a row's lanes must lie in one sub-group
        p->lds = group_lds_bytes(p->exact, pq, slack, hlog, p->qld, p->vis16, p->lanes);
    } else {
        p->lds = lds_bytes(p->exact, pq, p->qcap, hlog, p->qld, p->vis16, bloom_k, p->bag_lds, p->bglob);
    }
    if (p->lds > ix->max_lds)
        return fail(SONG_E_OVERFLOW, "search needs %zu bytes of LDS per query (> %zu)", p->lds, ix->max_lds);
    return SONG_OK;
}

// LDS visited-table size by budget, the fastest on the 1M x 128 workload: 2^12 slots (8 KB with
// the compact 16-bit table) keep the ~1-3k vertices a search discovers up to pq ~ 1000 in LDS
// at 3/4 load; the few queries that discover more continue in the global table, and queries
// whose global table overflows are re-run.
int song::default_hlog(int pq) { return pq > 1024 ? 13 : 12; }


SearchFn song::search_fn(int metric, const Plan &p, bool persistent) {
    if (p.duo && persistent) {
        switch (metric) {
        case SONG_L2: return duo_fn_l2(p.shape, p.bglob);
        case SONG_IP: return duo_fn_ip(p.shape, p.bglob);
        default: return duo_fn_cos(p.shape, p.bglob);
        }
    }
    if (p.lanes < 64) {
        switch (metric) {
        case SONG_L2: return group_fn_l2(p.exact, p.store, p.shape, p.lanes);
        case SONG_IP: return group_fn_ip(p.exact, p.store, p.shape, p.lanes);
        default: return group_fn_cos(p.exact, p.store, p.shape, p.lanes);
        }
    }
    const bool bagl = p.bag_lds && persistent;
    switch (metric) {
    case SONG_L2: return search_fn_l2(p.exact, p.store, p.shape, persistent, bagl, p.bglob);
    case SONG_IP: return search_fn_ip(p.exact, p.store, p.shape, persistent, bagl, p.bglob);
    default: return search_fn_cos(p.exact, p.store, p.shape, persistent, bagl, p.bglob);
    }
}

static int launch(song_index *ix, const Plan &p, SearchParams sp, unsigned count, hipStream_t st,
                  uint32_t *gvis) {
    SearchFn fn = search_fn(ix->metric, p);
    const size_t stride = bag_stride(p);
    int rc = ensure_gbag(ix, (void *)st, std::min(count, kChunk), stride, &sp.gbag);
    if (rc != SONG_OK) return rc;
    sp.gbag_stride = stride;
    sp.hlog = p.hlog;
    sp.slack = p.slack;
    sp.qcap = p.qcap;
    sp.gvis = gvis;
    sp.ghlog = kGhlog;
    sp.qld = p.qld;
    sp.wave_lds = (uint32_t)p.lds;
    sp.vis16 = p.vis16;
    sp.bloom_k = p.bloom_k;
    sp.idbits = ix->idbits;
    sp.data_lp = ix->d_lp[p.store];
    sp.scale = ix->d_scale[p.store];
    sp.lp_stride = ix->lp_stride[p.store];
    if (p.lds > 65536)
        HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds));
    for (unsigned q0 = 0; q0 < count; q0 += kChunk) {
        sp.qbase = (int)q0;
        hipLaunchKernelGGL(fn, dim3(std::min(kChunk, count - q0)), dim3(64), p.lds, st, sp);
        HIP_TRY(hipGetLastError());
    }
    return SONG_OK;
}

// FP32 accumulation is exact when every product and every partial sum is an integer
// below 2^24: integer-valued data and queries with dim * span^2 < 2^24 (L2) or
// dim * max|x| * max|q| < 2^24 (IP).
bool song::fp32_exact(const song_index *ix, int q_integer, float qmin, float qmax) {
    if (!ix->integer_valued || !q_integer || ix->metric == SONG_COS) return false;
    double lim = 16777216.0;
    if (ix->metric == SONG_L2) {
        double span = std::max((double)ix->vmax, (double)qmax) - std::min((double)ix->vmin, (double)qmin);
        return (double)ix->dim * span * span < lim;
    }
    double ax = std::max(std::fabs((double)ix->vmin), std::fabs((double)ix->vmax));
    double aq = std::max(std::fabs((double)qmin), std::fabs((double)qmax));
    return (double)ix->dim * ax * aq < lim;
}

extern "C" int song_search_device(song_index *ix, const float *d_queries, uint64_t nq, int pq, int k,
                                  const int *d_pq, const int *d_k, uint32_t *d_ids, float *d_dist, int k_ld,
                                  uint32_t *d_flags, uint32_t *d_stats, const song_search_opts *opts_in,
                                  void *stream) {
    song_query_params qp{};
    qp.pq = d_pq;
    qp.k = d_k;
    qp.n_entry = 1;
    return song_search_device_ex(ix, d_queries, nq, pq, k, &qp, d_ids, d_dist, k_ld, d_flags, d_stats, opts_in,
                                 stream);
}

extern "C" int song_search_device_ex(song_index *ix, const float *d_queries, uint64_t nq, int pq, int k,
                                     const song_query_params *qp_in, uint32_t *d_ids, float *d_dist, int k_ld,
                                     uint32_t *d_flags, uint32_t *d_stats, const song_search_opts *opts_in,
                                     void *stream) {
    song_query_params qp{};
    qp.n_entry = 1;
    if (qp_in) qp = *qp_in;
    const int *d_pq = qp.pq, *d_k = qp.k;
    const uint32_t *d_entry = qp.entry;
    const int n_entry = qp.entry ? qp.n_entry : 1, *d_steps = qp.steps;
    if (n_entry < 1 || n_entry > 64) return fail(SONG_E_ARG, "n_entry=%d not in [1,64]", n_entry);
    const uint64_t nrun = qp.subset ? qp.n_subset : nq;
    if (!ix) return fail(SONG_E_ARG, "index is null");
    if (nq == 0) return SONG_OK;
    if (!d_queries || !d_ids) return fail(SONG_E_ARG, "queries/ids are null");
    if (pq < 1 || k < 1 || k_ld < k) return fail(SONG_E_ARG, "pq=%d k=%d k_ld=%d", pq, k, k_ld);
    if (nq > 0x7fffffffull) return fail(SONG_E_ARG, "nq too large");
    song_search_opts opts;
    song_search_opts_default(&opts);
    if (opts_in) opts = *opts_in;
    if (opts.slack < 0) opts.slack = 0;
    if (opts.max_steps <= 0) opts.max_steps = 1000000;
    hipStream_t st = (hipStream_t)stream;

    if (opts.storage < SONG_STORE_F32 || opts.storage > SONG_STORE_FP8)
        return fail(SONG_E_ARG, "storage=%d", opts.storage);
    if (opts.visited_bits != 0 && opts.visited_bits != 16 && opts.visited_bits != 32)
        return fail(SONG_E_ARG, "visited_bits=%d not 0, 16 or 32", opts.visited_bits);
    bool exact = true;
    if (opts.storage != SONG_STORE_F32 || opts.mode == SONG_MODE_FP32) {
        exact = false;
    } else if (opts.mode == SONG_MODE_AUTO && ix->integer_valued && ix->metric != SONG_COS) {
        int qint = 0;
        float qmin = 0.f, qmax = 0.f;
        int rc = device_range(d_queries, nq * (uint64_t)ix->dim, &qint, &qmin, &qmax, st);
        if (rc != SONG_OK) return rc;
        exact = !fp32_exact(ix, qint, qmin, qmax);
    }

    uint32_t *gvis = nullptr;
    int rcg = ensure_gvis(ix, stream, (size_t)std::min<uint64_t>(nrun ? nrun : 1, kChunk), &gvis);
    if (rcg != SONG_OK) return rcg;
    int hlog = opts.hash_log2 > 0 ? opts.hash_log2 : default_hlog(pq);
    hlog = std::max(8, std::min(hlog, 16));
    if (opts.bloom_k != 0) hlog = opts.hash_log2 > 0 ? std::min(opts.hash_log2, 15) : 0;  // 0: sized for pq
    Plan plan;
    int rc = make_plan(ix, exact, opts.storage, pq, opts.slack, hlog, &plan, -1, opts.visited_bits, opts.bloom_k, 1, 0,
                       1, opts.best_in_global, nrun);
    if (rc != SONG_OK) return rc;

    uint32_t *flags = d_flags;
    uint32_t *tmp_flags = nullptr;
    if (!flags && opts.retry) {
        HIP_TRY(hipMalloc(&tmp_flags, nq * sizeof(uint32_t)));
        flags = tmp_flags;
    }
    SearchParams sp{};
    sp.data = ix->d_data;
    sp.graph = ix->d_graph;
    sp.norms = ix->d_norms;
    sp.queries = d_queries;
    sp.entry = d_entry;
    sp.n_entry = n_entry;
    sp.steps_arr = d_steps;
    sp.pq_arr = d_pq;
    sp.k_arr = d_k;
    sp.qidx = qp.subset;
    sp.out_ids = d_ids;
    sp.out_dist = d_dist;
    sp.out_flags = flags;
    sp.out_stats = d_stats;
    sp.n = ix->n;
    sp.dim = ix->dim;
    sp.ld = ix->ld;
    sp.gstride = ix->gstride;
    sp.nq = (int)nq;
    sp.pq = pq;
    sp.k = k;
    sp.out_ld = k_ld;
    sp.max_steps = opts.max_steps;
    sp.rerank = opts.rerank;
    if (nrun == 0) {
        if (tmp_flags) (void)hipFree(tmp_flags);
        return SONG_OK;
    }
    rc = launch(ix, plan, sp, (unsigned)nrun, st, gvis);
    if (rc == SONG_OK && opts.retry) {
        // re-run queries whose LDS tables overflowed, with 4x larger tables, until none remain
        std::vector<uint32_t> hf(nq);
        std::vector<uint32_t> run_ids;
        if (qp.subset) {
            run_ids.resize(nrun);
            if (hipMemcpyAsync(run_ids.data(), qp.subset, nrun * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
                rc = fail(SONG_E_HIP, "subset readback");
        }
        uint32_t *d_idx = nullptr;
        for (int round = 0; rc == SONG_OK; ++round) {
            if (hipMemcpyAsync(hf.data(), flags, nq * sizeof(uint32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess) {
                rc = fail(SONG_E_HIP, "flag readback");
                break;
            }
            std::vector<uint32_t> redo;
            bool vis = false, que = false;
            for (uint64_t j = 0; j < nrun; ++j) {
                const uint64_t i = run_ids.empty() ? j : run_ids[j];
                if (hf[i] & (SONG_FLAG_VISITED_OVF | SONG_FLAG_QUEUE_OVF)) {
                    redo.push_back((uint32_t)i);
                    vis |= (hf[i] & SONG_FLAG_VISITED_OVF) != 0;
                    que |= (hf[i] & SONG_FLAG_QUEUE_OVF) != 0;
                }
            }
            if (redo.empty()) break;
            if (round == 4) {
                rc = fail(SONG_E_OVERFLOW, "%zu queries still overflow after 4 re-runs", redo.size());
                break;
            }
            Plan p2;
            int h2 = vis ? plan.hlog + 2 : plan.hlog;
            int s2 = que ? plan.slack * 4 + 64 : plan.slack;
            rc = make_plan(ix, exact, opts.storage, pq, s2, std::min(h2, plan.bloom_k ? 15 : 16), &p2, -1,
                           opts.visited_bits, plan.bloom_k, 1, 0, 1, opts.best_in_global, redo.size());
            if (rc != SONG_OK) break;
            plan = p2;
            if (!d_idx && hipMalloc(&d_idx, nq * sizeof(uint32_t)) != hipSuccess) {
                rc = fail(SONG_E_HIP, "hipMalloc retry index");
                break;
            }
            if (hipMemcpyAsync(d_idx, redo.data(), redo.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess) {
                rc = fail(SONG_E_HIP, "retry index upload");
                break;
            }
            SearchParams sp2 = sp;
            sp2.qidx = d_idx;
            rc = launch(ix, plan, sp2, (unsigned)redo.size(), st, gvis);
            if (rc != SONG_OK) break;
        }
        if (d_idx) (void)hipFree(d_idx);
    }
    if (tmp_flags) (void)hipFree(tmp_flags);
    return rc;
}

extern "C" int song_search(song_index *ix, const float *queries, uint64_t nq, int pq, int k, uint64_t *out_ids,
                           float *out_dist, uint32_t *out_flags, const song_search_opts *opts) {
    if (!ix) return fail(SONG_E_ARG, "index is null");
    if (nq == 0) return SONG_OK;
    if (!queries || !out_ids) return fail(SONG_E_ARG, "queries/out_ids are null");
    if (pq < 1 || k < 1) return fail(SONG_E_ARG, "pq=%d k=%d", pq, k);
    // rows keep the caller's stride k; a search returns at most pq entries, so columns
    // [pq, k) are padding (UINT64_MAX, +inf) like a list shorter than k
    const int kk = std::min(k, pq);
    float *d_q = nullptr, *d_d = nullptr;
    uint32_t *d_i = nullptr, *d_f = nullptr;
    int rc = SONG_OK;
    size_t qb = nq * (size_t)ix->dim * sizeof(float), ib = nq * (size_t)k * sizeof(uint32_t);
    if (hipMalloc(&d_q, qb) != hipSuccess || hipMalloc(&d_i, ib) != hipSuccess ||
        hipMalloc(&d_f, nq * sizeof(uint32_t)) != hipSuccess || (out_dist && hipMalloc(&d_d, ib) != hipSuccess)) {
        rc = fail(SONG_E_HIP, "hipMalloc search buffers");
    }
    if (rc == SONG_OK && hipMemcpy(d_q, queries, qb, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(SONG_E_HIP, "upload queries");
    if (rc == SONG_OK) rc = song_search_device(ix, d_q, nq, pq, kk, nullptr, nullptr, d_i, d_d, k, d_f, nullptr, opts, nullptr);
    if (rc == SONG_OK && hipDeviceSynchronize() != hipSuccess) rc = fail(SONG_E_HIP, "search sync");
    if (rc == SONG_OK) {
        std::vector<uint32_t> ids(nq * (size_t)k);
        if (hipMemcpy(ids.data(), d_i, ib, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(SONG_E_HIP, "download ids");
        for (size_t i = 0; rc == SONG_OK && i < ids.size(); ++i)
            out_ids[i] = (ids[i] == 0xFFFFFFFFu || (int)(i % (size_t)k) >= kk) ? UINT64_MAX : (uint64_t)ids[i];
        if (rc == SONG_OK && out_dist && hipMemcpy(out_dist, d_d, ib, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(SONG_E_HIP, "download dists");
        for (size_t i = 0; rc == SONG_OK && out_dist && kk < k && i < ids.size(); ++i)
            if ((int)(i % (size_t)k) >= kk) out_dist[i] = INFINITY;
        if (rc == SONG_OK && out_flags &&
            hipMemcpy(out_flags, d_f, nq * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(SONG_E_HIP, "download flags");
    }
    if (d_q) (void)hipFree(d_q);
    if (d_i) (void)hipFree(d_i);
    if (d_d) (void)hipFree(d_d);
    if (d_f) (void)hipFree(d_f);
    return rc;
}

extern "C" int song_astar_multi_start_search_batch(const float *h_queries, uint64_t nq, int pq, uint64_t *results,
                                                   const float *h_data, const uint64_t *h_graph,
                                                   int vertex_offset_shift, uint64_t num, int dim, int metric) {
    song_index *ix = nullptr;
    int rc = song_index_create(h_data, num, dim, h_graph, vertex_offset_shift, metric, &ix);
    if (rc != SONG_OK) return rc;
    song_search_opts o;
    song_search_opts_default(&o);
    o.mode = SONG_MODE_EXACT;
    rc = song_search(ix, h_queries, nq, pq, pq, results, nullptr, nullptr, &o);
    song_index_destroy(ix);
    return rc;
}
