// Internal interface shared by the C ABI sources (song_api.hip, batch.hip): the index
// object, error reporting and launch planning.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <unordered_map>

#include "../../include/song_amd.h"
#include "search_inst.hpp"

namespace song {

// record the calling thread's last error text; returns code
int fail(int code, const char *fmt, ...);

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(SONG_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

}  // namespace song

struct song_index {
    int device = 0;
    uint64_t n = 0;
    int dim = 0, ld = 0, metric = 0, shift = 0, gstride = 0;
    int idbits = 32;  // ids are below 2^idbits
    float *d_data = nullptr;
    uint32_t *d_graph = nullptr;
    float *d_norms = nullptr;
    int integer_valued = 0;
    float vmin = 0.f, vmax = 0.f;
    size_t max_lds = 65536;
    int n_cu = 1;  // compute units of the index's device
    // visited spill tables and queue bags, one per launch slot, one set per stream: calls
    // on one stream run in order; calls on different streams use different buffers
    struct Spill {
        uint32_t *d = nullptr;
        size_t slots = 0;
        unsigned char *bag = nullptr;
        size_t bag_bytes = 0;
    };
    std::unordered_map<void *, Spill> spill;
    std::mutex spill_mu;
    uint8_t *d_lp[3] = {nullptr, nullptr, nullptr};  // low-precision rows by SONG_STORE_*
    float *d_scale[3] = {nullptr, nullptr, nullptr};
    size_t lp_stride[3] = {0, 0, 0};
};

namespace song {

// launch plan of one search configuration
struct Plan {
    bool exact;
    int store, shape, hlog, slack, qcap, qld;
    int vis16;   // compact visited table (16-bit slots)
    int bloom_k; // > 0: blocked Bloom filter (2^hlog 32-bit words, bloom_k bits per id) instead
    int lanes;   // lanes per query: 64 (one query per wave, search_query) or 16 / 32 (search_group)
    int bag_lds; // search_query's bag in LDS (qcap keys after the query) instead of global memory
    int duo;     // two waves per query (search_duo: search wave + helper wave)
    int pq;      // budget the layout is sized for
    int bglob;   // search_query's best-list front in global memory, after the slot's bag
    size_t lds;  // LDS bytes of one wave (lanes 64) or of one query's slice (lanes 16 / 32)
};

size_t lds_bytes(bool exact, int pq, int qcap, int hlog, int qld, int vis16, int bloom_k = 0, int bag_lds = 0,
                 int bglob = 0);
// bytes of one wave slot's global region: the queue bag (qcap keys, 256-byte aligned), then
// with bglob the best list's front (pq keys)
size_t bag_stride(const Plan &p);
// shape < 0: by row width; otherwise the given lanes-per-row shape
// visited_bits: 16 or 32 = slot width (16 only where the ids fit, else 32); 0 = 16 where they fit
// bloom_k: 0 = exact visited table; > 0 = blocked Bloom filter with that many bits per id and
// 2^hlog words; -1 = Bloom filter with bits (and, for hlog 0, size) from bloom_size
// queries_per_wave: 1 (search_query), 2 or 4 (search_group with 32 / 16 lanes per query; exact
// visited table only, persistent form only)
// bag_lds: the queue's bag in LDS (one query per wave, float32 rows, persistent launches)
// waves_per_query: 1, or 2 (search_duo: float32 mode, float32 rows, exact visited table, bag in global
// memory, persistent launches)
// best_in_global: 1 = the best list's front in the slot's global region (one query per wave, float32
// rows, bag in global memory), 0 = in LDS, -1 = by the launch's size: in global memory where allowed
// and the launch's nq_hint queries (0: unknown) outnumber the waves that fit the device at once with
// the front in LDS -- more waves per CU then pay for the front's global traffic; a launch that fits
// at once is latency-bound and keeps the front in LDS (same results either way)
int make_plan(const song_index *ix, bool exact, int store, int pq, int slack, int hlog, Plan *p, int shape = -1,
              int visited_bits = 0, int bloom_k = 0, int queries_per_wave = 1, int bag_lds = 0,
              int waves_per_query = 1, int best_in_global = -1, uint64_t nq_hint = 0);
// LDS bytes of one query's slice in search_group (its layout: visited | queue | best | stage | cid | qv)
size_t group_lds_bytes(bool exact, int pq, int slack, int hlog, int qld, int vis16, int lanes);
int default_hlog(int pq);
// Bloom filter sizing of the reference's auto_tune_bloom (auto_tune_bloom.cc:6-37, 65-86)
void bloom_size(int pq, int *hlog32, int *k);
SearchFn search_fn(int metric, const Plan &p, bool persistent = false);
// integer test and range of a device float array (synchronizes st)
int device_range(const float *d, uint64_t count, int *integer, float *vmin, float *vmax, hipStream_t st);
bool fp32_exact(const song_index *ix, int q_integer, float qmin, float qmax);

}  // namespace song
