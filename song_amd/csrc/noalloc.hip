// song_search_device_noalloc (include/song_amd.h): the plain / per-query search of
// song_search_device_ex with no device allocation, for callers that cannot allocate.
//
// Everything the search kernels need beyond the caller's inputs and outputs -- the queue
// bag and the visited spill table of every wave slot, and the persistent grid's work
// counter -- lives in `scratch`, one static device array per GPU.  Host code never passes
// the array's name to a kernel: the name is a host-side stand-in.  Its device address is
// taken once per device with hipGetSymbolAddress (on the first call with that device
// current) and that address is what the kernels receive.  The searches run on the
// persistent kernel (waves take queries from the work counter) with as many wave slots as
// the array holds for the call's configuration.  Every call on a GPU shares that GPU's
// array, so calls that overlap on one GPU race; the counter is cleared by a kernel at the
// start of every call, and the bags and spill tables are written before they are read by
// every query.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "index_internal.hpp"

using namespace song;

namespace {

constexpr size_t kScratchBytes = SONG_NOALLOC_SCRATCH_BYTES;
constexpr size_t kHeader = 256;  // work counter and count, padded to the array's alignment
constexpr int kGhlog = 13;       // spill table per wave slot: 2^13 u32, as the allocating search

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
struct DeviceScratch {
    unsigned char *ptr = nullptr;
    int n_cu = 0;
};
DeviceScratch g_dev[kMaxDevices];
std::mutex g_mu;

// the device address of this GPU's array (and its CU count), looked up once per device
int device_scratch(int dev, DeviceScratch *out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(SONG_E_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> g(g_mu);
    DeviceScratch &d = g_dev[dev];
    if (!d.ptr) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SONG_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        int ncu = 0;
        e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess || ncu < 1) return fail(SONG_E_HIP, "CU count: %s", hipGetErrorString(e));
        d.ptr = static_cast<unsigned char *>(p);
        d.n_cu = ncu;
    }
    *out = d;
    return SONG_OK;
}

// work_next = 0, work_count = count (a kernel, not a memset, so the call stays capturable)
__global__ void noalloc_init_kernel(uint32_t *ctr, uint32_t count) {
    if (threadIdx.x == 0) {
        ctr[0] = 0u;
        ctr[1] = count;
    }
}

}  // namespace

extern "C" int song_search_device_noalloc(song_index *ix, const float *d_queries, uint64_t nq, int pq, int k,
                                          const song_query_params *qp_in, uint32_t *d_ids, float *d_dist, int k_ld,
                                          uint32_t *d_flags, uint32_t *d_stats, const song_search_opts *opts_in,
                                          void *stream) {
    if (!ix) return fail(SONG_E_ARG, "index is null");
    song_query_params qp{};
    qp.n_entry = 1;
    if (qp_in) qp = *qp_in;
    const int n_entry = qp.entry ? qp.n_entry : 1;
    if (n_entry < 1 || n_entry > 64) return fail(SONG_E_ARG, "n_entry=%d not in [1,64]", n_entry);
    if (nq == 0) return SONG_OK;
    if (!d_queries || !d_ids) return fail(SONG_E_ARG, "queries/ids are null");
    if (pq < 1 || k < 1 || k_ld < k) return fail(SONG_E_ARG, "pq=%d k=%d k_ld=%d", pq, k, k_ld);
    if (nq > 0x7fffffffull) return fail(SONG_E_ARG, "nq too large");
    const uint64_t nrun = qp.subset ? qp.n_subset : nq;
    if (nrun == 0) return SONG_OK;
    song_search_opts opts;
    song_search_opts_default(&opts);
    if (opts_in) opts = *opts_in;
    if (opts.mode != SONG_MODE_EXACT && opts.mode != SONG_MODE_FP32)
        return fail(SONG_E_ARG, "the no-allocation search needs mode EXACT or FP32 (AUTO inspects the queries in "
                                "allocated memory; song_resolve_mode picks the mode beforehand)");
    if (opts.storage < SONG_STORE_F32 || opts.storage > SONG_STORE_FP8)
        return fail(SONG_E_ARG, "storage=%d", opts.storage);
    if (opts.visited_bits != 0 && opts.visited_bits != 16 && opts.visited_bits != 32)
        return fail(SONG_E_ARG, "visited_bits=%d not 0, 16 or 32", opts.visited_bits);
    if (opts.slack < 0) opts.slack = 0;
    if (opts.max_steps <= 0) opts.max_steps = 1000000;
    const bool exact = opts.storage == SONG_STORE_F32 && opts.mode == SONG_MODE_EXACT;
    int hlog = opts.hash_log2 > 0 ? opts.hash_log2 : default_hlog(pq);
    hlog = std::max(8, std::min(hlog, 16));
    if (opts.bloom_k != 0) hlog = opts.hash_log2 > 0 ? std::min(opts.hash_log2, 15) : 0;
    Plan plan;
    int rc = make_plan(ix, exact, opts.storage, pq, opts.slack, hlog, &plan, -1, opts.visited_bits, opts.bloom_k, 1, 0,
                       1, opts.best_in_global, nrun);
    if (rc != SONG_OK) return rc;

    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != ix->device) return fail(SONG_E_ARG, "the index lives on device %d, the current device is %d", ix->device, dev);
    DeviceScratch ds;
    rc = device_scratch(dev, &ds);
    if (rc != SONG_OK) return rc;

    // wave slots: one wave per workgroup, as many as fit the LDS of a CU and the array
    SearchFn fn = search_fn(ix->metric, plan, true);
    const size_t wave_lds = (plan.lds + 15) & ~(size_t)15;
    if (wave_lds > 65536)
        HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)wave_lds));
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, 64, wave_lds));
    if (per_cu < 1) return fail(SONG_E_OVERFLOW, "no wave of %zu LDS bytes fits a CU", wave_lds);
    const size_t bag = bag_stride(plan);
    const size_t spill = (size_t)4 << kGhlog;
    const size_t per_slot = bag + spill;
    const size_t fit = (kScratchBytes - kHeader) / per_slot;
    if (fit < (size_t)ds.n_cu)
        return fail(SONG_E_ARG,
                    "pq=%d slack=%d needs %zu scratch bytes per wave slot: fewer than one slot per CU fit the "
                    "%zu-byte no-allocation array (use song_search_device_ex)",
                    pq, opts.slack, per_slot, kScratchBytes);
    const size_t slots = std::min({fit, (size_t)per_cu * (size_t)ds.n_cu, (size_t)nrun});

    uint32_t *ctr = reinterpret_cast<uint32_t *>(ds.ptr);
    unsigned char *bags = ds.ptr + kHeader;
    uint32_t *gvis = reinterpret_cast<uint32_t *>(bags + slots * bag);

    SearchParams sp{};
    sp.data = ix->d_data;
    sp.graph = ix->d_graph;
    sp.norms = ix->d_norms;
    sp.queries = d_queries;
    sp.entry = qp.entry;
    sp.n_entry = n_entry;
    sp.steps_arr = qp.steps;
    sp.pq_arr = qp.pq;
    sp.k_arr = qp.k;
    sp.qidx = qp.subset;
    sp.out_ids = d_ids;
    sp.out_dist = d_dist;
    sp.out_flags = d_flags;
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
    sp.hlog = plan.hlog;
    sp.slack = plan.slack;
    sp.qcap = plan.qcap;
    sp.qld = plan.qld;
    sp.gvis = gvis;
    sp.ghlog = kGhlog;
    sp.wave_lds = (uint32_t)wave_lds;
    sp.vis16 = plan.vis16;
    sp.bloom_k = plan.bloom_k;
    sp.idbits = ix->idbits;
    sp.data_lp = ix->d_lp[plan.store];
    sp.scale = ix->d_scale[plan.store];
    sp.lp_stride = ix->lp_stride[plan.store];
    sp.gbag = bags;
    sp.gbag_stride = bag;
    sp.work_next = ctr;
    sp.work_count = ctr + 1;
    sp.qbase = 0;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(noalloc_init_kernel, dim3(1), dim3(64), 0, st, ctr, (uint32_t)nrun);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(fn, dim3((unsigned)slots), dim3(64), wave_lds, st, sp);
    HIP_TRY(hipGetLastError());
    return SONG_OK;
}
