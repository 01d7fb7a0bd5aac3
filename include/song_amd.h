/*
 * song_amd -- C ABI of the MI355X (gfx950) graph search.
 *
 * Plain C types only (no HIP or torch types): pointers, sizes, int codes.
 * Every entry point returns SONG_OK (0) or a negative SONG_E* code; the text
 * of the last error on the calling thread is song_last_error().
 *
 * Reference interfaces replaced (sunbelbd/song):
 *   song_index_create / song_search  <- WarpAStarAccelerator::astar_multi_start_search_batch
 *        (warp_astar_accelerator.h:347-423), split so the index is uploaded once
 *        instead of on every call (:353-356) and results come back as
 *        ascending (dist, id) lists like the CPU search (graph.h:209-215).
 *   song_astar_multi_start_search_batch <- the same static seam with its
 *        exact argument list (h_data, h_graph, shift, num, dim) for a one-shot
 *        call (uploads, searches, frees).
 *   KernelFixedDegreeGraph::search_top_k_batch (kernelgraph.h:269-272) calls
 *        either; song_amd/cli/kernelgraph.hpp is that binding written out.
 */
#ifndef SONG_AMD_H
#define SONG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SONG_OK = 0,
    SONG_E_ARG = -1,      /* invalid argument (sizes, metric, degree, id range) */
    SONG_E_HIP = -2,      /* HIP runtime error (allocation, launch, copy) */
    SONG_E_NODEV = -3,    /* no GPU visible */
    SONG_E_OVERFLOW = -4, /* a query still overflowed its LDS tables at the largest size */
};

/* Distance metrics (reference config flags -D__USE_{L2,IP,COS}_DIST, data.h:28-62). */
enum { SONG_L2 = 0, SONG_IP = 1, SONG_COS = 2 };

/*
 * Arithmetic of the search.
 *  SONG_MODE_EXACT  distances reproduce the reference CPU arithmetic bit for bit
 *                   (float products accumulated in double, data.h:28-62); ids
 *                   equal the CPU search's on the same graph/queries/pq.
 *  SONG_MODE_FP32   float accumulation; exact whenever every distance is an
 *                   integer below 2^24 (integer-valued data), else within float
 *                   rounding of the exact value.
 *  SONG_MODE_AUTO   FP32 when the index and the batch are integer-safe, else EXACT.
 */
enum { SONG_MODE_AUTO = 0, SONG_MODE_EXACT = 1, SONG_MODE_FP32 = 2 };

/* Per-query flag bits written to out_flags. */
enum {
    SONG_FLAG_VISITED_OVF = 1u, /* visited table filled (query re-run with a larger table) */
    SONG_FLAG_QUEUE_OVF = 2u,   /* tie group at the pq-th distance exceeded the slack */
    SONG_FLAG_STEP_CAP = 4u,    /* hit the 1e6 expansion cap of graph.h:183 (same as the CPU) */
    SONG_FLAG_BAD_LAUNCH = 0x80000000u, /* internal: the launch's LDS slice does not fit the
                                            kernel's layout; the query was not searched */
};

typedef struct song_index song_index;

/*
 * Row storage read by the search (precision form).  SONG_STORE_F32 reads the float32
 * rows.  BF16 / FP8 read a low-precision copy built by song_index_add_storage: code *
 * scale[row], scale a power of two for BF16 (lossless for integers up to 256) and
 * max|x| / 448 for FP8-E4M3; distances accumulate in FP32 and the final pq list is
 * re-ranked with exact FP32 distances (opts.rerank).
 */
enum { SONG_STORE_F32 = 0, SONG_STORE_BF16 = 1, SONG_STORE_FP8 = 2 };

typedef struct {
    int mode;      /* SONG_MODE_* (default AUTO); BF16/FP8 storage always accumulates in FP32 */
    int slack;     /* queue entries kept beyond pq for distance ties (default 128: fewer bag truncations) */
    int hash_log2; /* visited-table slots = 2^hash_log2; 0 = sized from pq */
    int max_steps; /* expansion cap; 0 = 1,000,000 (graph.h:183) */
    int retry;     /* re-run overflowed queries with larger tables (default 1) */
    int storage;   /* SONG_STORE_* (default F32) */
    int rerank;    /* low-precision storage: exact FP32 re-rank of the pq list (default 1) */
    int visited_bits; /* visited-table slot width: 16 (compact), 32, or 0 = compact whenever the ids
                         fit (default); results do not depend on it */
    int bloom_k;   /* visited set: 0 = the exact table (default); 1..9 = a blocked Bloom filter in its
                      place (the reference's BlockedBloomFilter, blocked_bloomfilter.h:9-82) with
                      bloom_k bits per id in one 32-bit word and 2^hash_log2 words (hash_log2 = 0:
                      sized by the reference's auto_tune_bloom closed form for pq); -1 = size and
                      bits both from that closed form.  Approximate: a false positive skips an
                      unvisited vertex, so ids may differ from the exact search (recall-judged) */
    int best_in_global; /* where the pq-best list's front (all but its last 64 entries) lives: 1 = the
                           query slot's global scratch, after its queue bag (frees pq keys of LDS per
                           wave, so more waves fit a CU; float32 rows), 0 = LDS, -1 = by the call's
                           size (default): global when its queries outnumber the waves that fit the
                           device at once with the front in LDS, else LDS (a call that fits at once
                           is latency-bound, and the front's global stores lengthen each step);
                           results do not depend on it */
} song_search_opts;

/* Fill opts with the defaults above. */
void song_search_opts_default(song_search_opts *opts);

/*
 * Upload an index once.  data: n x dim float32 row-major (bfsg.data layout,
 * data.h:164-175).  graph: n rows of (1 << shift) uint64, slot 0 = out-degree,
 * then neighbour ids (bfsg.graph layout, graph.h:259-271, README.md:84-89).
 * shift in [1,6]; degrees must be < (1 << shift) and ids < n.
 */
int song_index_create(const float *data, uint64_t n, int dim, const uint64_t *graph, int shift,
                      int metric, song_index **out);
int song_index_destroy(song_index *index);

/* 1 if every base coordinate is an integer (FP32 mode can be exact), else 0. */
int song_index_integer_valued(const song_index *index);

/* Build the BF16 or FP8-E4M3 copy of the rows (with per-row scales) on the GPU. */
int song_index_add_storage(song_index *index, int storage);

/*
 * Copy back a low-precision storage for inspection: codes n x row_bytes (row_bytes =
 * dim rounded up to 8 BF16 / 16 FP8 elements, times the element size) and scales n.
 */
int song_index_get_storage(const song_index *index, int storage, void *codes, float *scales, uint64_t *row_bytes);

/*
 * Batched search with host buffers.  queries: nq x dim float32 (dense,
 * organize_point order, data.h:106-114).  Each query returns the first k
 * entries of its pq-sized result list, ascending by (dist, id): out_ids nq x k
 * (UINT64_MAX padding when fewer were found, and in columns pq..k-1 when k > pq),
 * out_dist nq x k float32 (+inf padding; nullable), out_flags nq (nullable).
 */
int song_search(song_index *index, const float *queries, uint64_t nq, int pq, int k,
                uint64_t *out_ids, float *out_dist, uint32_t *out_flags, const song_search_opts *opts);

/*
 * Batched search with device buffers on `stream` (a hipStream_t, NULL = default).
 * d_queries nq x dim float32; d_pq / d_k: per-query budgets (nullable: use pq
 * and k for all); d_ids nq x k_ld uint32 (0xFFFFFFFF padding); d_dist nq x k_ld
 * float32 (nullable); d_flags nq uint32 (nullable); d_stats nq x 8 uint32
 * (expansions, distances computed, adjacency words read = sum of 1 + degree
 * over expanded vertices, adjacency-row prefetch hits, settle step = the last
 * expansion that changed the first k results, queue-head refills from the bag,
 * bag truncations, 0; nullable).  pq is the
 * maximum budget.
 * With opts->retry set (the default) the call reads the flags back, re-runs
 * overflowed queries and synchronizes the stream; with retry = 0 it only
 * enqueues the search and the caller checks d_flags.
 */
int song_search_device(song_index *index, const float *d_queries, uint64_t nq, int pq, int k,
                       const int *d_pq, const int *d_k, uint32_t *d_ids, float *d_dist, int k_ld,
                       uint32_t *d_flags, uint32_t *d_stats, const song_search_opts *opts, void *stream);

/*
 * Per-query parameters of song_search_device_ex (all nullable).  No counterpart in the
 * reference, which searches every query with the compiled TOPK from vertex 0
 * (warp_astar_accelerator.h:166-176, 386); used by the learned and batching forms.
 */
typedef struct {
    const int *pq;          /* nq budgets (each <= the call's pq) */
    const int *k;           /* nq result counts (each <= its pq and <= k_ld) */
    const uint32_t *entry;  /* nq x n_entry start vertices, in place of vertex 0 */
    int n_entry;            /* 1..64 */
    const int *steps;       /* nq expansion caps */
    const uint32_t *subset; /* search only these query indices (results land at those rows) */
    uint64_t n_subset;
} song_query_params;

/* song_search_device with per-query parameters (qp nullable = the plain call). */
int song_search_device_ex(song_index *index, const float *d_queries, uint64_t nq, int pq, int k,
                          const song_query_params *qp, uint32_t *d_ids, float *d_dist, int k_ld, uint32_t *d_flags,
                          uint32_t *d_stats, const song_search_opts *opts, void *stream);

/*
 * Batching form.  No counterpart in the reference, whose launch is one block per query
 * with the compiled TOPK for all (warp_astar_accelerator.h:386); it replaces that launch
 * for batches with uneven k and pq per query.
 *
 * A batch object owns every buffer its launches use, so song_batch_run only enqueues
 * kernels (no allocation, no memset, no host synchronization) and can be captured in a
 * hipGraph; song_batch_run_graph captures it once per argument set and replays it.
 * One run: a bucket kernel sorts the queries into tiers by their pq, one persistent
 * kernel per tier searches its list (waves take queries from a device work counter,
 * LDS sized for the tier; tiers on forked streams joined by events), and one
 * persistent kernel re-runs queries that overflowed with the retry configuration.  Each query's result equals song_search_device's at the
 * same (k, pq), entries, step cap, mode and storage.
 */
#define SONG_MAX_TIERS 8

enum { SONG_FLAG_BAD_PARAMS = 8u }; /* batch: pq outside every tier (query not searched) */

typedef struct {
    int pq_max;          /* tier of the queries with pq in (previous pq_max, pq_max] */
    int hash_log2;       /* visited table: 2^hash_log2 LDS slots per wave (8..16), spilling to a
                            global table; 0 = sized from pq_max */
    int slack;           /* heap depth: queue entries kept beyond pq for distance ties (>= 0) */
    int shape;           /* lanes per candidate row: 0..4 = 4x2, 4x8, 4x16, 16x16, 64x16
                            (lanes x 16-byte units); -1 = by row width */
    int waves_per_group; /* queries in flight per workgroup, one per wave: 1, 2 or 4 */
    int groups_per_cu;   /* persistent grid = groups_per_cu x compute units; 0 = all that fit */
    int visited_bits;    /* visited-table slot width: 16 (compact), 32, or 0 = compact whenever the
                            ids fit; results do not depend on it */
    int storage;         /* rows this tier reads: SONG_STORE_*, or -1 = the batch's cfg.storage.  A
                            query re-run after an overflow is searched with the retry tier's storage */
    int bloom_k;         /* visited set as song_search_opts.bloom_k (hash_log2 = the filter's words);
                            the re-run of overflowed queries always uses the exact table */
    int queries_per_wave; /* 1 (0): one query per 64-lane wave; 2 or 4: sub-groups of 32 or 16 lanes,
                             each searching its own query (the reference's N_MULTIQUERY,
                             warp_astar_accelerator.h:27-30) with its queue and best list in LDS;
                             same results; exact visited table and graph rows of 32 slots only;
                             shape at most 3.  The re-run uses one query per wave */
    int multiprobe;      /* with queries_per_wave 2 or 4: expansions per step, 1 (0) or 2 (the next queue
                            head too, its neighbours evaluated in the same gather: the reference's
                            N_MULTIPROBE, warp_astar_accelerator.h:184-197).  2 changes the expansion
                            order, so results are approximate (recall-judged) */
    int bag_in_lds;      /* one query per wave, float32 rows: 1 keeps the queue's overflow bag in the
                            wave's LDS instead of global memory (same results; for latency-bound
                            tiers with few waves per CU, whose refills then wait on LDS, not HBM) */
    int waves_per_query; /* 1 (0), or 2: each query gets a search wave and a helper wave that computes,
                            ahead of the search, the distances of the predicted next expansion's
                            unvisited neighbours into an LDS cache the search wave reads (same results;
                            FP32 mode, float32 rows, exact visited table; waves_per_group is then 1) */
    int best_in_global;  /* as song_search_opts.best_in_global, but -1 (the default) = 0, LDS: a tier's
                            share of the batch is not known when the batch is created, and the autotuner
                            sets 1 where it measures faster (one query per wave, float32 rows, bag in
                            global memory; the re-run keeps the front in LDS); same results */
} song_tier;

typedef struct {
    int mode;      /* SONG_MODE_EXACT or SONG_MODE_FP32 (song_resolve_mode picks like AUTO) */
    int storage;   /* SONG_STORE_* */
    int rerank;    /* low-precision storage: exact FP32 re-rank of the pq list */
    int max_steps; /* expansion cap of every query; 0 = 1,000,000 */
    int spill_log2; /* global visited spill table per wave: 2^spill_log2 u32 (10..18, default 14;
                       the re-run's tables are 4x larger) */
    int n_tiers;   /* 1..SONG_MAX_TIERS, ascending pq_max */
    song_tier tier[SONG_MAX_TIERS];
    song_tier retry; /* re-run of overflowed queries (pq_max ignored: the largest tier's) */
    int parallel_tiers; /* 1: tiers run on forked streams, largest pq first, so one tier's
                           workgroups fill the slots another's finished waves free; 0: in
                           order on the run's stream */
    int priority_first; /* 1: with parallel tiers, the largest-pq tier's waves run at raised
                           issue priority (its searches are the longest) */
} song_batch_config;

typedef struct song_batch song_batch;

/* One tier up to pq_max with default knobs; the re-run with 4x the filter, heap depth and
 * spill table, on a full-occupancy grid. */
void song_batch_config_default(song_batch_config *cfg, int pq_max);

/* The mode SONG_MODE_AUTO would pick for these device queries (synchronizes stream). */
int song_resolve_mode(const song_index *index, const float *d_queries, uint64_t nq, int *mode, void *stream);

/* Allocate a batch for at most nq_max queries per run. */
int song_batch_create(song_index *index, const song_batch_config *cfg, uint64_t nq_max, song_batch **out);
int song_batch_destroy(song_batch *batch);

/*
 * Enqueue one batched search on `stream` (a hipStream_t; NULL = default).  Arguments as
 * song_search_device_ex (qp nullable; qp->subset unsupported); pq / k apply to queries
 * without per-query values.  Per-query entry lists may be padded with 0xFFFFFFFF.  A
 * query whose pq fits no tier gets SONG_FLAG_BAD_PARAMS and 0xFFFFFFFF ids; one that
 * still overflows after the retry keeps its overflow flag.  d_flags is required.
 * Runs of one batch object must not overlap (they share its buffers).
 */
int song_batch_run(song_batch *batch, const float *d_queries, uint64_t nq, int pq, int k, const song_query_params *qp,
                   uint32_t *d_ids, float *d_dist, int k_ld, uint32_t *d_flags, uint32_t *d_stats, void *stream);

/* song_batch_run through a hipGraph: captured on the first call with these arguments
 * (stream must not be NULL), replayed by later calls with the same arguments. */
int song_batch_run_graph(song_batch *batch, const float *d_queries, uint64_t nq, int pq, int k,
                         const song_query_params *qp, uint32_t *d_ids, float *d_dist, int k_ld, uint32_t *d_flags,
                         uint32_t *d_stats, void *stream);

/* Queries per tier and re-run count of the last run (n_tiers + 1 values; synchronizes). */
int song_batch_counts(song_batch *batch, uint32_t *out, void *stream);

/* The Bloom filter size of the reference's auto_tune_bloom (auto_tune_bloom.cc:65-86) for a
 * budget pq: 2^hash_log2 32-bit words and bloom_k bits per id (song_search_opts.bloom_k = -1). */
int song_bloom_size(int pq, int *hash_log2, int *bloom_k);

/* Persistent grid chosen for tier t (t = n_tiers: the retry), in workgroups. */
int song_batch_grid(const song_batch *batch, int t, int *groups, int *waves_per_group, int *lds_per_wave);

/*
 * song_search_device_ex without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc).  The kernels' scratch (queue bags, visited spill
 * tables, work counter) lives in one static device array of SONG_NOALLOC_SCRATCH_BYTES on
 * each GPU the library runs on, held until the process ends; the searches run on a
 * persistent grid with as many wave slots as that array holds for the call's pq, slack and
 * mode.  Differences from song_search_device_ex: opts->mode must be EXACT or FP32 (AUTO
 * inspects the queries in allocated memory; song_resolve_mode picks beforehand),
 * opts->retry is ignored (overflowed queries keep their flags, as with retry = 0), and a
 * configuration whose per-slot scratch leaves fewer than one slot per CU is rejected with
 * SONG_E_ARG.  Results, flags and stats equal song_search_device_ex's with retry = 0.
 * Calls that overlap on one GPU race on the shared array.
 */
#define SONG_NOALLOC_SCRATCH_BYTES (64u << 20)
int song_search_device_noalloc(song_index *index, const float *d_queries, uint64_t nq, int pq, int k,
                               const song_query_params *qp, uint32_t *d_ids, float *d_dist, int k_ld,
                               uint32_t *d_flags, uint32_t *d_stats, const song_search_opts *opts, void *stream);

/*
 * One-shot replacement of WarpAStarAccelerator::astar_multi_start_search_batch
 * (warp_astar_accelerator.h:347): dense queries (nq x dim), pq = the compiled
 * TOPK of the reference, results nq x pq uint64 ascending by (dist, id).
 */
int song_astar_multi_start_search_batch(const float *h_queries, uint64_t nq, int pq, uint64_t *results,
                                        const float *h_data, const uint64_t *h_graph,
                                        int vertex_offset_shift, uint64_t num, int dim, int metric);

const char *song_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
