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
 *        either; see INTEGRATION.md for the host-side binding.
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
    int slack;     /* queue entries kept beyond pq for distance ties (default 32) */
    int hash_log2; /* visited-table slots = 2^hash_log2; 0 = sized from pq */
    int max_steps; /* expansion cap; 0 = 1,000,000 (graph.h:183) */
    int retry;     /* re-run overflowed queries with larger tables (default 1) */
    int storage;   /* SONG_STORE_* (default F32) */
    int rerank;    /* low-precision storage: exact FP32 re-rank of the pq list (default 1) */
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
 * (k <= pq) entries of its pq-sized result list, ascending by (dist, id):
 * out_ids nq x k (UINT64_MAX padding when fewer were found), out_dist nq x k
 * float32 (nullable), out_flags nq (nullable).
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
 * expansion that changed the first k results, 0, 0, 0; nullable).  pq is the
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
