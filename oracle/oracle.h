/*
 * CPU oracle for the SONG graph search -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU algorithm, used as the checker
 * for the HIP search path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in song_amd/ links it.
 *
 * Every function cites the reference file:line it restates.  Types follow
 * reference config.h:11-14 (value_t = float, dist_t = double, idx_t = size_t).
 *
 * Pinning status (how each function is checked independently of itself):
 *   oracle_distance / oracle_pair_distance   closed forms in tests + golden
 *                                            ids written by the reference
 *                                            binary song.cpu (all 3 metrics)
 *   oracle_search                            golden top-k ids from song.cpu,
 *                                            brute-force on tiny graphs,
 *                                            hand-worked 6-vertex trace
 *   oracle_build                             SHA-256 of song.cpu's bfsg.graph
 *   oracle_parse_libsvm_line                 hand-written lines in tests
 *   oracle_bruteforce_knn                    independent numpy float64 argsort
 */
#ifndef SONG_ORACLE_H
#define SONG_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_L2 = 0, ORACLE_IP = 1, ORACLE_COS = 2 };

typedef struct {
    int64_t expansions;   /* explore_cnt, graph.h:191 */
    int64_t visited;      /* visited.size() at exit = distances computed, graph.h:202-204 */
} oracle_stats;

/* Distance of base row x to query q: data.h:28-62 (generic template, x = get(a)). */
double oracle_distance(int metric, const float *x, const float *q, int dim);

/* Pairwise specialisations data.h:178-215 (a = row owner, b = other row). */
double oracle_pair_distance(int metric, const float *a, const float *b, int dim);

/*
 * FixedDegreeGraph::astar_multi_start_search, graph.h:169-216.
 * graph: n rows of (1<<shift) uint64 entries, slot 0 = degree.
 * n_visible: data->curr_vertices() (n after load, data.h:171; vertex index
 *            during build, data.h:140).  Search starts at vertex 0 if n_visible > 0.
 * Writes min(k, found) ids ascending by (dist, id) into out_ids (and the
 * double distances into out_dists if non-NULL); returns the count.
 */
int oracle_search(const float *data, uint64_t n_visible, int dim, const uint64_t *graph, int shift,
                  int metric, const float *query, int k, uint64_t *out_ids, double *out_dists,
                  oracle_stats *stats);

/*
 * Batch of independent searches on `threads` POSIX threads (static split of
 * queries, as a CPU baseline).  out_ids is nq x k row-major; rows with fewer
 * than k results are padded with UINT64_MAX.  stats (nullable) gets per-query
 * counters.  Returns 0.
 */
int oracle_search_batch(const float *data, uint64_t n, int dim, const uint64_t *graph, int shift,
                        int metric, const float *queries, uint64_t nq, int k, uint64_t *out_ids,
                        double *out_dists, oracle_stats *stats, int threads);

/*
 * Sequential insertion build of FixedDegreeGraph (graph.h:32-167 with
 * config.h:21-27: SEARCH_DEGREE 15, FIXED_DEGREE 31, shift 5, budget 150),
 * driven as main.cc:26-32 (Data::add then add_vertex per row).
 * graph_out: n << 5 uint64, zero-initialised by the caller.  Returns 0.
 */
int oracle_build(const float *data, uint64_t n, int dim, int metric, uint64_t *graph_out);

/*
 * One LIBSVM line, Parser::tokenize + Parser::parse (parser.h:17-43):
 * token positions are every ':' or ' '; pairs (i, i+1) give "%d" index
 * (minus one_based) and "%lf" value cast to float.  Returns the pair count
 * (at most cap pairs are stored).
 */
int oracle_parse_libsvm_line(const char *line, int one_based, int *idx, float *val, int cap);

/* Exact k nearest neighbours by float64 brute force (ids ascending by (d, id)). */
int oracle_bruteforce_knn(const float *data, uint64_t n, int dim, int metric, const float *queries,
                          uint64_t nq, int k, uint64_t *out_ids, double *out_dists, int threads);

#ifdef __cplusplus
}
#endif
#endif
