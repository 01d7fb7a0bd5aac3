/*
 * song_amd host library (CPU side of the drop-in): graph construction, LIBSVM
 * parsing and the raw bfsg.data / bfsg.graph files.  Plain C ABI; returns
 * SONG_OK (0) or a negative SONG_E* code (values as in song_amd.h).
 *
 * Reference interfaces replaced (sunbelbd/song):
 *   song_graph_build      <- the `song build` loop: Parser callback -> Data::add ->
 *                            FixedDegreeGraph::add_vertex (main.cc:26-32, graph.h:149-167,
 *                            kernelgraph.h:150-167).  Same bytes as the reference
 *                            bfsg.graph; runs on several threads.
 *   song_libsvm_parse     <- Parser (parser.h:12-64): one-based "label idx:val ..." lines,
 *                            values read as %lf and cast to float.
 *   song_file_write/read  <- Data::dump/load (data.h:164-175) and the graph dump/load
 *                            (graph.h:259-271): raw little-endian arrays, no header.
 *   song_libsvm_parse_mt  <- the same Parser, with the file cut into chunks at line starts
 *                            and parsed on several threads (same values bit for bit).
 *   song_libsvm_write, song_query_read: no counterpart (the reference reads LIBSVM text
 *                            only, parser.h:57-62); a writer for the reference's input
 *                            format and binary query files (raw float32 rows, .fvecs, .fbin).
 */
#ifndef SONG_HOST_H
#define SONG_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t searches;      /* speculative searches run in parallel */
    uint64_t redone;        /* searches repeated because a row they read changed */
    double seconds;
} song_build_stats;

/*
 * Build the fixed-degree graph of config.h:21-27 (SEARCH_DEGREE 15, FIXED_DEGREE 31,
 * FIXED_DEGREE_SHIFT 5, CONSTRUCT_SEARCH_BUDGET 150) over rows 0..n-1 inserted in
 * order.  graph_out: n * 32 uint64 (written completely).  threads <= 0: all cores.
 * metric: 0 l2, 1 ip, 2 cos.  stats nullable.
 */
int song_graph_build(const float *data, uint64_t n, int dim, int metric, int threads, uint64_t *graph_out,
                     song_build_stats *stats);

/*
 * Parse a LIBSVM text file into a dense row-major float matrix of `rows` x `dim`
 * (rows beyond the file stay zero; lines beyond `rows` are an error, as are
 * indices >= dim).  *rows_read gets the line count (on an index error: the number of
 * lines before the first bad one; the matrix is then unspecified).  Records are read as
 * the reference's fgets loop reads them: a last line without '\n' is a record, a line of
 * 10,000,000 bytes or more is split.  Runs on SONG_THREADS (else min(16, cores)) threads.
 */
int song_libsvm_parse(const char *path, int one_based, uint64_t rows, int dim, float *out, uint64_t *rows_read);
/* song_libsvm_parse on `threads` threads (<= 0: SONG_THREADS, else min(16, cores)). */
int song_libsvm_parse_mt(const char *path, int one_based, uint64_t rows, int dim, int threads, float *out,
                         uint64_t *rows_read);

/*
 * Write rows as one-based LIBSVM text ("0 i:v ...", zeros omitted): integer values as
 * integers, others with 9 significant digits, so song_libsvm_parse reads every value back
 * exactly.  threads as song_libsvm_parse_mt.
 */
int song_libsvm_write(const char *path, const float *data, uint64_t rows, int dim, int threads);

/* Query file formats of song_query_read. */
enum {
    SONG_QUERY_LIBSVM = 0, /* one-based LIBSVM text (the reference's format) */
    SONG_QUERY_F32 = 1,    /* raw float32 rows of dim values, no header (the bfsg.data layout) */
    SONG_QUERY_FVECS = 2,  /* per row: int32 dim, then dim float32 */
    SONG_QUERY_FBIN = 3,   /* uint32 rows, uint32 dim, then rows x dim float32 */
};

/*
 * Read a query file into dense rows (out: max_rows x dim float32).  The size is validated
 * against the format (whole rows, every fvecs row's dimension, the fbin header); more rows
 * than max_rows is an error.  out == NULL: only *rows_read = the file's row count.
 */
int song_query_read(const char *path, int format, int dim, uint64_t max_rows, float *out, uint64_t *rows_read);

/* Raw file helpers: write/read exactly `bytes` bytes; a short file is an error. */
int song_file_write(const char *path, const void *buf, uint64_t bytes);
int song_file_read(const char *path, void *buf, uint64_t bytes);

const char *song_host_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
