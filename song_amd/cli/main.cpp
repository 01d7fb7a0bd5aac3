// `song` command line, drop-in for the reference binaries (main.cu:53-100, main.cc:49-105):
//   song <build/test> <build_data> <query_data> <search_top_k> <row> <dim> <return_top_k> <l2/ip/cos>
// build: parse the one-based LIBSVM rows, build the graph (byte-identical to the reference
//        insertion build), write bfsg.data and bfsg.graph to the working directory.
// test:  load bfsg.data / bfsg.graph, search every query of the LIBSVM file (or a binary
//        .f32 / .fvecs / .fbin query file) in one GPU batch, print the first <return_top_k>
//        ids of each result, "%zu " each, one line per query (main.cu:28-33).
// The search budget pq is SONG_PQ when fill_parameters.sh compiled one in (the reference
// GPU binary's TOPK), otherwise <search_top_k> (the reference CPU binary's k).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "kernelgraph.hpp"
#if __has_include("song_params.h")
#include "song_params.h"
#endif

using namespace song_cli;

static void usage(char **argv) {
    printf("Usage: %s <build/test> <build_data> <query_data> <search_top_k> <row> <dim> <return_top_k> <l2/ip/cos>\n",
           argv[0]);
}

static bool ends_with(const std::string &s, const char *suf) {
    const size_t n = strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

static int query_format(const std::string &path) {
    if (ends_with(path, ".f32")) return SONG_QUERY_F32;
    if (ends_with(path, ".fvecs")) return SONG_QUERY_FVECS;
    if (ends_with(path, ".fbin")) return SONG_QUERY_FBIN;
    return SONG_QUERY_LIBSVM;
}

template <int M>
static int run(const std::string &mode, char **argv, size_t row, int dim, int pq, int display) {
    KernelFixedDegreeGraph<M> g(row, dim, pq);
    if (mode == "build") {
        uint64_t got = 0;
        check_host(song_libsvm_parse(argv[2], 1, row, dim, g.data(), &got), "parse");
        g.set_rows(got);
        fprintf(stderr, "Writing the graph and data...");
        g.dump_data();
        fprintf(stderr, "...");
        g.dump();
        fprintf(stderr, "done\n");
        return 0;
    }
    fprintf(stderr, "Loading the graph and data...");
    g.load_data();
    fprintf(stderr, "...");
    g.load();
    fprintf(stderr, "done\n");
    // the query file: LIBSVM text (the reference's format), or binary rows by extension
    // (.f32 raw float32 rows, .fvecs, .fbin); count the records, read them into dense rows,
    // search as one batch
    const int fmt = query_format(argv[3]);
    uint64_t lines = 0;
    check_host(song_query_read(argv[3], fmt, dim, 0, nullptr, &lines), "count queries");
    std::vector<float> dense(std::max<uint64_t>(lines, 1) * (size_t)dim);
    uint64_t nq = 0;
    check_host(song_query_read(argv[3], fmt, dim, std::max<uint64_t>(lines, 1), dense.data(), &nq), "read queries");
    std::vector<std::vector<std::pair<int, value_t>>> queries(nq);
    for (uint64_t i = 0; i < nq; ++i)
        for (int j = 0; j < dim; ++j)
            if (dense[i * dim + j] != 0.0f) queries[i].push_back({j, dense[i * dim + j]});
    std::vector<std::vector<idx_t>> results;
    g.search_top_k_batch(queries, pq, results);
    for (const auto &r : results) {
        for (size_t j = 0; j < r.size() && (int)j < display; ++j) printf("%zu ", r[j]);
        printf("\n");
    }
    return 0;
}

int main(int argc, char **argv) {
    if (argc != 9) {
        usage(argv);
        return 1;
    }
    const size_t row = atoll(argv[5]);
    const int dim = atoi(argv[6]);
    const int display = atoi(argv[7]);
    const std::string metric = argv[8], mode = argv[1];
    int pq = atoi(argv[4]);
#ifdef SONG_PQ
    pq = SONG_PQ;
#endif
#ifdef SONG_DIM
    if (dim != SONG_DIM) fprintf(stderr, "warning: dim %d differs from the compiled %d\n", dim, SONG_DIM);
#endif
    if ((mode != "build" && mode != "test") || row == 0 || dim <= 0) {
        usage(argv);
        return 1;
    }
    if (mode == "test" && pq <= 0) {
        fprintf(stderr, "pq (search_top_k) must be positive\n");
        return 1;
    }
    try {
        if (metric == "l2") return run<0>(mode, argv, row, dim, pq, display);
        if (metric == "ip") return run<1>(mode, argv, row, dim, pq, display);
        if (metric == "cos") return run<2>(mode, argv, row, dim, pq, display);
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    usage(argv);
    return 1;
}
