// Drop-in C++ host interface of the graph index, over the song_amd C ABIs.
//
// Mirrors the reference's GraphWrapper virtual interface (kernelgraph.h:14-22,
// graph.h:11-19) so callers of search_top_k / search_top_k_batch keep their code:
//   add_vertex(id, sparse point)       rows are buffered; the graph is built (in parallel,
//                                      byte-identical to the reference insertion build) by
//                                      song_graph_build the first time it is needed
//   search_top_k(query, k, result)     one query on the GPU
//   search_top_k_batch(queries, k, r)  the batched GPU search (kernelgraph.h:269-272)
//   dump(file) / load(file)            raw bfsg.graph, n * 32 size_t (graph.h:259-271)
// The search budget is pq (the reference's compiled TOPK); k is how many ids each
// result keeps (the reference GPU path ignores k and returns TOPK ids).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/song_amd.h"
#include "../../include/song_host.h"

namespace song_cli {

typedef float value_t;  // reference config.h:11
typedef size_t idx_t;   // reference config.h:14

class GraphWrapper {
  public:
    virtual void add_vertex(idx_t vertex_id, std::vector<std::pair<int, value_t>> &point) = 0;
    virtual void search_top_k(const std::vector<std::pair<int, value_t>> &query, int k, std::vector<idx_t> &result) = 0;
    virtual void dump(std::string file = "bfsg.graph") = 0;
    virtual void load(std::string file = "bfsg.graph") = 0;
    virtual void search_top_k_batch(const std::vector<std::vector<std::pair<int, value_t>>> &queries, int k,
                                    std::vector<std::vector<idx_t>> &results) = 0;
    virtual ~GraphWrapper() {}
};

inline void check_host(int rc, const char *what) {
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + song_host_last_error());
}
inline void check_gpu(int rc, const char *what) {
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + song_last_error());
}

// dist_type: 0 l2, 1 ip, 2 cos (the reference's template argument, kernelgraph.h:24)
template <int dist_type>
class KernelFixedDegreeGraph : public GraphWrapper {
  public:
    KernelFixedDegreeGraph(size_t num, int dim, int pq) : num_(num), dim_(dim), pq_(pq) {
        data_.assign(num * (size_t)dim, 0.0f);
        edges_.assign(num << 5, 0);
    }
    ~KernelFixedDegreeGraph() override {
        if (index_) song_index_destroy(index_);
    }

    float *data() { return data_.data(); }
    size_t num() const { return num_; }
    int dim() const { return dim_; }
    // rows added so far are visible (the reference's Data::add keeps curr_num = last id, data.h:140)
    void set_rows(size_t rows) { rows_ = rows; }

    void add_vertex(idx_t vertex_id, std::vector<std::pair<int, value_t>> &point) override {
        if (vertex_id >= num_) throw std::runtime_error("vertex id beyond the row count");
        float *r = data_.data() + vertex_id * (size_t)dim_;
        for (const auto &p : point)
            if (p.first >= 0 && p.first < dim_) r[p.first] = p.second;
        rows_ = std::max(rows_, (size_t)vertex_id + 1);
        built_ = false;
    }

    void build(int threads = 0) {
        if (built_) return;
        std::vector<uint64_t> g(rows_ << 5, 0);
        check_host(song_graph_build(data_.data(), rows_, dim_, dist_type, threads, g.data(), nullptr), "graph build");
        std::fill(edges_.begin(), edges_.end(), 0);
        std::copy(g.begin(), g.end(), edges_.begin());
        built_ = true;
        drop_index();
    }

    void search_top_k(const std::vector<std::pair<int, value_t>> &query, int k, std::vector<idx_t> &result) override {
        std::vector<std::vector<std::pair<int, value_t>>> one(1, query);
        std::vector<std::vector<idx_t>> r;
        search_top_k_batch(one, k, r);
        result = r.empty() ? std::vector<idx_t>() : r[0];
    }

    void search_top_k_batch(const std::vector<std::vector<std::pair<int, value_t>>> &queries, int k,
                            std::vector<std::vector<idx_t>> &results) override {
        build();
        upload();
        const int keep = (k > 0 && k < pq_) ? k : pq_;
        std::vector<float> dense(queries.size() * (size_t)dim_, 0.0f);  // organize_point, data.h:106-114
        for (size_t i = 0; i < queries.size(); ++i)
            for (const auto &p : queries[i])
                if (p.first >= 0 && p.first < dim_) dense[i * dim_ + p.first] = p.second;
        std::vector<uint64_t> ids(queries.size() * (size_t)keep);
        song_search_opts o;
        song_search_opts_default(&o);
        check_gpu(song_search(index_, dense.data(), queries.size(), pq_, keep, ids.data(), nullptr, nullptr, &o),
                  "search");
        results.assign(queries.size(), std::vector<idx_t>());
        for (size_t i = 0; i < queries.size(); ++i)
            for (int j = 0; j < keep; ++j) {
                const uint64_t v = ids[i * keep + j];
                if (v == UINT64_MAX) break;
                results[i].push_back((idx_t)v);
            }
    }

    void dump(std::string file = "bfsg.graph") override {
        build();
        check_host(song_file_write(file.c_str(), edges_.data(), edges_.size() * sizeof(uint64_t)), "graph dump");
    }
    void load(std::string file = "bfsg.graph") override {
        check_host(song_file_read(file.c_str(), edges_.data(), edges_.size() * sizeof(uint64_t)), "graph load");
        rows_ = num_;
        built_ = true;
        drop_index();
    }
    void dump_data(std::string file = "bfsg.data") {
        check_host(song_file_write(file.c_str(), data_.data(), data_.size() * sizeof(float)), "data dump");
    }
    void load_data(std::string file = "bfsg.data") {
        check_host(song_file_read(file.c_str(), data_.data(), data_.size() * sizeof(float)), "data load");
        rows_ = num_;
        drop_index();
    }

  private:
    void upload() {
        if (index_) return;
        check_gpu(song_index_create(data_.data(), num_, dim_, edges_.data(), 5, dist_type, &index_), "index upload");
    }
    void drop_index() {
        if (index_) song_index_destroy(index_);
        index_ = nullptr;
    }

    size_t num_, rows_ = 0;
    int dim_, pq_;
    bool built_ = false;
    std::vector<float> data_;
    std::vector<uint64_t> edges_;
    song_index *index_ = nullptr;
};

}  // namespace song_cli
