// Graph construction for `song build`, byte-identical to the reference insertion build
// (graph.h:32-167 driven by main.cc:26-32) but run on several threads.
//
// The reference inserts vertices one at a time: vertex v searches the graph built so far
// (budget 150, from vertex 0), keeps its 15 nearest as out-edges, ranks them, and links
// itself into each of those neighbours' rows.  Vertex v's search only reads the rows of the
// vertices it expands, so searches for a window of consecutive vertices can run in parallel
// against the graph as it stood at the window start.  Commits then happen in order; a
// speculative result is kept only if none of the rows its search read was changed by an
// earlier commit of the same window, otherwise that search is repeated on the current
// graph.  The committed graph is therefore exactly the sequential one.
//
// Distances use the reference arithmetic (data.h:28-62, 178-215): float products
// accumulated in double in dimension order; the cosine norms are one fmaf chain per row.
// For integer-valued data whose products and sums fit int32 exactly, an int16 copy gives
// the same values with a vectorisable integer loop.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/song_host.h"
#include "host_common.hpp"

namespace {

constexpr int kSearchDegree = 15;    // config.h:26 SEARCH_DEGREE
constexpr int kFixedDegree = 31;     // config.h:22 FIXED_DEGREE
constexpr int kShift = 5;            // config.h:23 FIXED_DEGREE_SHIFT
constexpr int kBudget = 150;         // config.h:27 CONSTRUCT_SEARCH_BUDGET
constexpr int kMaxStep = 1000000;    // graph.h:183

struct Metric {
    const float *data = nullptr;
    int dim = 0, metric = 0;
    bool int_mode = false;
    std::vector<int16_t> i16;
    std::vector<float> norms;

    void init(const float *d, uint64_t n, int dm, int mt) {
        data = d;
        dim = dm;
        metric = mt;
        float amax = 0.f;
        bool integer = true;
        for (uint64_t i = 0; i < n * (uint64_t)dim && integer; ++i) {
            float x = d[i];
            if (x != std::nearbyint(x) || std::fabs(x) > 2047.f) integer = false;
            amax = std::max(amax, std::fabs(x));
        }
        // every product and partial sum an exact int32 (and an exact float product)
        double bound = metric == 0 ? (double)dim * (2.0 * amax) * (2.0 * amax) : (double)dim * amax * amax;
        int_mode = integer && metric != 2 && bound < 2147483647.0;
        if (int_mode) {
            i16.resize(n * (uint64_t)dim);
            for (uint64_t i = 0; i < i16.size(); ++i) i16[i] = (int16_t)d[i];
        }
        if (metric == 2) {
            norms.resize(n);
            for (uint64_t v = 0; v < n; ++v) {
                const float *r = d + v * (uint64_t)dim;
                float s = 0.f;
                for (int i = 0; i < dim; ++i) s = std::fma(r[i], r[i], s);
                norms[v] = s;
            }
        }
    }

    void prefetch(uint32_t v) const {
        const char *p = int_mode ? (const char *)(i16.data() + (uint64_t)v * dim)
                                 : (const char *)(data + (uint64_t)v * dim);
        size_t bytes = (size_t)dim * (int_mode ? 2 : 4);
        for (size_t o = 0; o < bytes; o += 64) __builtin_prefetch(p + o);
    }

    // distance with a = the measured vertex (row owner / candidate), b = the other row
    double pair(uint32_t a, uint32_t b) const {
        if (int_mode) {
            const int16_t *x = i16.data() + (uint64_t)a * dim, *y = i16.data() + (uint64_t)b * dim;
            int32_t s = 0;
            if (metric == 0) {
                for (int i = 0; i < dim; ++i) {
                    int32_t t = (int32_t)x[i] - (int32_t)y[i];
                    s += t * t;
                }
                return (double)s;
            }
            for (int i = 0; i < dim; ++i) s += (int32_t)x[i] * (int32_t)y[i];
            return -(double)s;
        }
        const float *x = data + (uint64_t)a * dim, *y = data + (uint64_t)b * dim;
        double ret = 0.0;
        if (metric == 0) {
            for (int i = 0; i < dim; ++i) {
                float diff = x[i] - y[i];
                float p = diff * diff;
                ret += (double)p;
            }
            return ret;
        }
        if (metric == 1) {
            for (int i = 0; i < dim; ++i) {
                float p = x[i] * y[i];
                ret -= (double)p;
            }
            return ret;
        }
        for (int i = 0; i < dim; ++i) {
            float p = x[i] * y[i];
            ret += (double)p;
        }
        int sign = ret < 0 ? 1 : -1;
        return (double)sign * (ret * ret / (double)norms[a] / (double)norms[b]);
    }
};

using DP = std::pair<double, uint32_t>;

struct SearchCtx {
    std::vector<uint32_t> stamp;  // visited marks: stamp[u] == epoch
    uint32_t epoch = 0;
    std::vector<DP> q, topk;
    std::vector<uint32_t> fresh;

    void init(uint64_t n) { stamp.assign(n, 0u); }
    uint32_t next_epoch() {
        if (++epoch == 0u) {
            std::fill(stamp.begin(), stamp.end(), 0u);
            epoch = 1u;
        }
        return epoch;
    }
};

// graph.h:169-216 with query = row `qv`, visible vertices 0..qv-1 (data.h:140).
// Writes up to k ids ascending by (d, id) into out; the ids of expanded vertices
// (the rows this search read) go to `reads`.
int search_build(const Metric &M, const uint64_t *edges, uint32_t qv, int k, uint32_t *out,
                 std::vector<uint32_t> *reads, SearchCtx &C) {
    const uint32_t ep = C.next_epoch();
    C.q.clear();
    C.topk.clear();
    if (reads) reads->clear();
    if (qv == 0) return 0;
    C.stamp[0] = ep;
    C.q.push_back(DP(M.pair(0, qv), 0u));
    auto gt = std::greater<DP>();
    for (int iter = 0; iter < kMaxStep && !C.q.empty(); ++iter) {
        DP now = C.q.front();
        if ((int)C.topk.size() == k && C.topk.front().first < now.first) break;
        std::pop_heap(C.q.begin(), C.q.end(), gt);
        C.q.pop_back();
        C.topk.push_back(now);
        std::push_heap(C.topk.begin(), C.topk.end());
        if ((int)C.topk.size() > k) {
            std::pop_heap(C.topk.begin(), C.topk.end());
            C.topk.pop_back();
        }
        if (reads) reads->push_back(now.second);
        const uint64_t *row = edges + ((uint64_t)now.second << kShift);
        const uint64_t deg = row[0];
        C.fresh.clear();
        for (uint64_t i = 0; i < deg; ++i) {
            uint32_t nb = (uint32_t)row[i + 1];
            if (C.stamp[nb] == ep) continue;
            C.stamp[nb] = ep;
            C.fresh.push_back(nb);
            M.prefetch(nb);
        }
        for (uint32_t nb : C.fresh) {
            C.q.push_back(DP(M.pair(nb, qv), nb));
            std::push_heap(C.q.begin(), C.q.end(), gt);
        }
    }
    int cnt = (int)C.topk.size();
    std::sort_heap(C.topk.begin(), C.topk.end());
    for (int i = 0; i < cnt; ++i) out[i] = C.topk[i].second;
    return cnt;
}

struct Graph {
    uint64_t *edges;
    std::vector<double> edge_dist;
    const Metric *M;

    // graph.h:98-114: Hoare partition on edge_dist with int cursors, swapping edges along
    void qsort(size_t l, size_t r) {
        size_t mid = (l + r) >> 1;
        int i = (int)l, j = (int)r;
        double k = edge_dist[mid];
        do {
            while (edge_dist[i] < k) ++i;
            while (k < edge_dist[j]) --j;
            if (i <= j) {
                std::swap(edge_dist[i], edge_dist[j]);
                std::swap(edges[i], edges[j]);
                ++i;
                --j;
            }
        } while (i <= j);
        if ((size_t)i < r) qsort((size_t)i, r);
        if ((long long)l < (long long)j) qsort(l, (size_t)j);
    }
    // graph.h:75-85, 116-124
    void rank_edges(size_t off) {
        uint64_t deg = edges[off];
        uint32_t owner = (uint32_t)(off >> kShift);
        for (uint64_t i = 0; i < deg; ++i) edge_dist[off + i + 1] = M->pair(owner, (uint32_t)edges[off + i + 1]);
        qsort(off + 1, off + deg);
    }
    // graph.h:32-55; returns true if the row changed
    bool rank_and_switch(uint32_t v, uint32_t u) {
        double cur = M->pair(v, u);
        size_t off = (size_t)v << kShift;
        if (cur >= edge_dist[off + edges[off]]) return false;
        edges[off + edges[off]] = u;
        edge_dist[off + edges[off]] = cur;
        for (size_t i = off + edges[off] - 1; i > off; --i) {
            if (edge_dist[i] > edge_dist[i + 1]) {
                std::swap(edge_dist[i], edge_dist[i + 1]);
                std::swap(edges[i], edges[i + 1]);
            } else {
                break;
            }
        }
        return true;
    }
    // graph.h:126-137
    bool add_edge(uint32_t v, uint32_t u) {
        size_t off = (size_t)v << kShift;
        if (edges[off] < (uint64_t)kFixedDegree) {
            ++edges[off];
            edges[off + edges[off]] = u;
            if (edges[off] == (uint64_t)kFixedDegree) rank_edges(off);
            return true;
        }
        return rank_and_switch(v, u);
    }
    // graph.h:149-167 after the search; marks changed rows with stamp v
    void commit(uint32_t v, const uint32_t *nb, int got, std::vector<int64_t> &mod) {
        int num = std::min(got, kSearchDegree);
        size_t off = (size_t)v << kShift;
        edges[off] = (uint64_t)num;
        for (int i = 0; i < num; ++i) edges[off + i + 1] = nb[i];
        rank_edges(off);
        mod[v] = v;
        for (int i = 0; i < num; ++i)
            if (add_edge(nb[i], v)) mod[nb[i]] = v;
    }
};

// fixed set of workers running a parallel-for per window
class Pool {
  public:
    explicit Pool(int n) : n_(n) {
        for (int t = 1; t < n_; ++t) th_.emplace_back([this, t] { loop(t); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    void run(uint64_t count, const std::function<void(int, uint64_t)> &fn) {
        {
            std::lock_guard<std::mutex> g(mu_);
            fn_ = &fn;
            count_ = count;
            next_.store(0);
            busy_ = n_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        work(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return busy_ == 0; });
    }

  private:
    void work(int t) {
        for (;;) {
            uint64_t i = next_.fetch_add(1);
            if (i >= count_) break;
            (*fn_)(t, i);
        }
    }
    void loop(int t) {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
            }
            work(t);
            std::lock_guard<std::mutex> g(mu_);
            if (--busy_ == 0) done_cv_.notify_one();
        }
    }
    int n_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int, uint64_t)> *fn_ = nullptr;
    uint64_t count_ = 0, gen_ = 0;
    std::atomic<uint64_t> next_{0};
    int busy_ = 0;
    bool stop_ = false;
};

}  // namespace

extern "C" int song_graph_build(const float *data, uint64_t n, int dim, int metric, int threads, uint64_t *graph_out,
                                song_build_stats *stats) {
    if (!data || !graph_out) return host_fail(-1, "data/graph_out is null");
    if (n == 0 || n >= 0xFFFFFFFFull || dim < 1) return host_fail(-1, "bad sizes n=%llu dim=%d", (unsigned long long)n, dim);
    if (metric < 0 || metric > 2) return host_fail(-1, "metric=%d", metric);
    auto t0 = std::chrono::steady_clock::now();
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, 64);

    Metric M;
    M.init(data, n, dim, metric);
    std::memset(graph_out, 0, sizeof(uint64_t) * (n << kShift));
    Graph G{graph_out, std::vector<double>(n << kShift, 0.0), &M};

    std::vector<SearchCtx> ctx(threads);
    for (auto &c : ctx) c.init(n);
    std::vector<int64_t> mod(n, -1);
    // Window size adapts to the conflict rate: a search is redone (sequentially) when an
    // earlier commit of its window changed a row it read, so a window holds between
    // `threads` and 8 x `threads` searches: halved when over a quarter of the last window
    // was redone, doubled when under a twentieth was.
    const uint64_t Wmin = threads == 1 ? 1 : (uint64_t)threads, Wmax = threads == 1 ? 1 : (uint64_t)threads * 8;
    uint64_t W = threads == 1 ? 1 : (uint64_t)threads * 2;
    std::vector<std::vector<uint32_t>> reads(Wmax);
    std::vector<uint32_t> res(Wmax * kBudget);
    std::vector<int> got(Wmax);
    uint64_t searches = 0, redone = 0;
    Pool pool(threads);

    // SONG_BUILD_PROGRESS set: a line on stderr about every 20 s of building
    const bool progress = std::getenv("SONG_BUILD_PROGRESS") != nullptr;
    auto last_report = t0;
    for (uint64_t s = 0, e = 0; s < n; s = e) {
        e = std::min(n, s + W);
        const uint64_t redone0 = redone;
        if (progress) {
            const auto now = std::chrono::steady_clock::now();
            if (std::chrono::duration<double>(now - last_report).count() >= 20.0) {
                last_report = now;
                std::fprintf(stderr, "[build] %llu / %llu vertices, %.0f s, %llu redone, window %llu\n",
                             (unsigned long long)s, (unsigned long long)n,
                             std::chrono::duration<double>(now - t0).count(), (unsigned long long)redone,
                             (unsigned long long)W);
                std::fflush(stderr);
            }
        }
        if (threads > 1 && e - s > 1) {
            pool.run(e - s, [&](int t, uint64_t i) {
                uint32_t v = (uint32_t)(s + i);
                got[i] = search_build(M, graph_out, v, kBudget, &res[i * kBudget], &reads[i], ctx[t]);
            });
            searches += e - s;
        }
        for (uint64_t v = s; v < e; ++v) {
            const uint64_t i = v - s;
            bool valid = threads > 1 && e - s > 1;
            if (valid)
                for (uint32_t u : reads[i])
                    if (mod[u] >= (int64_t)s) {
                        valid = false;
                        break;
                    }
            if (!valid) {
                got[i] = search_build(M, graph_out, (uint32_t)v, kBudget, &res[i * kBudget], nullptr, ctx[0]);
                if (threads > 1 && e - s > 1) ++redone;
            }
            G.commit((uint32_t)v, &res[i * kBudget], got[i], mod);
        }
        if (threads > 1) {
            const uint64_t r = redone - redone0, w = e - s;
            if (r * 4 > w) W = std::max(Wmin, W / 2);
            else if (r * 20 < w) W = std::min(Wmax, W * 2);
        }
    }
    if (stats) {
        stats->searches = searches;
        stats->redone = redone;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return 0;
}
