/*
 * CPU oracle for the SONG graph search -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this file's library.  See oracle.h for the pinning status of each
 * function.  Built with -ffp-contract=off so that every rounding step is the
 * one written here: float products are rounded to float, then accumulated in
 * double, exactly as the reference's data.h does (a float product converted to
 * double cannot be contracted into an FMA).  The cosine norms use an explicit
 * fmaf() because the reference build (Makefile:8, -O3 -march=native) contracts
 * `lena += a*a` into a fused multiply-add.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- distances */

/* data.h:28-37 l2_distance: diff = float(pa[i] - v[i]); ret(double) += float(diff*diff) */
static double l2_dist(const float *x, const float *q, int dim) {
    double ret = 0.0;
    for (int i = 0; i < dim; ++i) {
        float diff = x[i] - q[i];
        float p = diff * diff;
        ret += (double)p;
    }
    return ret;
}

/* data.h:39-47 negative_inner_prod_distance: ret(double) -= float(pa[i]*v[i]) */
static double ip_dist(const float *x, const float *q, int dim) {
    double ret = 0.0;
    for (int i = 0; i < dim; ++i) {
        float p = x[i] * q[i];
        ret -= (double)p;
    }
    return ret;
}

/*
 * data.h:49-62 negative_cosine_distance:
 *   ret(double) += float(pa[i]*v[i]); lena = fmaf(pa[i],pa[i],lena); lenv = fmaf(v[i],v[i],lenv)
 *   sign = ret < 0 ? 1 : -1; return sign * (ret*ret/lena/lenv)   (double arithmetic)
 */
static double cos_dist(const float *x, const float *q, int dim) {
    double ret = 0.0;
    float lena = 0.0f, lenv = 0.0f;
    for (int i = 0; i < dim; ++i) {
        float p = x[i] * q[i];
        ret += (double)p;
        lena = fmaf(x[i], x[i], lena);
        lenv = fmaf(q[i], q[i], lenv);
    }
    int sign = ret < 0 ? 1 : -1;
    return (double)sign * (ret * ret / (double)lena / (double)lenv);
}

double oracle_distance(int metric, const float *x, const float *q, int dim) {
    switch (metric) {
    case ORACLE_L2: return l2_dist(x, q, dim);
    case ORACLE_IP: return ip_dist(x, q, dim);
    default: return cos_dist(x, q, dim);
    }
}

/* data.h:178-215: same arithmetic with a = row owner, b = other row. */
double oracle_pair_distance(int metric, const float *a, const float *b, int dim) {
    return oracle_distance(metric, a, b, dim);
}

/* ------------------------------------------------------- heaps of (d, id) */

typedef struct {
    double d;
    uint64_t id;
} pair_t;

/* std::pair<double,size_t> operator< : lexicographic. */
static inline int pair_lt(pair_t a, pair_t b) {
    return a.d < b.d || (!(b.d < a.d) && a.id < b.id);
}

typedef struct {
    pair_t *v;
    size_t n, cap;
    int is_max; /* 1: max-heap (std::priority_queue default), 0: min-heap (std::greater) */
} heap_t;

static inline int heap_before(const heap_t *h, pair_t a, pair_t b) {
    return h->is_max ? pair_lt(b, a) : pair_lt(a, b);
}

static void heap_push(heap_t *h, pair_t x) {
    if (h->n == h->cap) {
        h->cap = h->cap ? h->cap * 2 : 64;
        h->v = (pair_t *)realloc(h->v, h->cap * sizeof(pair_t));
    }
    size_t i = h->n++;
    while (i > 0) {
        size_t p = (i - 1) / 2;
        if (!heap_before(h, x, h->v[p])) break;
        h->v[i] = h->v[p];
        i = p;
    }
    h->v[i] = x;
}

static void heap_pop(heap_t *h) {
    pair_t x = h->v[--h->n];
    size_t i = 0, n = h->n;
    for (;;) {
        size_t l = 2 * i + 1, r = l + 1, b = i;
        pair_t bv = x;
        if (l < n && heap_before(h, h->v[l], bv)) { b = l; bv = h->v[l]; }
        if (r < n && heap_before(h, h->v[r], bv)) { b = r; bv = h->v[r]; }
        if (b == i) break;
        h->v[i] = h->v[b];
        i = b;
    }
    if (n) h->v[i] = x;
}

/* ------------------------------------------------------- visited id set */

typedef struct {
    uint64_t *slot; /* id + 1, 0 = empty */
    size_t cap, n;
} idset_t;

static inline size_t id_hash(uint64_t x, size_t mask) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    return (size_t)x & mask;
}

static void idset_init(idset_t *s, size_t cap) {
    s->cap = cap;
    s->n = 0;
    s->slot = (uint64_t *)calloc(cap, sizeof(uint64_t));
}

static int idset_has(const idset_t *s, uint64_t id) {
    size_t m = s->cap - 1, h = id_hash(id, m);
    while (s->slot[h]) {
        if (s->slot[h] == id + 1) return 1;
        h = (h + 1) & m;
    }
    return 0;
}

static void idset_add(idset_t *s, uint64_t id);

static void idset_grow(idset_t *s) {
    idset_t t;
    idset_init(&t, s->cap * 2);
    for (size_t i = 0; i < s->cap; ++i)
        if (s->slot[i]) idset_add(&t, s->slot[i] - 1);
    free(s->slot);
    *s = t;
}

static void idset_add(idset_t *s, uint64_t id) {
    if (2 * (s->n + 1) > s->cap) idset_grow(s);
    size_t m = s->cap - 1, h = id_hash(id, m);
    while (s->slot[h]) {
        if (s->slot[h] == id + 1) return;
        h = (h + 1) & m;
    }
    s->slot[h] = id + 1;
    s->n++;
}

/* ----------------------------------------------------------------- search */

/* graph.h:169-216 FixedDegreeGraph::astar_multi_start_search */
int oracle_search(const float *data, uint64_t n_visible, int dim, const uint64_t *graph, int shift,
                  int metric, const float *query, int k, uint64_t *out_ids, double *out_dists,
                  oracle_stats *stats) {
    heap_t q = {0}, topk = {0};
    q.is_max = 0;
    topk.is_max = 1;
    idset_t visited;
    idset_init(&visited, 1024);

    /* :175-181 -- one start point, always vertex 0, if any vertex is visible */
    if (n_visible > 0) {
        idset_add(&visited, 0);
        pair_t s = {oracle_distance(metric, data, query, dim), 0};
        heap_push(&q, s);
    }
    const int max_step = 1000000; /* :183 */
    int64_t explore_cnt = 0;
    for (int iter = 0; iter < max_step && q.n > 0; ++iter) {
        pair_t now = q.v[0];
        /* :188 stop: |topk| == k and topk.top().first < now.first (strict, distance only) */
        if ((int64_t)topk.n == k && topk.v[0].d < now.d) break;
        ++explore_cnt;
        heap_pop(&q);
        heap_push(&topk, now);
        if ((int64_t)topk.n > k) heap_pop(&topk);
        const uint64_t *row = graph + (now.id << shift);
        uint64_t degree = row[0];
        for (uint64_t i = 0; i < degree; ++i) { /* :198-205, row order */
            uint64_t nb = row[i + 1];
            if (idset_has(&visited, nb)) continue;
            pair_t c = {oracle_distance(metric, data + nb * (uint64_t)dim, query, dim), nb};
            heap_push(&q, c);
            idset_add(&visited, nb);
        }
    }
    /* :209-215 result ascending by (d, id) */
    int cnt = (int)topk.n;
    int outn = cnt < k ? cnt : k;
    pair_t *tmp = (pair_t *)malloc((cnt ? cnt : 1) * sizeof(pair_t));
    for (int i = cnt - 1; i >= 0; --i) {
        tmp[i] = topk.v[0];
        heap_pop(&topk);
    }
    for (int i = 0; i < outn; ++i) {
        out_ids[i] = tmp[i].id;
        if (out_dists) out_dists[i] = tmp[i].d;
    }
    if (stats) {
        stats->expansions = explore_cnt;
        stats->visited = (int64_t)visited.n;
    }
    free(tmp);
    free(q.v);
    free(topk.v);
    free(visited.slot);
    return outn;
}

typedef struct {
    const float *data;
    uint64_t n;
    int dim;
    const uint64_t *graph;
    int shift, metric;
    const float *queries;
    uint64_t q0, q1;
    int k;
    uint64_t *out_ids;
    double *out_dists;
    oracle_stats *stats;
} batch_job_t;

static void *batch_worker(void *arg) {
    batch_job_t *j = (batch_job_t *)arg;
    for (uint64_t qi = j->q0; qi < j->q1; ++qi) {
        uint64_t *ids = j->out_ids + qi * (uint64_t)j->k;
        double *ds = j->out_dists ? j->out_dists + qi * (uint64_t)j->k : NULL;
        int got = oracle_search(j->data, j->n, j->dim, j->graph, j->shift, j->metric,
                                j->queries + qi * (uint64_t)j->dim, j->k, ids, ds,
                                j->stats ? j->stats + qi : NULL);
        for (int i = got; i < j->k; ++i) {
            ids[i] = UINT64_MAX;
            if (ds) ds[i] = INFINITY;
        }
    }
    return NULL;
}

int oracle_search_batch(const float *data, uint64_t n, int dim, const uint64_t *graph, int shift,
                        int metric, const float *queries, uint64_t nq, int k, uint64_t *out_ids,
                        double *out_dists, oracle_stats *stats, int threads) {
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > nq) threads = nq ? (int)nq : 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    batch_job_t *jobs = (batch_job_t *)malloc(sizeof(batch_job_t) * threads);
    for (int t = 0; t < threads; ++t) {
        batch_job_t j = {data, n, dim, graph, shift, metric, queries,
                         nq * t / threads, nq * (t + 1) / threads, k, out_ids, out_dists, stats};
        jobs[t] = j;
        if (threads == 1) batch_worker(&jobs[t]);
        else pthread_create(&th[t], NULL, batch_worker, &jobs[t]);
    }
    if (threads > 1)
        for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
    return 0;
}

/* ------------------------------------------------------------------ build */

#define SEARCH_DEGREE 15            /* config.h:26 */
#define FIXED_DEGREE 31             /* config.h:22 */
#define FIXED_DEGREE_SHIFT 5        /* config.h:23 */
#define CONSTRUCT_SEARCH_BUDGET 150 /* config.h:27 */

typedef struct {
    const float *data;
    int dim, metric;
    uint64_t *edges;
    double *edge_dist;
} builder_t;

static double bpair(const builder_t *b, uint64_t a, uint64_t c) {
    return oracle_pair_distance(b->metric, b->data + a * (uint64_t)b->dim, b->data + c * (uint64_t)b->dim,
                                b->dim);
}

/* graph.h:98-114 -- Hoare-style quicksort on edge_dist with `int` cursors, swapping
 * edges alongside; the tie order of this exact routine shapes the graph bytes. */
static void bqsort(builder_t *b, size_t l, size_t r) {
    size_t mid = (l + r) >> 1;
    int i = (int)l, j = (int)r;
    double k = b->edge_dist[mid];
    do {
        while (b->edge_dist[i] < k) ++i;
        while (k < b->edge_dist[j]) --j;
        if (i <= j) {
            double td = b->edge_dist[i];
            b->edge_dist[i] = b->edge_dist[j];
            b->edge_dist[j] = td;
            uint64_t te = b->edges[i];
            b->edges[i] = b->edges[j];
            b->edges[j] = te;
            ++i;
            --j;
        }
    } while (i <= j);
    if ((size_t)i < r) bqsort(b, (size_t)i, r);
    if ((long long)l < (long long)j) bqsort(b, l, (size_t)j);
}

/* graph.h:75-85, 116-124 rank_edges */
static void rank_edges(builder_t *b, size_t offset) {
    uint64_t deg = b->edges[offset];
    uint64_t owner = offset >> FIXED_DEGREE_SHIFT;
    for (uint64_t i = 0; i < deg; ++i)
        b->edge_dist[offset + i + 1] = bpair(b, owner, b->edges[offset + i + 1]);
    bqsort(b, offset + 1, offset + deg);
}

/* graph.h:32-55 rank_and_switch_ordered */
static void rank_and_switch(builder_t *b, uint64_t v, uint64_t u) {
    double cur = bpair(b, v, u);
    size_t offset = (size_t)v << FIXED_DEGREE_SHIFT;
    if (cur >= b->edge_dist[offset + b->edges[offset]]) return;
    b->edges[offset + b->edges[offset]] = u;
    b->edge_dist[offset + b->edges[offset]] = cur;
    for (size_t i = offset + b->edges[offset] - 1; i > offset; --i) {
        if (b->edge_dist[i] > b->edge_dist[i + 1]) {
            double td = b->edge_dist[i];
            b->edge_dist[i] = b->edge_dist[i + 1];
            b->edge_dist[i + 1] = td;
            uint64_t te = b->edges[i];
            b->edges[i] = b->edges[i + 1];
            b->edges[i + 1] = te;
        } else {
            break;
        }
    }
}

/* graph.h:126-137 add_edge */
static void add_edge(builder_t *b, uint64_t v, uint64_t u) {
    size_t offset = (size_t)v << FIXED_DEGREE_SHIFT;
    if (b->edges[offset] < FIXED_DEGREE) {
        ++b->edges[offset];
        b->edges[offset + b->edges[offset]] = u;
        if (b->edges[offset] == FIXED_DEGREE) rank_edges(b, offset);
    } else {
        rank_and_switch(b, v, u);
    }
}

int oracle_build(const float *data, uint64_t n, int dim, int metric, uint64_t *graph_out) {
    builder_t b = {data, dim, metric, graph_out, NULL};
    b.edge_dist = (double *)calloc((size_t)n << FIXED_DEGREE_SHIFT, sizeof(double));
    uint64_t nb[CONSTRUCT_SEARCH_BUDGET];
    for (uint64_t v = 0; v < n; ++v) {
        /* main.cc:30 Data::add sets curr_num = max(curr_num, v) = v (data.h:140);
         * graph.h:149-167 add_vertex */
        const float *point = data + v * (uint64_t)dim;
        int got = oracle_search(data, v, dim, graph_out, FIXED_DEGREE_SHIFT, metric, point,
                                CONSTRUCT_SEARCH_BUDGET, nb, NULL, NULL);
        int num = got < SEARCH_DEGREE ? got : SEARCH_DEGREE;
        size_t offset = (size_t)v << FIXED_DEGREE_SHIFT;
        graph_out[offset] = (uint64_t)num;
        for (int i = 0; i < num; ++i) graph_out[offset + i + 1] = nb[i];
        rank_edges(&b, offset);
        for (int i = 0; i < num; ++i) add_edge(&b, nb[i], v);
    }
    free(b.edge_dist);
    return 0;
}

/* ----------------------------------------------------------------- parser */

/* parser.h:17-43 */
int oracle_parse_libsvm_line(const char *line, int one_based, int *idx, float *val, int cap) {
    size_t len = strlen(line);
    size_t ntok = 0, tcap = 64;
    size_t *tok = (size_t *)malloc(tcap * sizeof(size_t));
    for (size_t i = 0; i < len; ++i) {
        if (line[i] == ':' || line[i] == ' ') {
            if (ntok == tcap) {
                tcap *= 2;
                tok = (size_t *)realloc(tok, tcap * sizeof(size_t));
            }
            tok[ntok++] = i;
        }
    }
    int cnt = 0;
    for (size_t i = 0; i + 1 < ntok; i += 2) {
        int index = 0;
        double tmp = 0.0;
        sscanf(line + tok[i] + 1, "%d", &index);
        index -= one_based;
        sscanf(line + tok[i + 1] + 1, "%lf", &tmp);
        if (cnt < cap) {
            idx[cnt] = index;
            val[cnt] = (float)tmp;
        }
        ++cnt;
    }
    free(tok);
    return cnt;
}

/* ------------------------------------------------------- exact brute force */

static double exact_dist(int metric, const float *x, const float *q, int dim) {
    double s = 0.0, xx = 0.0, qq = 0.0;
    for (int i = 0; i < dim; ++i) {
        double a = x[i], b = q[i];
        if (metric == ORACLE_L2) s += (a - b) * (a - b);
        else {
            s += a * b;
            xx += a * a;
            qq += b * b;
        }
    }
    if (metric == ORACLE_L2) return s;
    if (metric == ORACLE_IP) return -s;
    return (s < 0 ? 1.0 : -1.0) * (s * s / xx / qq);
}

typedef struct {
    const float *data;
    uint64_t n;
    int dim, metric;
    const float *queries;
    uint64_t q0, q1;
    int k;
    uint64_t *ids;
    double *ds;
} bf_job_t;

static void *bf_worker(void *arg) {
    bf_job_t *j = (bf_job_t *)arg;
    heap_t h = {0};
    h.is_max = 1;
    for (uint64_t qi = j->q0; qi < j->q1; ++qi) {
        h.n = 0;
        const float *q = j->queries + qi * (uint64_t)j->dim;
        for (uint64_t v = 0; v < j->n; ++v) {
            pair_t p = {exact_dist(j->metric, j->data + v * (uint64_t)j->dim, q, j->dim), v};
            if ((int)h.n < j->k) heap_push(&h, p);
            else if (pair_lt(p, h.v[0])) {
                heap_pop(&h);
                heap_push(&h, p);
            }
        }
        int cnt = (int)h.n;
        for (int i = j->k - 1; i >= cnt; --i) {
            j->ids[qi * j->k + i] = UINT64_MAX;
            if (j->ds) j->ds[qi * j->k + i] = INFINITY;
        }
        for (int i = cnt - 1; i >= 0; --i) {
            j->ids[qi * j->k + i] = h.v[0].id;
            if (j->ds) j->ds[qi * j->k + i] = h.v[0].d;
            heap_pop(&h);
        }
    }
    free(h.v);
    return NULL;
}

int oracle_bruteforce_knn(const float *data, uint64_t n, int dim, int metric, const float *queries,
                          uint64_t nq, int k, uint64_t *out_ids, double *out_dists, int threads) {
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > nq) threads = nq ? (int)nq : 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    bf_job_t *jobs = (bf_job_t *)malloc(sizeof(bf_job_t) * threads);
    for (int t = 0; t < threads; ++t) {
        bf_job_t j = {data, n, dim, metric, queries, nq * t / threads, nq * (t + 1) / threads, k,
                      out_ids, out_dists};
        jobs[t] = j;
        if (threads == 1) bf_worker(&jobs[t]);
        else pthread_create(&th[t], NULL, bf_worker, &jobs[t]);
    }
    if (threads > 1)
        for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
    return 0;
}
