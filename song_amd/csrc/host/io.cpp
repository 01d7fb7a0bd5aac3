// Input path of the drop-in host (song_host.h): LIBSVM parsing (parallel, with the
// reference Parser's semantics), a LIBSVM writer, binary query formats and the raw bfsg
// files.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/song_host.h"
#include "host_common.hpp"

static thread_local std::string g_host_err;

int host_fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_host_err = buf;
    return code;
}

extern "C" const char *song_host_last_error(void) { return g_host_err.c_str(); }

int host_threads(int threads) {
    if (threads > 0) return threads;
    if (const char *e = getenv("SONG_THREADS")) {
        const int t = atoi(e);
        if (t > 0) return t;
    }
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hc ? hc : 1u));
}

namespace {

// fgets reads lines into a 10,000,000-byte buffer (parser.h:15), so a longer line arrives
// as several records; such files take the sequential path, which reads them the same way
constexpr int kMaxLine = 10000000;

// One record (NUL-terminated, as fgets leaves it) per parser.h:18-43: every ':' or ' ' is a
// token boundary; token pairs (i, i+1) give an index ("%d" = strtol cast to int, minus
// one_based) and a value ("%lf" = strtod, cast to float).  A field with no number reads as 0.
// Returns the first out-of-range index + one_based, or INT_MIN when all are in range.
int parse_record(const char *b, int one_based, int dim, float *row, std::vector<size_t> &tok) {
    tok.clear();
    for (size_t i = 0; b[i]; ++i)
        if (b[i] == ':' || b[i] == ' ') tok.push_back(i);
    for (size_t i = 0; i + 1 < tok.size(); i += 2) {
        const char *s1 = b + tok[i] + 1;
        char *e1 = nullptr;
        long iv = strtol(s1, &e1, 10);
        int index = e1 == s1 ? 0 : (int)iv;
        const char *s2 = b + tok[i + 1] + 1;
        char *e2 = nullptr;
        double tmp = strtod(s2, &e2);
        if (e2 == s2) tmp = 0.0;
        index -= one_based;
        if (index < 0 || index >= dim) return index + one_based;
        row[index] = (float)tmp;
    }
    return INT_MIN;
}

int parse_sequential(const char *path, int one_based, uint64_t rows, int dim, float *out, uint64_t *rows_read) {
    FILE *fp = fopen(path, "r");
    if (!fp) return host_fail(-1, "File not found at (%s)", path);
    std::unique_ptr<char[]> buff(new char[kMaxLine]);
    std::vector<size_t> tok;
    uint64_t idx = 0;
    std::memset(out, 0, sizeof(float) * rows * (uint64_t)dim);
    int rc = 0;
    while (fgets(buff.get(), kMaxLine, fp)) {
        if (idx >= rows) {
            rc = host_fail(-1, "%s: more than %llu lines", path, (unsigned long long)rows);
            break;
        }
        const int bad = parse_record(buff.get(), one_based, dim, out + idx * (uint64_t)dim, tok);
        if (bad != INT_MIN) {
            rc = host_fail(-1, "%s line %llu: feature index %d outside [0,%d)", path, (unsigned long long)idx + 1, bad,
                           dim);
            break;
        }
        ++idx;
    }
    fclose(fp);
    if (rows_read) *rows_read = idx;
    return rc;
}

struct Mapped {
    const char *p = nullptr;
    size_t n = 0;
    ~Mapped() {
        if (p && n) munmap((void *)p, n);
    }
};

int map_file(const char *path, Mapped *m) {
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return host_fail(-1, "File not found at (%s)", path);
    struct stat st;
    if (fstat(fd, &st) != 0) {
        close(fd);
        return host_fail(-1, "cannot stat %s", path);
    }
    m->n = (size_t)st.st_size;
    if (m->n) {
        void *p = mmap(nullptr, m->n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (p == MAP_FAILED) {
            close(fd);
            m->n = 0;
            return host_fail(-1, "cannot map %s: %s", path, strerror(errno));
        }
        madvise(p, m->n, MADV_SEQUENTIAL);
        m->p = static_cast<const char *>(p);
    }
    close(fd);
    return 0;
}

// records of [s, e) as fgets would read them (each line with its '\n'; a last line without one
// is a record too); *longest = longest record in bytes
uint64_t count_records(const char *s, const char *e, size_t *longest) {
    uint64_t c = 0;
    size_t lng = 0;
    for (const char *p = s; p < e;) {
        const char *nl = static_cast<const char *>(memchr(p, '\n', (size_t)(e - p)));
        const char *end = nl ? nl + 1 : e;
        lng = std::max(lng, (size_t)(end - p));
        ++c;
        p = end;
    }
    *longest = lng;
    return c;
}

}  // namespace

// Parser (parser.h:12-64): one record per line, parsed in parallel.  The file is mapped and
// cut into one chunk per thread at line starts; a first pass counts each chunk's records,
// their prefix sums give each chunk's first row, and a second pass parses every record into
// its row (copied with a terminating NUL, as fgets hands it to the reference's sscanf calls).
// Results equal the sequential reading bit for bit.
extern "C" int song_libsvm_parse_mt(const char *path, int one_based, uint64_t rows, int dim, int threads, float *out,
                                    uint64_t *rows_read) {
    if (!path || !out || dim < 1) return host_fail(-1, "bad arguments");
    Mapped m;
    int rc = map_file(path, &m);
    if (rc) return rc;
    const int T = (int)std::max<size_t>(1, std::min<size_t>((size_t)host_threads(threads), m.n / (1 << 16) + 1));
    std::vector<const char *> cut(T + 1);
    cut[0] = m.p;
    cut[T] = m.p + m.n;
    for (int t = 1; t < T; ++t) {
        const char *c = m.p + m.n * (size_t)t / (size_t)T;
        c = std::max(c, cut[t - 1]);
        const char *nl = c < cut[T] ? static_cast<const char *>(memchr(c, '\n', (size_t)(cut[T] - c))) : nullptr;
        cut[t] = nl ? nl + 1 : cut[T];
    }
    std::vector<uint64_t> cnt(T, 0), first(T + 1, 0);
    std::vector<size_t> longest(T, 0);
    {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] { cnt[t] = count_records(cut[t], cut[t + 1], &longest[t]); });
        for (auto &x : th) x.join();
    }
    if (*std::max_element(longest.begin(), longest.end()) >= (size_t)kMaxLine - 1)
        return parse_sequential(path, one_based, rows, dim, out, rows_read);
    for (int t = 0; t < T; ++t) first[t + 1] = first[t] + cnt[t];
    const uint64_t total = first[T];
    std::vector<uint64_t> bad_line(T, UINT64_MAX);
    std::vector<int> bad_index(T, 0);
    {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                // rows past the file's records stay zero (split among the threads)
                const uint64_t z0 = std::min(rows, total + (rows > total ? (rows - total) * t / T : 0));
                const uint64_t z1 = std::min(rows, total + (rows > total ? (rows - total) * (t + 1) / T : 0));
                if (z1 > z0) std::memset(out + z0 * (uint64_t)dim, 0, sizeof(float) * (z1 - z0) * (uint64_t)dim);
                std::vector<char> buf;
                std::vector<size_t> tok;
                uint64_t r = first[t];
                for (const char *p = cut[t]; p < cut[t + 1] && r < rows; ++r) {
                    const char *nl = static_cast<const char *>(memchr(p, '\n', (size_t)(cut[t + 1] - p)));
                    const char *end = nl ? nl + 1 : cut[t + 1];
                    buf.assign(p, end);
                    buf.push_back('\0');
                    float *row = out + r * (uint64_t)dim;
                    std::memset(row, 0, sizeof(float) * (size_t)dim);
                    const int bad = parse_record(buf.data(), one_based, dim, row, tok);
                    if (bad != INT_MIN) {
                        bad_line[t] = r;
                        bad_index[t] = bad;
                        return;
                    }
                    p = end;
                }
            });
        for (auto &x : th) x.join();
    }
    // the first error in file order, as the sequential reading reports it
    for (int t = 0; t < T; ++t) {
        if (bad_line[t] == UINT64_MAX) continue;
        if (rows_read) *rows_read = bad_line[t];
        return host_fail(-1, "%s line %llu: feature index %d outside [0,%d)", path,
                         (unsigned long long)bad_line[t] + 1, bad_index[t], dim);
    }
    if (total > rows) {
        if (rows_read) *rows_read = rows;
        return host_fail(-1, "%s: more than %llu lines", path, (unsigned long long)rows);
    }
    if (rows_read) *rows_read = total;
    return 0;
}

extern "C" int song_libsvm_parse(const char *path, int one_based, uint64_t rows, int dim, float *out,
                                 uint64_t *rows_read) {
    return song_libsvm_parse_mt(path, one_based, rows, dim, 0, out, rows_read);
}

// "0 i:v ..." per row, zeros omitted, one-based indices; integer values as integers, others
// with 9 significant digits (enough for the float to read back exactly through %lf and a cast)
extern "C" int song_libsvm_write(const char *path, const float *data, uint64_t rows, int dim, int threads) {
    if (!path || (!data && rows) || dim < 1) return host_fail(-1, "bad arguments");
    FILE *fp = fopen(path, "wb");
    if (!fp) return host_fail(-1, "cannot open %s for writing", path);
    const int T = host_threads(threads);
    const uint64_t block = 8192;
    std::vector<std::string> part(T);
    int rc = 0;
    for (uint64_t r0 = 0; r0 < rows && rc == 0; r0 += block * (uint64_t)T) {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                std::string &s = part[t];
                s.clear();
                char tmp[64];
                const uint64_t a = std::min(rows, r0 + block * t), b = std::min(rows, r0 + block * (t + 1));
                for (uint64_t r = a; r < b; ++r) {
                    s += '0';
                    const float *x = data + r * (uint64_t)dim;
                    for (int j = 0; j < dim; ++j) {
                        const float v = x[j];
                        if (v == 0.0f) continue;
                        int n;
                        if (v == std::rint(v) && std::fabs(v) < 2147483648.0f)
                            n = snprintf(tmp, sizeof(tmp), " %d:%d", j + 1, (int)v);
                        else
                            n = snprintf(tmp, sizeof(tmp), " %d:%.9g", j + 1, (double)v);
                        s.append(tmp, (size_t)n);
                    }
                    s += '\n';
                }
            });
        for (auto &x : th) x.join();
        for (int t = 0; t < T && rc == 0; ++t)
            if (!part[t].empty() && fwrite(part[t].data(), 1, part[t].size(), fp) != part[t].size())
                rc = host_fail(-1, "short write to %s", path);
    }
    if (fclose(fp) != 0 && rc == 0) rc = host_fail(-1, "short write to %s", path);
    return rc;
}

// Query files (no counterpart in the reference, whose only query format is LIBSVM text,
// parser.h:57-62).  out == NULL: only the row count.
extern "C" int song_query_read(const char *path, int format, int dim, uint64_t max_rows, float *out,
                               uint64_t *rows_read) {
    if (!path || dim < 1 || !rows_read) return host_fail(-1, "bad arguments");
    *rows_read = 0;
    if (format == SONG_QUERY_LIBSVM) {
        if (!out) {
            Mapped m;
            int rc = map_file(path, &m);
            if (rc) return rc;
            size_t lng = 0;
            *rows_read = count_records(m.p, m.p + m.n, &lng);
            return 0;
        }
        return song_libsvm_parse(path, 1, max_rows, dim, out, rows_read);
    }
    if (format != SONG_QUERY_F32 && format != SONG_QUERY_FVECS && format != SONG_QUERY_FBIN)
        return host_fail(-1, "query format %d", format);
    FILE *fp = fopen(path, "rb");
    if (!fp) return host_fail(-1, "File not found at (%s)", path);
    std::unique_ptr<FILE, int (*)(FILE *)> guard(fp, fclose);
    fseeko(fp, 0, SEEK_END);
    const uint64_t size = (uint64_t)ftello(fp);
    fseeko(fp, 0, SEEK_SET);
    const uint64_t rb = 4ull * (uint64_t)dim;
    uint64_t n = 0, head = 0, stride = rb;
    if (format == SONG_QUERY_F32) {
        if (size % rb) return host_fail(-1, "%s: %llu bytes is not a whole number of %d-float rows", path,
                                        (unsigned long long)size, dim);
        n = size / rb;
    } else if (format == SONG_QUERY_FVECS) {
        stride = rb + 4;
        if (size % stride)
            return host_fail(-1, "%s: %llu bytes is not a whole number of fvecs rows of dim %d", path,
                             (unsigned long long)size, dim);
        n = size / stride;
    } else {
        uint32_t h[2];
        if (size < 8 || fread(h, 4, 2, fp) != 2) return host_fail(-1, "%s: no fbin header", path);
        if ((int)h[1] != dim) return host_fail(-1, "%s: fbin dim %u, expected %d", path, h[1], dim);
        if (size != 8 + (uint64_t)h[0] * rb)
            return host_fail(-1, "%s: %llu bytes, the fbin header says %u rows of %d floats", path,
                             (unsigned long long)size, h[0], dim);
        n = h[0];
        head = 8;
    }
    if (!out) {
        *rows_read = n;
        return 0;
    }
    if (n > max_rows) return host_fail(-1, "%s: %llu rows, room for %llu", path, (unsigned long long)n,
                                       (unsigned long long)max_rows);
    fseeko(fp, (off_t)head, SEEK_SET);
    if (format == SONG_QUERY_FVECS) {
        for (uint64_t r = 0; r < n; ++r) {
            int32_t d = 0;
            if (fread(&d, 4, 1, fp) != 1 || fread(out + r * (uint64_t)dim, 4, (size_t)dim, fp) != (size_t)dim)
                return host_fail(-1, "short read from %s", path);
            if (d != dim) return host_fail(-1, "%s row %llu: dimension %d, expected %d", path,
                                           (unsigned long long)r, d, dim);
        }
    } else if (n && fread(out, rb, (size_t)n, fp) != (size_t)n) {
        return host_fail(-1, "short read from %s", path);
    }
    *rows_read = n;
    return 0;
}

extern "C" int song_file_write(const char *path, const void *buf, uint64_t bytes) {
    FILE *fp = fopen(path, "wb");
    if (!fp) return host_fail(-1, "cannot open %s for writing", path);
    size_t w = bytes ? fwrite(buf, 1, bytes, fp) : 0;
    int c = fclose(fp);
    if (w != bytes || c != 0) return host_fail(-1, "short write to %s", path);
    return 0;
}

extern "C" int song_file_read(const char *path, void *buf, uint64_t bytes) {
    FILE *fp = fopen(path, "rb");
    if (!fp) return host_fail(-1, "File not found at (%s)", path);
    fseeko(fp, 0, SEEK_END);
    long long size = (long long)ftello(fp);
    fseeko(fp, 0, SEEK_SET);
    if (size != (long long)bytes) {
        fclose(fp);
        return host_fail(-1, "%s has %lld bytes, expected %llu", path, size, (unsigned long long)bytes);
    }
    size_t r = bytes ? fread(buf, 1, bytes, fp) : 0;
    fclose(fp);
    if (r != bytes) return host_fail(-1, "short read from %s", path);
    return 0;
}
