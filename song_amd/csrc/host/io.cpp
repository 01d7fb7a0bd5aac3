// LIBSVM parsing and raw bfsg files for the drop-in host (song_host.h).
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
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

// Parser (parser.h:12-64): every ':' or ' ' is a token boundary; token pairs (i, i+1)
// give an index ("%d", minus one_based) and a value ("%lf" cast to float).  Lines are
// read with fgets into a 10,000,000-byte buffer like the reference (MAX_LINE).
extern "C" int song_libsvm_parse(const char *path, int one_based, uint64_t rows, int dim, float *out,
                                 uint64_t *rows_read) {
    if (!path || !out || dim < 1) return host_fail(-1, "bad arguments");
    FILE *fp = fopen(path, "r");
    if (!fp) return host_fail(-1, "File not found at (%s)", path);
    const int kMaxLine = 10000000;
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
        const char *b = buff.get();
        tok.clear();
        for (size_t i = 0; b[i]; ++i)
            if (b[i] == ':' || b[i] == ' ') tok.push_back(i);
        float *row = out + idx * (uint64_t)dim;
        for (size_t i = 0; i + 1 < tok.size(); i += 2) {
            int index = 0;
            double tmp = 0.0;
            sscanf(b + tok[i] + 1, "%d", &index);
            index -= one_based;
            sscanf(b + tok[i + 1] + 1, "%lf", &tmp);
            if (index < 0 || index >= dim) {
                rc = host_fail(-1, "%s line %llu: feature index %d outside [0,%d)", path, (unsigned long long)idx + 1,
                               index + one_based, dim);
                break;
            }
            row[index] = (float)tmp;
        }
        if (rc) break;
        ++idx;
    }
    fclose(fp);
    if (rows_read) *rows_read = idx;
    return rc;
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
    fseek(fp, 0, SEEK_END);
    long long size = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    if (size != (long long)bytes) {
        fclose(fp);
        return host_fail(-1, "%s has %lld bytes, expected %llu", path, size, (unsigned long long)bytes);
    }
    size_t r = bytes ? fread(buf, 1, bytes, fp) : 0;
    fclose(fp);
    if (r != bytes) return host_fail(-1, "short read from %s", path);
    return 0;
}
