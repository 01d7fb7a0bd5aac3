// Learned form: the entry-vertex / stop-step MLP on the matrix cores (include/song_learned.h).
//
// Matrix products use v_mfma_f32_16x16x16_bf16 (BF16 inputs, FP32 accumulation).  Fragment
// layout of one 16x16x16 product for wave lane l (verified by the GEMM test against numpy):
//   A[i][k]: i = l % 16, k = 4*(l/16) + j     (j = 0..3, the 4 bf16 of the lane's operand)
//   B[k][n]: n = l % 16, k = 4*(l/16) + j
//   D[i][n]: n = l % 16, i = 4*(l/16) + j
// Training runs as separate kernels (GEMM with bias/ReLU/mask epilogues, softmax
// cross-entropy, column sums, Adam); inference is one fused kernel, 16 queries per workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/song_learned.h"

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

static thread_local std::string g_mlp_err;

static int mfail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_mlp_err = buf;
    return code;
}

#define MHIP(expr)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return mfail(-2, "%s: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

extern "C" const char *song_mlp_last_error(void) { return g_mlp_err.c_str(); }

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ---------------------------------------------------------------- GEMM

struct GemmArgs {
    int M, N, K;
    const uint16_t *A;
    long long sam, sak;
    const uint16_t *B;
    long long sbk, sbn;
    float *C;  // float32 output or null
    int ldc;
    uint16_t *Cb;  // bf16 output or null
    int ldcb;
    const float *bias;      // per column, or null
    const uint16_t *mask;   // zero the output where mask[m][n] <= 0 (ReLU backward), or null
    int ldm;
    int relu;
    float alpha;
};

constexpr int GBM = 64, GBN = 64, GBK = 16;

// C[m][n] = alpha * sum_k A(m,k) B(k,n) (+ bias[n]) (ReLU) (mask), A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn].  256 threads: 2x2 waves of 32x32, each 2x2 MFMA tiles.
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
    __shared__ uint16_t As[GBM][GBK + 4];
    __shared__ uint16_t Bs[GBN][GBK + 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    v4f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += GBK) {
        for (int e = tid; e < GBM * GBK; e += 256) {
            const int r = e / GBK, c = e % GBK, m = m0 + r, k = k0 + c;
            As[r][c] = (m < g.M && k < g.K) ? g.A[(long long)m * g.sam + (long long)k * g.sak] : (uint16_t)0;
        }
        for (int e = tid; e < GBN * GBK; e += 256) {
            const int r = e / GBK, c = e % GBK, n = n0 + r, k = k0 + c;
            Bs[r][c] = (n < g.N && k < g.K) ? g.B[(long long)k * g.sbk + (long long)n * g.sbn] : (uint16_t)0;
        }
        __syncthreads();
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) {
                v4s a, b;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    a[j] = (short)As[wm + mi * 16 + (lane & 15)][4 * (lane >> 4) + j];
                    b[j] = (short)Bs[wn + ni * 16 + (lane & 15)][4 * (lane >> 4) + j];
                }
                acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[mi][ni], 0, 0, 0);
            }
        __syncthreads();
    }
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = m0 + wm + mi * 16 + 4 * (lane >> 4) + j;
                const int n = n0 + wn + ni * 16 + (lane & 15);
                if (m < g.M && n < g.N) {
                    float v = g.alpha * acc[mi][ni][j];
                    if (g.bias) v += g.bias[n];
                    if (g.relu) v = fmaxf(v, 0.0f);
                    if (g.mask && bf2f(g.mask[(long long)m * g.ldm + n]) <= 0.0f) v = 0.0f;
                    if (g.C) g.C[(long long)m * g.ldc + n] = v;
                    if (g.Cb) g.Cb[(long long)m * g.ldcb + n] = f2bf(v);
                }
            }
}

static int gemm(const GemmArgs &g, hipStream_t st) {
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return 0;
    dim3 grid((g.N + GBN - 1) / GBN, (g.M + GBM - 1) / GBM);
    hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, st, g);
    MHIP(hipGetLastError());
    return 0;
}

static GemmArgs gemm_args(int M, int N, int K, const uint16_t *A, long long sam, long long sak, const uint16_t *B,
                          long long sbk, long long sbn) {
    GemmArgs g{};
    g.M = M;
    g.N = N;
    g.K = K;
    g.A = A;
    g.sam = sam;
    g.sak = sak;
    g.B = B;
    g.sbk = sbk;
    g.sbn = sbn;
    g.alpha = 1.0f;
    return g;
}

extern "C" int song_gemm_bf16(int M, int N, int K, const uint16_t *d_a, int sam, int sak, const uint16_t *d_b, int sbk,
                              int sbn, float *d_c, int ldc, float alpha, void *stream) {
    GemmArgs g = gemm_args(M, N, K, d_a, sam, sak, d_b, sbk, sbn);
    g.C = d_c;
    g.ldc = ldc;
    g.alpha = alpha;
    return gemm(g, (hipStream_t)stream);
}

// ---------------------------------------------------------------- loss, sums, Adam

// One 64-lane block per row: softmax cross-entropy over the first C logits against the
// label, squared error of logit C against the stop target; writes dZ (bf16) and adds the
// batch means of both losses to loss[0..1].
__global__ __launch_bounds__(64) void loss_kernel(const float *Z, int ldz, int C, const int *labels, const float *stop_t,
                                                  int B, float stop_w, uint16_t *dZ, int lddz, float *loss) {
    const int row = blockIdx.x, lane = threadIdx.x;
    const float *z = Z + (long long)row * ldz;
    float mx = -INFINITY;
    for (int c = lane; c < C; c += 64) mx = fmaxf(mx, z[c]);
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    float se = 0.0f;
    for (int c = lane; c < C; c += 64) se += expf(z[c] - mx);
    for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o);
    const int lab = labels[row];
    const float inv = 1.0f / (float)B;
    for (int c = lane; c < C; c += 64) {
        const float p = expf(z[c] - mx) / se;
        dZ[(long long)row * lddz + c] = f2bf((p - (c == lab ? 1.0f : 0.0f)) * inv);
    }
    if (lane == 0) {
        const float diff = z[C] - stop_t[row];
        dZ[(long long)row * lddz + C] = f2bf(2.0f * stop_w * diff * inv);
        atomicAdd(&loss[0], (logf(se) + mx - z[lab]) * inv);
        atomicAdd(&loss[1], diff * diff * inv);
    }
}

// out[n] = sum over rows of G[m][n] (bf16 rows)
__global__ void colsum_kernel(const uint16_t *G, int M, int N, int ldg, float *out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int m = 0; m < M; ++m) s += bf2f(G[(long long)m * ldg + n]);
    out[n] = s;
}

__global__ void adam_kernel(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                            float c1, float c2, float eps) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void to_bf16_kernel(const float *x, uint16_t *y, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// ---------------------------------------------------------------- features

__global__ __launch_bounds__(64) void features_kernel(const float *Q, int dim, const float *mean, const float *inv_std,
                                                      const float *v0, float f0, float f1, float f2, float f3,
                                                      float f4, float f5, float f6, float f7, uint16_t *X,
                                                      int width) {
    const int row = blockIdx.x, lane = threadIdx.x;
    const float *q = Q + (long long)row * dim;
    uint16_t *x = X + (long long)row * width;
    float s = 0.f, s2 = 0.f, dv = 0.f, mn = INFINITY, mx = -INFINITY, nz = 0.f;
    for (int i = lane; i < dim; i += 64) {
        const float v = q[i];
        x[i] = f2bf((v - mean[i]) * inv_std[i]);
        s += v;
        s2 += v * v;
        const float d = v - v0[i];
        dv += d * d;
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
        nz += (v == 0.0f) ? 1.0f : 0.0f;
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o);
        s2 += __shfl_xor(s2, o);
        dv += __shfl_xor(dv, o);
        mn = fminf(mn, __shfl_xor(mn, o));
        mx = fmaxf(mx, __shfl_xor(mx, o));
        nz += __shfl_xor(nz, o);
    }
    if (lane < 8) {
        const float meanq = s / dim;
        const float var = fmaxf(s2 / dim - meanq * meanq, 0.0f);
        float f;
        switch (lane) {
        case 0: f = sqrtf(s2) * f0; break;
        case 1: f = sqrtf(dv) * f1; break;
        case 2: f = meanq * f2; break;
        case 3: f = sqrtf(var) * f3; break;
        case 4: f = mn * f4; break;
        case 5: f = mx * f5; break;
        case 6: f = nz / dim * f6; break;
        default: f = f7; break;
        }
        x[dim + lane] = f2bf(f);
    }
    for (int i = dim + 8 + lane; i < width; i += 64) x[i] = 0;
}

// ---------------------------------------------------------------- fused inference

// queries per workgroup: 16 rows (MI = 1 row tile) -- a 10k-query batch launches 625 workgroups,
// two or three per CU, and each wave's chain of MFMAs and softmax updates is half that of 32 rows;
// the output weights are read twice as often, from L2.  For 10k queries: 16 rows 197 us, 32 rows
// 256 us, 64 rows 447 us (157 workgroups, one wave per SIMD)
constexpr int IQ = 16;
constexpr int MI = IQ / 16;
// waves per workgroup: they split the output layer's columns (64 per wave and chunk) and the
// hidden units (slices of 64); with 16 rows, 2 waves per group measured 319 us and 8 waves 206 us
constexpr int IW = 4;
constexpr int IT = 64 * IW;  // threads per workgroup

struct InferArgs {
    const uint16_t *X;
    int batch, width, hidden, C;
    const uint16_t *W1, *W2, *W3;
    const uint16_t *W3i;  // W3 in the k-quad interleaved layout, npad columns
    int npad;
    const float *b1, *b2, *b3;
    int *top2;  // batch x 2 pivot indices (best first)
    float *conf, *stop, *logits;
    const int *nbr;  // C x n_nbr pivot neighbourhoods (row a starts with a) or null
    int n_nbr;
};

// One dense layer on the LDS tile: accumulators of in[IQ rows] . W[n]^T for the 64-column
// slice [n0, n0+64) of this wave (MI x 4 MFMA tiles).  K is a multiple of 16, so every
// fragment (4 consecutive k of one row or column) is one aligned 8-byte load; the next
// k-step's weight fragments are loaded while the current step's MFMAs run (a deeper ring of
// in-flight weight loads measured no faster).
__device__ __forceinline__ void tile_layer(const uint16_t *in, int ldin, int K, const uint16_t *W, int N, int n0,
                                           v4f (&acc)[MI][4]) {
    const int lane = threadIdx.x & 63;
    const int kq = 4 * (lane >> 4);
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    const uint16_t *wrow[4];
    bool ok[4];
    v4s bf[4];
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
        const int n = n0 + ni * 16 + (lane & 15);
        ok[ni] = n < N;
        wrow[ni] = W + (long long)(ok[ni] ? n : 0) * K + kq;
        bf[ni] = ok[ni] ? *reinterpret_cast<const v4s *>(wrow[ni]) : v4s{0, 0, 0, 0};
    }
    const uint16_t *arow = in + (lane & 15) * ldin + kq;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bool more = k0 + 16 < K;
        v4s nb[4];
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
            nb[ni] = (ok[ni] && more) ? *reinterpret_cast<const v4s *>(wrow[ni] + k0 + 16) : v4s{0, 0, 0, 0};
#pragma unroll
        for (int mi = 0; mi < MI; ++mi) {
            const v4s af = *reinterpret_cast<const v4s *>(arow + mi * 16 * ldin + k0);
#pragma unroll
            for (int ni = 0; ni < 4; ++ni)
                acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bf[ni], acc[mi][ni], 0, 0, 0);
        }
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) bf[ni] = nb[ni];
    }
}

// As tile_layer, with the weights in the k-quad interleaved layout (interleave_kernel): a lane
// group's 16 fragments of one step are one contiguous 128-byte run (four full cache lines per
// load instead of 32 bytes of sixteen).  The lanes take the same k as in tile_layer, so the
// accumulators are identical.
__device__ __forceinline__ void tile_layer_il(const uint16_t *in, int ldin, int K, const uint16_t *Wi, int N, int npad,
                                              int n0, v4f (&acc)[MI][4]) {
    const int lane = threadIdx.x & 63;
    const int q = lane >> 4, kq = 4 * q;
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    const long long sstride = 16LL * npad;  // elements per k-step: 4 quads x npad columns x 4
    const uint16_t *wb[4];
    bool ok[4];
    v4s bf[4];
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
        const int n = n0 + ni * 16 + (lane & 15);
        ok[ni] = n < N;
        wb[ni] = Wi + ((long long)q * npad + (ok[ni] ? n : 0)) * 4;
        bf[ni] = ok[ni] ? *reinterpret_cast<const v4s *>(wb[ni]) : v4s{0, 0, 0, 0};
    }
    const uint16_t *arow = in + (lane & 15) * ldin + kq;
    for (int k0 = 0, st = 0; k0 < K; k0 += 16, ++st) {
        const bool more = k0 + 16 < K;
        v4s nb[4];
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
            nb[ni] = (ok[ni] && more) ? *reinterpret_cast<const v4s *>(wb[ni] + (st + 1) * sstride) : v4s{0, 0, 0, 0};
#pragma unroll
        for (int mi = 0; mi < MI; ++mi) {
            const v4s af = *reinterpret_cast<const v4s *>(arow + mi * 16 * ldin + k0);
#pragma unroll
            for (int ni = 0; ni < 4; ++ni)
                acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bf[ni], acc[mi][ni], 0, 0, 0);
        }
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) bf[ni] = nb[ni];
    }
}

// running softmax state of a row: top two (value, index) and the sum of exp(v - v1)
struct Soft {
    float v1, v2, s;
    int a1, a2;
};
__device__ __forceinline__ bool better(float v, int a, float w, int b) { return v > w || (v == w && a < b); }
__device__ __forceinline__ Soft soft_empty() { return Soft{-INFINITY, -INFINITY, 0.0f, 0x7fffffff, 0x7fffffff}; }
// the (value, index) pair into the top two, without branches
__device__ __forceinline__ void top2_add(Soft &t, float v, int a) {
    const bool b1 = better(v, a, t.v1, t.a1), b2 = better(v, a, t.v2, t.a2);
    const float nv2 = b1 ? t.v1 : (b2 ? v : t.v2);
    const int na2 = b1 ? t.a1 : (b2 ? a : t.a2);
    t.v1 = b1 ? v : t.v1;
    t.a1 = b1 ? a : t.a1;
    t.v2 = nv2;
    t.a2 = na2;
}
// four values of one row at once (-inf where absent): one rescale of the running sum to the new
// maximum, four exponentials, then the top two
__device__ __forceinline__ void soft_add4(Soft &t, const float (&v)[4], const int (&a)[4]) {
    const float m = fmaxf(fmaxf(v[0], v[1]), fmaxf(v[2], v[3]));
    if (m == -INFINITY) return;
    const float ref = fmaxf(m, t.v1);
    float s = t.v1 == -INFINITY ? 0.0f : t.s * expf(t.v1 - ref);
#pragma unroll
    for (int i = 0; i < 4; ++i) s += expf(v[i] - ref);  // exp(-inf) = 0
    t.s = s;
    // the sum is kept relative to the top value: ref becomes v1 below whenever m > v1
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (v[i] != -INFINITY) top2_add(t, v[i], a[i]);
}
__device__ __forceinline__ Soft soft_merge(const Soft &x, const Soft &y) {
    if (y.v1 == -INFINITY) return x;
    if (x.v1 == -INFINITY) return y;
    Soft t;
    const bool xf = better(x.v1, x.a1, y.v1, y.a1);
    // field by field (a conditional over the two structs would select their addresses, which
    // puts them in scratch memory)
    Soft hi, lo;
    hi.v1 = xf ? x.v1 : y.v1;
    hi.v2 = xf ? x.v2 : y.v2;
    hi.s = xf ? x.s : y.s;
    hi.a1 = xf ? x.a1 : y.a1;
    hi.a2 = xf ? x.a2 : y.a2;
    lo.v1 = xf ? y.v1 : x.v1;
    lo.v2 = xf ? y.v2 : x.v2;
    lo.s = xf ? y.s : x.s;
    lo.a1 = xf ? y.a1 : x.a1;
    lo.a2 = xf ? y.a2 : x.a2;
    t.v1 = hi.v1;
    t.a1 = hi.a1;
    t.s = hi.s + lo.s * expf(lo.v1 - hi.v1);
    if (better(hi.v2, hi.a2, lo.v1, lo.a1)) {
        t.v2 = hi.v2;
        t.a2 = hi.a2;
    } else {
        t.v2 = lo.v1;
        t.a2 = lo.a1;
    }
    return t;
}

__global__ __launch_bounds__(IT) void infer_kernel(InferArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = blockIdx.x * IQ;
    const int ldx = a.width + 8, ldh = a.hidden + 8;
    uint16_t *X = reinterpret_cast<uint16_t *>(smem);
    uint16_t *H1 = X + IQ * ldx;
    uint16_t *H2 = H1 + IQ * ldh;
    Soft *part = reinterpret_cast<Soft *>(H2 + IQ * ldh);  // [IW waves][IQ]
    Soft *run = part + IW * IQ;                             // [IQ]
    float *stp = reinterpret_cast<float *>(run + IQ);       // [IQ]

    for (int e = tid; e < IQ * a.width; e += IT) {
        const int r = e / a.width, c = e % a.width;
        X[r * ldx + c] = (r0 + r < a.batch) ? a.X[(long long)(r0 + r) * a.width + c] : (uint16_t)0;
    }
    for (int r = tid; r < IQ; r += IT) run[r] = soft_empty();
    __syncthreads();

    v4f acc[MI][4];
    // hidden layers: the waves split the hidden units in slices of 64
    for (int layer = 0; layer < 2; ++layer) {
        const uint16_t *in = layer == 0 ? X : H1;
        const int ldin = layer == 0 ? ldx : ldh, K = layer == 0 ? a.width : a.hidden;
        const uint16_t *W = layer == 0 ? a.W1 : a.W2;
        const float *bias = layer == 0 ? a.b1 : a.b2;
        uint16_t *out = layer == 0 ? H1 : H2;
        for (int n0 = wave * 64; n0 < a.hidden; n0 += IT) {
            tile_layer(in, ldin, K, W, a.hidden, n0, acc);
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                for (int ni = 0; ni < 4; ++ni)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int r = mi * 16 + 4 * (lane >> 4) + j, n = n0 + ni * 16 + (lane & 15);
                        if (n < a.hidden) out[r * ldh + n] = f2bf(fmaxf(acc[mi][ni][j] + bias[n], 0.0f));
                    }
        }
        __syncthreads();
    }

    // output layer in chunks of 64 columns per wave: each lane keeps a running softmax state (top
    // two and the sum) for each of its rows over the columns it owns; the 16 lanes of a row and
    // the waves combine once at the end
    const int NO = a.C + 1;
    Soft ts[MI][4];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int j = 0; j < 4; ++j) ts[mi][j] = soft_empty();
    for (int c0 = 0; c0 < NO; c0 += IT) {
        const int n0 = c0 + wave * 64;
        tile_layer_il(H2, ldh, a.hidden, a.W3i, NO, a.npad, n0, acc);
#pragma unroll
        for (int mi = 0; mi < MI; ++mi)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = mi * 16 + 4 * (lane >> 4) + j;
                float vv[4];
                int nn[4];
#pragma unroll
                for (int ni = 0; ni < 4; ++ni) {
                    const int n = n0 + ni * 16 + (lane & 15);
                    nn[ni] = n;
                    vv[ni] = -INFINITY;
                    if (n >= NO) continue;
                    const float v = acc[mi][ni][j] + a.b3[n];
                    if (a.logits && r0 + r < a.batch) a.logits[(long long)(r0 + r) * NO + n] = v;
                    if (n == a.C) stp[r] = v;
                    else vv[ni] = v;
                }
                soft_add4(ts[mi][j], vv, nn);
            }
    }
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = mi * 16 + 4 * (lane >> 4) + j;
            Soft t = ts[mi][j];
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                Soft u;
                u.v1 = __shfl_xor(t.v1, o);
                u.v2 = __shfl_xor(t.v2, o);
                u.s = __shfl_xor(t.s, o);
                u.a1 = __shfl_xor(t.a1, o);
                u.a2 = __shfl_xor(t.a2, o);
                t = soft_merge(t, u);
            }
            if ((lane & 15) == 0) part[wave * IQ + r] = t;
        }
    __syncthreads();
    if (tid < IQ) {
        Soft t = soft_empty();
        for (int w = 0; w < IW; ++w) t = soft_merge(t, part[w * IQ + tid]);
        run[tid] = t;
    }
    __syncthreads();
    // confidence: softmax mass of the best pivot's neighbourhood, its logits recomputed from
    // the last hidden layer by TPR threads per row (entries e = quarter, quarter + TPR, ...);
    // without neighbourhoods, the mass of the top two
    {
        constexpr int TPR = IT / IQ;
        const int row = tid / TPR, quarter = tid % TPR;
        const Soft t = run[row];
        float mass = 0.0f;
        if (a.nbr && t.a1 >= 0 && t.a1 < a.C) {  // (no valid best pivot: non-finite logits)
            for (int e = quarter; e < a.n_nbr; e += TPR) {
                const int p = a.nbr[(long long)t.a1 * a.n_nbr + e];
                const uint16_t *w = a.W3 + (long long)p * a.hidden;
                const uint16_t *hrow = H2 + row * ldh;
                float z = a.b3[p];
                for (int h = 0; h < a.hidden; h += 4) {
                    const v4s wv = *reinterpret_cast<const v4s *>(w + h);
                    const v4s hv = *reinterpret_cast<const v4s *>(hrow + h);
#pragma unroll
                    for (int q = 0; q < 4; ++q) z += bf2f((uint16_t)hv[q]) * bf2f((uint16_t)wv[q]);
                }
                mass += expf(fminf(z - t.v1, 0.0f));
            }
        } else if (!a.nbr && quarter == 0) {
            mass = 1.0f + (t.v2 == -INFINITY ? 0.0f : expf(t.v2 - t.v1));
        }
#pragma unroll
        for (int o = 1; o < TPR; o <<= 1) mass += __shfl_xor(mass, o);
        if (quarter == 0 && r0 + row < a.batch) {
            a.top2[2 * (r0 + row)] = t.a1;
            a.top2[2 * (r0 + row) + 1] = t.a2;
            a.conf[r0 + row] = fminf(mass / t.s, 1.0f);
            a.stop[r0 + row] = stp[row];
        }
    }
}

// ---------------------------------------------------------------- host object

struct song_mlp {
    int in_dim = 0, width = 0, hidden = 0, C = 0, max_batch = 0;
    long long off[6] = {0}, cnt[6] = {0}, total = 0;  // W1 b1 W2 b2 W3 b3
    float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;
    uint16_t *Wb[3] = {nullptr, nullptr, nullptr};
    uint16_t *W3i = nullptr;  // inference copy of the output weights, k-quad interleaved (infer_kernel)
    int npad = 0;             // its column count: C + 1 rounded up to 16
    uint16_t *H1 = nullptr, *H2 = nullptr, *dZ = nullptr, *dH2 = nullptr, *dH1 = nullptr;
    float *Z = nullptr, *loss = nullptr;
    int t = 0;
};

static uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// The output weights W [N][K] (bf16, row-major) in the layout the inference GEMM reads: fragment
// (k-step s, k-quad q, column n) = W[n][16s + 4q .. +4) at Wi[((4s + q) * npad + n) * 4], zero for
// n >= N, so the 16 columns one lane group reads per step are one aligned 128-byte run.
__global__ __launch_bounds__(256) void interleave_kernel(const uint16_t *W, uint16_t *Wi, int N, int K, int npad) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)K * npad) return;
    const int t = (int)(e & 3);
    const long long g = e >> 2;
    const int n = (int)(g % npad);
    const int sq = (int)(g / npad);
    const int k = 16 * (sq >> 2) + 4 * (sq & 3) + t;
    Wi[e] = n < N ? W[(long long)n * K + k] : (uint16_t)0;
}

static int refresh_bf16(song_mlp *m, hipStream_t st) {
    for (int l = 0; l < 3; ++l) {
        const long long n = m->cnt[2 * l];
        hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, m->P + m->off[2 * l],
                           m->Wb[l], n);
    }
    const long long ni = (long long)m->hidden * m->npad;
    hipLaunchKernelGGL(interleave_kernel, dim3((unsigned)((ni + 255) / 256)), dim3(256), 0, st, m->Wb[2], m->W3i,
                       m->C + 1, m->hidden, m->npad);
    MHIP(hipGetLastError());
    return 0;
}

extern "C" int song_mlp_create(int in_dim, int hidden, int n_classes, int max_batch, uint64_t seed, song_mlp **out) {
    if (!out || in_dim < 1 || hidden < 16 || hidden % 16 != 0 || n_classes < 2 || max_batch < 1)
        return mfail(-1, "bad MLP shape in=%d hidden=%d classes=%d batch=%d", in_dim, hidden, n_classes, max_batch);
    song_mlp *m = new song_mlp();
    m->in_dim = in_dim;
    m->width = (in_dim + 15) & ~15;
    m->hidden = hidden;
    m->C = n_classes;
    m->max_batch = max_batch;
    const long long sizes[6] = {(long long)hidden * m->width, hidden, (long long)hidden * hidden, hidden,
                                (long long)(n_classes + 1) * hidden, n_classes + 1};
    long long o = 0;
    for (int i = 0; i < 6; ++i) {
        m->off[i] = o;
        m->cnt[i] = sizes[i];
        o += sizes[i];
    }
    m->total = o;
    std::vector<float> init(o, 0.0f);
    uint64_t s = seed;
    const int fan[3] = {in_dim, hidden, hidden};
    for (int l = 0; l < 3; ++l) {
        const float lim = std::sqrt(6.0f / (float)fan[l]);
        const int rows = (int)(sizes[2 * l] / (l == 0 ? m->width : hidden));
        const int cols = l == 0 ? m->width : hidden;
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) {
                const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
                init[m->off[2 * l] + (long long)r * cols + c] = (l == 0 && c >= in_dim) ? 0.0f : (float)((2.0 * u - 1.0) * lim);
            }
    }
    auto bad = [&](const char *what) {
        song_mlp_destroy(m);
        return mfail(-2, "hipMalloc %s", what);
    };
    if (hipMalloc(&m->P, o * 4) || hipMalloc(&m->G, o * 4) || hipMalloc(&m->Mo, o * 4) || hipMalloc(&m->Vo, o * 4))
        return bad("params");
    for (int l = 0; l < 3; ++l)
        if (hipMalloc(&m->Wb[l], m->cnt[2 * l] * 2)) return bad("bf16 weights");
    m->npad = (n_classes + 1 + 15) & ~15;
    if (hipMalloc(&m->W3i, (size_t)m->npad * hidden * 2)) return bad("interleaved output weights");
    const long long B = max_batch, NO = n_classes + 1;
    if (hipMalloc(&m->H1, B * hidden * 2) || hipMalloc(&m->H2, B * hidden * 2) || hipMalloc(&m->dZ, B * NO * 2) ||
        hipMalloc(&m->dH2, B * hidden * 2) || hipMalloc(&m->dH1, B * hidden * 2) || hipMalloc(&m->Z, B * NO * 4) ||
        hipMalloc(&m->loss, 8))
        return bad("activations");
    if (hipMemcpy(m->P, init.data(), o * 4, hipMemcpyHostToDevice) || hipMemset(m->Mo, 0, o * 4) ||
        hipMemset(m->Vo, 0, o * 4) || hipMemset(m->G, 0, o * 4)) {
        song_mlp_destroy(m);
        return mfail(-2, "init upload");
    }
    if (refresh_bf16(m, 0) || hipDeviceSynchronize()) {
        song_mlp_destroy(m);
        return mfail(-2, "init bf16");
    }
    *out = m;
    return 0;
}

extern "C" int song_mlp_destroy(song_mlp *m) {
    if (!m) return 0;
    void *ptrs[] = {m->P, m->G, m->Mo, m->Vo, m->Wb[0], m->Wb[1], m->Wb[2], m->W3i, m->H1, m->H2, m->dZ, m->dH2,
                    m->dH1, m->Z, m->loss};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete m;
    return 0;
}

extern "C" int song_mlp_input_width(const song_mlp *m) { return m ? m->width : 0; }

extern "C" int song_mlp_param_count(const song_mlp *m, uint64_t *count) {
    if (!m || !count) return mfail(-1, "null");
    *count = (uint64_t)m->total;
    return 0;
}

extern "C" int song_mlp_get_params(const song_mlp *m, float *host) {
    if (!m || !host) return mfail(-1, "null");
    MHIP(hipMemcpy(host, m->P, m->total * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int song_mlp_set_params(song_mlp *m, const float *host) {
    if (!m || !host) return mfail(-1, "null");
    MHIP(hipMemcpy(m->P, host, m->total * 4, hipMemcpyHostToDevice));
    if (refresh_bf16(m, 0)) return -2;
    MHIP(hipDeviceSynchronize());
    return 0;
}

extern "C" int song_mlp_loss_grad(song_mlp *m, const uint16_t *d_x, const int *d_labels, const float *d_stop, int batch,
                                  float stop_weight, float *loss_out, float *grads_host, void *stream) {
    if (!m || !d_x || !d_labels || !d_stop) return mfail(-1, "null argument");
    if (batch < 1 || batch > m->max_batch) return mfail(-1, "batch %d outside [1, %d]", batch, m->max_batch);
    hipStream_t st = (hipStream_t)stream;
    const int W = m->width, Hd = m->hidden, NO = m->C + 1;
    float *W1 = m->P + m->off[0], *b1 = m->P + m->off[1], *b2 = m->P + m->off[3], *b3 = m->P + m->off[5];
    (void)W1;
    // forward
    GemmArgs g = gemm_args(batch, Hd, W, d_x, W, 1, m->Wb[0], 1, W);
    g.Cb = m->H1;
    g.ldcb = Hd;
    g.bias = b1;
    g.relu = 1;
    if (gemm(g, st)) return -2;
    g = gemm_args(batch, Hd, Hd, m->H1, Hd, 1, m->Wb[1], 1, Hd);
    g.Cb = m->H2;
    g.ldcb = Hd;
    g.bias = b2;
    g.relu = 1;
    if (gemm(g, st)) return -2;
    g = gemm_args(batch, NO, Hd, m->H2, Hd, 1, m->Wb[2], 1, Hd);
    g.C = m->Z;
    g.ldc = NO;
    g.bias = b3;
    if (gemm(g, st)) return -2;
    MHIP(hipMemsetAsync(m->loss, 0, 8, st));
    hipLaunchKernelGGL(loss_kernel, dim3(batch), dim3(64), 0, st, m->Z, NO, m->C, d_labels, d_stop, batch, stop_weight,
                       m->dZ, NO, m->loss);
    MHIP(hipGetLastError());
    // backward: dW3 = dZ^T H2, db3, dH2 = dZ W3 * (H2 > 0), ...
    float *G = m->G;
    g = gemm_args(NO, Hd, batch, m->dZ, 1, NO, m->H2, Hd, 1);
    g.C = G + m->off[4];
    g.ldc = Hd;
    if (gemm(g, st)) return -2;
    hipLaunchKernelGGL(colsum_kernel, dim3((NO + 255) / 256), dim3(256), 0, st, m->dZ, batch, NO, NO, G + m->off[5]);
    g = gemm_args(batch, Hd, NO, m->dZ, NO, 1, m->Wb[2], Hd, 1);
    g.Cb = m->dH2;
    g.ldcb = Hd;
    g.mask = m->H2;
    g.ldm = Hd;
    if (gemm(g, st)) return -2;
    g = gemm_args(Hd, Hd, batch, m->dH2, 1, Hd, m->H1, Hd, 1);
    g.C = G + m->off[2];
    g.ldc = Hd;
    if (gemm(g, st)) return -2;
    hipLaunchKernelGGL(colsum_kernel, dim3((Hd + 255) / 256), dim3(256), 0, st, m->dH2, batch, Hd, Hd, G + m->off[3]);
    g = gemm_args(batch, Hd, Hd, m->dH2, Hd, 1, m->Wb[1], Hd, 1);
    g.Cb = m->dH1;
    g.ldcb = Hd;
    g.mask = m->H1;
    g.ldm = Hd;
    if (gemm(g, st)) return -2;
    g = gemm_args(Hd, W, batch, m->dH1, 1, Hd, d_x, W, 1);
    g.C = G + m->off[0];
    g.ldc = W;
    if (gemm(g, st)) return -2;
    hipLaunchKernelGGL(colsum_kernel, dim3((Hd + 255) / 256), dim3(256), 0, st, m->dH1, batch, Hd, Hd, G + m->off[1]);
    MHIP(hipGetLastError());
    if (loss_out) {
        MHIP(hipMemcpyAsync(loss_out, m->loss, 8, hipMemcpyDeviceToHost, st));
        MHIP(hipStreamSynchronize(st));
    }
    if (grads_host) {
        MHIP(hipMemcpyAsync(grads_host, G, m->total * 4, hipMemcpyDeviceToHost, st));
        MHIP(hipStreamSynchronize(st));
    }
    return 0;
}

extern "C" int song_mlp_adam_step(song_mlp *m, float lr, void *stream) {
    if (!m) return mfail(-1, "null");
    hipStream_t st = (hipStream_t)stream;
    m->t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)m->t), c2 = 1.0f - std::pow(b2, (float)m->t);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((m->total + 255) / 256)), dim3(256), 0, st, m->P, m->G, m->Mo,
                       m->Vo, m->total, lr, b1, b2, c1, c2, 1e-8f);
    MHIP(hipGetLastError());
    return refresh_bf16(m, st);
}

extern "C" int song_mlp_infer(song_mlp *m, const uint16_t *d_x, int batch, int *d_top2, float *d_conf, float *d_stop,
                              float *d_logits, void *stream) {
    return song_mlp_infer_nbr(m, d_x, batch, nullptr, 0, d_top2, d_conf, d_stop, d_logits, stream);
}

extern "C" int song_mlp_infer_nbr(song_mlp *m, const uint16_t *d_x, int batch, const int *d_nbr, int n_nbr, int *d_top2,
                                  float *d_conf, float *d_stop, float *d_logits, void *stream) {
    if (d_nbr && (n_nbr < 1 || n_nbr > 64)) return mfail(-1, "n_nbr=%d not in [1,64]", n_nbr);
    if (!m || !d_x || !d_top2 || !d_conf || !d_stop) return mfail(-1, "null argument");
    if (batch < 1) return 0;
    InferArgs a{};
    a.X = d_x;
    a.batch = batch;
    a.width = m->width;
    a.hidden = m->hidden;
    a.C = m->C;
    a.W1 = m->Wb[0];
    a.W2 = m->Wb[1];
    a.W3 = m->Wb[2];
    a.W3i = m->W3i;
    a.npad = m->npad;
    a.b1 = m->P + m->off[1];
    a.b2 = m->P + m->off[3];
    a.b3 = m->P + m->off[5];
    a.top2 = d_top2;
    a.conf = d_conf;
    a.stop = d_stop;
    a.logits = d_logits;
    a.nbr = d_nbr;
    a.n_nbr = n_nbr;
    const size_t lds = (size_t)IQ * (m->width + 8) * 2 + 2 * (size_t)IQ * (m->hidden + 8) * 2 +
                       ((IW + 1) * IQ) * sizeof(Soft) + IQ * 4;
    if (lds > 65536) MHIP(hipFuncSetAttribute((const void *)infer_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(infer_kernel, dim3((batch + IQ - 1) / IQ), dim3(IT), lds, (hipStream_t)stream, a);
    MHIP(hipGetLastError());
    return 0;
}

__global__ __launch_bounds__(256) void plan_kernel(const int *top2, const float *conf, const float *stop, int batch,
                                                   const int *nbr, int n_nbr, const int *pivot_vertex, int npiv, float thr,
                                                   float scale, float add, int fcap, int pq_l, int pq_f, int *pq,
                                                   int *entry, int *caps, unsigned char *ok) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= batch) return;
    const int best = top2[2 * i];
    // no valid best pivot (non-finite logits): the fallback
    const bool good = conf[i] >= thr && best >= 0 && best < npiv;
    int *row = entry + (long long)i * n_nbr;
    for (int e = 0; e < n_nbr; ++e) row[e] = good ? pivot_vertex[nbr[(long long)best * n_nbr + e]] : -1;
    const float c = fminf(fmaxf(scale * (exp2f(stop[i]) - 1.0f) + add, 1.0f), 1.0e6f);
    caps[i] = good ? (int)c : (fcap > 0 ? fcap : 1000000);
    pq[i] = good ? pq_l : pq_f;
    ok[i] = good ? 1 : 0;
}

extern "C" int song_mlp_plan(const int *d_top2, const float *d_conf, const float *d_stop, int batch, const int *d_nbr,
                             int n_nbr, const int *d_pivot_vertex, int n_pivots, float conf_threshold,
                             float step_scale, float step_add, int fallback_cap, int pq_learned, int pq_fallback, int *d_pq,
                             int *d_entry, int *d_caps, unsigned char *d_ok, void *stream) {
    if (!d_top2 || !d_conf || !d_stop || !d_nbr || !d_pivot_vertex || !d_pq || !d_entry || !d_caps || !d_ok)
        return mfail(-1, "null argument");
    if (n_nbr < 1 || pq_learned < 1 || pq_fallback < 1 || fallback_cap < 0)
        return mfail(-1, "bad plan arguments n_nbr=%d pq=%d/%d cap=%d", n_nbr, pq_learned, pq_fallback, fallback_cap);
    if (batch < 1) return 0;
    hipLaunchKernelGGL(plan_kernel, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, (hipStream_t)stream, d_top2,
                       d_conf, d_stop, batch, d_nbr, n_nbr, d_pivot_vertex, n_pivots, conf_threshold, step_scale,
                       step_add,
                       fallback_cap, pq_learned, pq_fallback, d_pq, d_entry, d_caps, d_ok);
    MHIP(hipGetLastError());
    return 0;
}

extern "C" int song_mlp_features(const float *d_queries, int nq, int dim, const float *d_mean, const float *d_inv_std,
                                 const float *d_v0, const float *fs, uint16_t *d_x, int width, void *stream) {
    if (!d_queries || !d_mean || !d_inv_std || !d_v0 || !fs || !d_x) return mfail(-1, "null argument");
    if (width < dim + 8) return mfail(-1, "width %d < dim + 8", width);
    if (nq < 1) return 0;
    hipLaunchKernelGGL(features_kernel, dim3(nq), dim3(64), 0, (hipStream_t)stream, d_queries, dim, d_mean, d_inv_std,
                       d_v0, fs[0], fs[1], fs[2], fs[3], fs[4], fs[5], fs[6], fs[7], d_x, width);
    MHIP(hipGetLastError());
    return 0;
}
