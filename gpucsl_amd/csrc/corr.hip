// Correlation matrix on MFMA: C = corrcoef(X) for X (n x p, row-major).
//
// Replaces gpucsl/pc/helpers.py:58-59 (np.corrcoef on the host).  The Gram
// matrix of the centred data is computed with v_mfma_f64_16x16x4_f64 in FP64 so
// that C64 serves the FP64 recheck and C32 (correctly rounded from C64) feeds the
// FP32 CI tests.  Split-K partial Gram matrices are summed in a fixed order, so
// results do not depend on scheduling.
#include "common.h"

namespace gpucsl {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int CT = 64;   // output tile edge
constexpr int KT = 16;   // k-slab per LDS stage
constexpr int LDA = CT + 4;

// partial[s][a] = sum of X[k][a] over rows k in split s (sequential in k).
template <typename T>
__global__ __launch_bounds__(256) void colsum_kernel(const T *__restrict__ X, int64_t n, int p,
                                                     int64_t rows_per_split,
                                                     double *__restrict__ partial) {
    const int a = blockIdx.x * 256 + threadIdx.x;
    const int s = blockIdx.y;
    if (a >= p) return;
    const int64_t r0 = s * rows_per_split;
    int64_t r1 = r0 + rows_per_split;
    if (r1 > n) r1 = n;
    double acc = 0.0;
    for (int64_t k = r0; k < r1; ++k) acc += (double)X[k * p + a];
    partial[(size_t)s * p + a] = acc;
}

__global__ void mean_kernel(const double *__restrict__ partial, int splits, int64_t n, int p,
                            double *__restrict__ mean) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= p) return;
    double acc = 0.0;
    for (int s = 0; s < splits; ++s) acc += partial[(size_t)s * p + a];
    mean[a] = acc / (double)n;
}

// Gp[s] (p x p, only tiles with ta <= tb written) = Zs^T Zs over k-range s,
// with Z = X - mean.  Block: 256 threads = 4 waves in a 2 x 2 arrangement,
// each wave owns a 32 x 32 sub-tile (2 x 2 MFMA 16x16 blocks).  Raw X slabs of
// KT rows are double-buffered in LDS: the global loads of slab s+1 are in
// flight while the MFMAs of slab s run, and centring (x - mean, in FP64)
// happens when a fragment is read.
template <typename T>
__global__ __launch_bounds__(256) void gram_mfma_kernel(const T *__restrict__ X, int64_t n, int p,
                                                        const double *__restrict__ mean,
                                                        int64_t k_per_split,
                                                        double *__restrict__ Gp) {
    __shared__ T As[2][KT][LDA];
    __shared__ T Bs[2][KT][LDA];

    // upper-triangle tile id -> (ta, tb), ta <= tb, row-major over the triangle
    const int tiles = (p + CT - 1) / CT;
    int t = blockIdx.x, ta = 0;
    while (t >= tiles - ta) {
        t -= tiles - ta;
        ++ta;
    }
    const int tb = ta + t;
    const int a0 = ta * CT, b0 = tb * CT;
    const int split = blockIdx.y;
    const int64_t k0 = split * k_per_split;
    int64_t k1 = k0 + k_per_split;
    if (k1 > n) k1 = n;
    const int nslab = k1 > k0 ? (int)((k1 - k0 + KT - 1) / KT) : 0;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr = (wave >> 1) * 32;  // wave row offset in tile
    const int wc = (wave & 1) * 32;   // wave col offset in tile
    const int ci = lane & 15;

    f64x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

    double ma[2], mb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int a = a0 + wr + i * 16 + ci, b = b0 + wc + i * 16 + ci;
        ma[i] = a < p ? mean[a] : 0.0;
        mb[i] = b < p ? mean[b] : 0.0;
    }

    // loader: 256 threads cover KT rows x 64 cols as KT/4 passes of 4 rows
    const int lcol = tid & 63;
    const int lrow = tid >> 6;
    const int ga = a0 + lcol, gb = b0 + lcol;
    T ra[KT / 4], rb[KT / 4];

    auto load = [&](int s) {
        const int64_t kk = k0 + (int64_t)s * KT;
#pragma unroll
        for (int pass = 0; pass < KT / 4; ++pass) {
            const int64_t k = kk + pass * 4 + lrow;
            const bool kv = k < k1;
            ra[pass] = (kv && ga < p) ? X[k * p + ga] : T(0);
            rb[pass] = (kv && gb < p) ? X[k * p + gb] : T(0);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int pass = 0; pass < KT / 4; ++pass) {
            As[buf][pass * 4 + lrow][lcol] = ra[pass];
            Bs[buf][pass * 4 + lrow][lcol] = rb[pass];
        }
    };

    if (nslab > 0) {
        load(0);
        store(0);
    }
    __syncthreads();
    for (int s = 0; s < nslab; ++s) {
        const int cur = s & 1;
        if (s + 1 < nslab) load(s + 1);
        const int64_t kk = k0 + (int64_t)s * KT;
#pragma unroll
        for (int kq = 0; kq < KT / 4; ++kq) {
            const int kr = kq * 4 + (lane >> 4);
            const bool kv = kk + kr < k1;
            double af[2], bf[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                af[i] = kv ? (double)As[cur][kr][wr + i * 16 + ci] - ma[i] : 0.0;
#pragma unroll
            for (int j = 0; j < 2; ++j)
                bf[j] = kv ? (double)Bs[cur][kr][wc + j * 16 + ci] - mb[j] : 0.0;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        if (s + 1 < nslab) store(cur ^ 1);
        __syncthreads();
    }

    // D layout of v_mfma_f64_16x16x4_f64 (checked on gfx950): accumulator item v
    // of a lane holds column (lane & 15) and row (lane >> 4) + 4 * v.
    double *out = Gp + (size_t)split * p * p;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = b0 + wc + j * 16 + ci;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = a0 + wr + i * 16 + (lane >> 4) + 4 * v;
                if (row < p && col < p) out[(size_t)row * p + col] = acc[i][j][v];
            }
        }
}

// sd[a] = sqrt(G_aa / (n-1)) with G_aa summed over splits in order.
__global__ void sd_kernel(const double *__restrict__ Gp, int splits, int64_t n, int p,
                          double *__restrict__ sd) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= p) return;
    double g = 0.0;
    for (int s = 0; s < splits; ++s) g += Gp[(size_t)s * p * p + (size_t)a * p + a];
    sd[a] = sqrt(g / (double)(n - 1));
}

// C64[a][b] = clip((G_ab / (n-1)) / sd_lo / sd_hi, -1, 1) with lo = min(a, b):
// numpy corrcoef's operation order for its upper triangle, mirrored so that C is
// exactly symmetric; C32 = (float)C64.
__global__ __launch_bounds__(256) void normalize_kernel(const double *__restrict__ Gp, int splits,
                                                        int64_t n, int p,
                                                        const double *__restrict__ sd,
                                                        double *__restrict__ C64,
                                                        float *__restrict__ C32) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    const int a = blockIdx.y;
    if (b >= p) return;
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    double g = 0.0;
    for (int s = 0; s < splits; ++s) g += Gp[(size_t)s * p * p + (size_t)lo * p + hi];
    double c = g / (double)(n - 1);
    c = c / sd[lo];
    c = c / sd[hi];
    if (c > 1.0) c = 1.0;
    if (c < -1.0) c = -1.0;
    if (C64) C64[(size_t)a * p + b] = c;
    C32[(size_t)a * p + b] = (float)c;
}

__global__ void c64_to_c32_kernel(const double *__restrict__ C64, size_t count,
                                  float *__restrict__ C32) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) C32[i] = (float)C64[i];
}

struct CorrPlan {
    int row_splits;
    int64_t rows_per_split;
    int k_splits;
    int64_t k_per_split;
    int tiles_upper;
    size_t off_partial, off_mean, off_sd, off_gram, total;
};

static CorrPlan corr_plan(int64_t n, int p) {
    CorrPlan c{};
    c.row_splits = (int)((n + 127) / 128);
    if (c.row_splits > 64) c.row_splits = 64;
    if (c.row_splits < 1) c.row_splits = 1;
    c.rows_per_split = (n + c.row_splits - 1) / c.row_splits;
    const int tiles = (p + CT - 1) / CT;
    c.tiles_upper = tiles * (tiles + 1) / 2;
    int ks = (1024 + c.tiles_upper - 1) / c.tiles_upper;
    int64_t kmax = (n + 255) / 256;  // at least 256 rows per split
    if (ks > kmax) ks = (int)kmax;
    if (ks > 16) ks = 16;
    if (ks < 1) ks = 1;
    c.k_splits = ks;
    c.k_per_split = (n + ks - 1) / ks;
    c.k_per_split = (c.k_per_split + KT - 1) / KT * KT;
    size_t off = 0;
    c.off_partial = off;
    off = align_up(off + sizeof(double) * (size_t)c.row_splits * p, 256);
    c.off_mean = off;
    off = align_up(off + sizeof(double) * (size_t)p, 256);
    c.off_sd = off;
    off = align_up(off + sizeof(double) * (size_t)p, 256);
    c.off_gram = off;
    off = align_up(off + sizeof(double) * (size_t)ks * p * p, 256);
    c.total = off;
    return c;
}

size_t corr_workspace_bytes(int64_t n, int p) { return corr_plan(n, p).total; }

int corr_launch(const void *X, int dtype, int64_t n, int p, double *C64, float *C32, void *ws,
                size_t ws_bytes, hipStream_t s) {
    if (!X || !C32 || !ws || n < 2 || p < 1) return set_error(ERR_ARG, "corr: bad arguments");
    if (dtype != 0 && dtype != 1) return set_error(ERR_ARG, "corr: dtype must be 0 (f32) or 1 (f64)");
    const CorrPlan c = corr_plan(n, p);
    if (ws_bytes < c.total)
        return set_error(ERR_WORKSPACE, "corr: workspace %zu < required %zu", ws_bytes, c.total);
    char *base = (char *)ws;
    double *partial = (double *)(base + c.off_partial);
    double *mean = (double *)(base + c.off_mean);
    double *sd = (double *)(base + c.off_sd);
    double *gram = (double *)(base + c.off_gram);
    dim3 g1((p + 255) / 256, c.row_splits);
    if (dtype == 0)
        GPUCSL_LAUNCH("corr_colsum", colsum_kernel<float>, g1, dim3(256), 0, s, (const float *)X, n, p,
                           c.rows_per_split, partial);
    else
        GPUCSL_LAUNCH("corr_colsum", colsum_kernel<double>, g1, dim3(256), 0, s, (const double *)X, n, p,
                           c.rows_per_split, partial);
    GPUCSL_LAUNCH("corr_mean", mean_kernel, dim3((p + 255) / 256), dim3(256), 0, s, partial, c.row_splits, n,
                       p, mean);
    dim3 g2(c.tiles_upper, c.k_splits);
    if (dtype == 0)
        GPUCSL_LAUNCH("corr_gram_mfma_f64", gram_mfma_kernel<float>, g2, dim3(256), 0, s, (const float *)X, n, p, mean,
                           c.k_per_split, gram);
    else
        GPUCSL_LAUNCH("corr_gram_mfma_f64", gram_mfma_kernel<double>, g2, dim3(256), 0, s, (const double *)X, n, p,
                           mean, c.k_per_split, gram);
    GPUCSL_LAUNCH("corr_sd", sd_kernel, dim3((p + 255) / 256), dim3(256), 0, s, gram, c.k_splits, n, p, sd);
    GPUCSL_LAUNCH("corr_normalize", normalize_kernel, dim3((p + 255) / 256, p), dim3(256), 0, s, gram, c.k_splits,
                       n, p, sd, C64, C32);
    return check_hip(hipGetLastError(), "corr launch");
}

int c64_to_c32_launch(const double *C64, int p, float *C32, hipStream_t s) {
    if (!C64 || !C32 || p < 1) return set_error(ERR_ARG, "c64_to_c32: bad arguments");
    const size_t cnt = (size_t)p * p;
    size_t blocks = (cnt + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    GPUCSL_LAUNCH("corr_c64_to_c32", c64_to_c32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, C64, cnt, C32);
    return check_hip(hipGetLastError(), "c64_to_c32 launch");
}

}  // namespace gpucsl
