// Correlation matrix on MFMA: C = corrcoef(X) for X (n x p, row-major), for a
// table of jobs (blockIdx.z selects the job).
//
// Replaces gpucsl/pc/helpers.py:58-59 (np.corrcoef on the host).  The Gram
// matrix of the centred data is computed with v_mfma_f64_16x16x4_f64 in FP64 so
// that C64 serves the FP64 recheck and C32 (correctly rounded from C64) feeds the
// FP32 CI tests.  Split-K partial Gram matrices are summed in a fixed order, so
// results do not depend on scheduling.
#include "job.h"

namespace gpucsl {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int CT = 64;   // output tile edge
constexpr int KT = 16;   // k-slab per LDS stage
constexpr int LDA = CT + 4;

__global__ void write_job_kernel(Job *dst, Job src) { *dst = src; }

__host__ __device__ inline int upper_tiles(int p) {
    const int t = (p + CT - 1) / CT;
    return t * (t + 1) / 2;
}

// partial[s][a] = sum of X[k][a] over rows k of split s (sequential in k).
template <typename T>
__global__ __launch_bounds__(256) void colsum_kernel(const Job *__restrict__ jobs, int row_splits) {
    const Job &J = jobs[blockIdx.z];
    const int p = J.p;
    const int64_t n = J.n;
    const int a = blockIdx.x * 256 + threadIdx.x;
    const int s = blockIdx.y;
    if (a >= p) return;
    const T *X = (const T *)J.X;
    const int64_t per = (n + row_splits - 1) / row_splits;
    const int64_t r0 = s * per;
    const int64_t r1 = r0 + per < n ? r0 + per : n;
    double acc = 0.0;
    for (int64_t k = r0; k < r1; ++k) acc += (double)X[k * p + a];
    carve_corr(J.corr_ws, p).partial[(size_t)s * p + a] = acc;
}

__global__ void mean_kernel(const Job *__restrict__ jobs, int row_splits) {
    const Job &J = jobs[blockIdx.z];
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= J.p) return;
    CorrWs c = carve_corr(J.corr_ws, J.p);
    double acc = 0.0;
    for (int s = 0; s < row_splits; ++s) acc += c.partial[(size_t)s * J.p + a];
    c.mean[a] = acc / (double)J.n;
}

// gram[s] (p x p, only tiles with ta <= tb written) = Zs^T Zs over k-range s,
// with Z = X - mean.  Block: 256 threads = 4 waves in a 2 x 2 arrangement,
// each wave owns a 32 x 32 sub-tile (2 x 2 MFMA 16x16 blocks).  Raw X slabs of
// KT rows are double-buffered in LDS: the global loads of slab s+1 are in
// flight while the MFMAs of slab s run, and centring (x - mean, in FP64)
// happens when a fragment is read.
template <typename T>
__global__ __launch_bounds__(256) void gram_mfma_kernel(const Job *__restrict__ jobs) {
    __shared__ T As[2][KT][LDA];
    __shared__ T Bs[2][KT][LDA];

    const Job &J = jobs[blockIdx.z];
    const int p = J.p;
    const int64_t n = J.n;
    if ((int)blockIdx.x >= upper_tiles(p) || (int)blockIdx.y >= J.k_splits) return;
    const T *__restrict__ X = (const T *)J.X;
    CorrWs cw = carve_corr(J.corr_ws, p);
    const double *__restrict__ mean = cw.mean;

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
    int64_t k_per_split = (n + J.k_splits - 1) / J.k_splits;
    k_per_split = (k_per_split + KT - 1) / KT * KT;
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
    double *out = cw.gram + (size_t)split * p * p;
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
__global__ void sd_kernel(const Job *__restrict__ jobs) {
    const Job &J = jobs[blockIdx.z];
    const int p = J.p;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= p) return;
    CorrWs c = carve_corr(J.corr_ws, p);
    double g = 0.0;
    for (int s = 0; s < J.k_splits; ++s) g += c.gram[(size_t)s * p * p + (size_t)a * p + a];
    c.sd[a] = sqrt(g / (double)(J.n - 1));
}

// C64[a][b] = clip((G_ab / (n-1)) / sd_lo / sd_hi, -1, 1) with lo = min(a, b):
// numpy corrcoef's operation order for its upper triangle, mirrored so that C is
// exactly symmetric; C32 = (float)C64.
__global__ __launch_bounds__(256) void normalize_kernel(const Job *__restrict__ jobs) {
    const Job &J = jobs[blockIdx.z];
    const int p = J.p;
    const int b = blockIdx.x * 256 + threadIdx.x;
    const int a = blockIdx.y;
    if (b >= p || a >= p) return;
    CorrWs cw = carve_corr(J.corr_ws, p);
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    double g = 0.0;
    for (int s = 0; s < J.k_splits; ++s) g += cw.gram[(size_t)s * p * p + (size_t)lo * p + hi];
    double c = g / (double)(J.n - 1);
    c = c / cw.sd[lo];
    c = c / cw.sd[hi];
    if (c > 1.0) c = 1.0;
    if (c < -1.0) c = -1.0;
    if (J.C64) J.C64[(size_t)a * p + b] = c;
    J.C32[(size_t)a * p + b] = (float)c;
}

__global__ void c64_to_c32_kernel(const Job *__restrict__ jobs) {
    const Job &J = jobs[blockIdx.z];
    const size_t count = (size_t)J.p * J.p;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) J.C32[i] = (float)J.C64[i];
}

// ----------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------
static int corr_row_splits(int64_t max_n) {
    int rs = (int)((max_n + 127) / 128);
    if (rs > CORR_ROW_SPLITS_MAX) rs = CORR_ROW_SPLITS_MAX;
    return rs < 1 ? 1 : rs;
}

// split-K factor that gives one dataset >= 1024 workgroups
int corr_k_splits(int64_t n, int p) {
    const int tu = upper_tiles(p);
    int ks = (1024 + tu - 1) / tu;
    const int64_t kmax = (n + 255) / 256;
    if (ks > kmax) ks = (int)kmax;
    if (ks > 16) ks = 16;
    return ks < 1 ? 1 : ks;
}

int corr_batch_launch(const Job *jobs_host, const Job *jobs_dev, int nj, hipStream_t s) {
    if (nj < 1) return OK;
    int maxp = 0, maxks = 1, maxtiles = 1, dtype = jobs_host[0].x_dtype;
    int64_t maxn = 0;
    for (int b = 0; b < nj; ++b) {
        const Job &J = jobs_host[b];
        if (!J.X || !J.C32 || !J.corr_ws || J.n < 2 || J.p < 1 || J.k_splits < 1)
            return set_error(ERR_ARG, "corr: job %d has bad arguments", b);
        if (J.x_dtype != dtype) return set_error(ERR_ARG, "corr: jobs of one launch must share x_dtype");
        maxp = J.p > maxp ? J.p : maxp;
        maxn = J.n > maxn ? J.n : maxn;
        maxks = J.k_splits > maxks ? J.k_splits : maxks;
        maxtiles = upper_tiles(J.p) > maxtiles ? upper_tiles(J.p) : maxtiles;
    }
    if (dtype != 0 && dtype != 1) return set_error(ERR_ARG, "corr: dtype must be 0 (f32) or 1 (f64)");
    const int rs = corr_row_splits(maxn);
    const int cb = (maxp + 255) / 256;
    if (dtype == 0)
        GPUCSL_LAUNCH("corr_colsum", colsum_kernel<float>, dim3(cb, rs, nj), dim3(256), 0, s, jobs_dev, rs);
    else
        GPUCSL_LAUNCH("corr_colsum", colsum_kernel<double>, dim3(cb, rs, nj), dim3(256), 0, s, jobs_dev, rs);
    GPUCSL_LAUNCH("corr_mean", mean_kernel, dim3(cb, 1, nj), dim3(256), 0, s, jobs_dev, rs);
    if (dtype == 0)
        GPUCSL_LAUNCH("corr_gram_mfma_f64", gram_mfma_kernel<float>, dim3(maxtiles, maxks, nj), dim3(256), 0, s,
                      jobs_dev);
    else
        GPUCSL_LAUNCH("corr_gram_mfma_f64", gram_mfma_kernel<double>, dim3(maxtiles, maxks, nj), dim3(256), 0, s,
                      jobs_dev);
    GPUCSL_LAUNCH("corr_sd", sd_kernel, dim3(cb, 1, nj), dim3(256), 0, s, jobs_dev);
    GPUCSL_LAUNCH("corr_normalize", normalize_kernel, dim3(cb, maxp, nj), dim3(256), 0, s, jobs_dev);
    return check_hip(hipGetLastError(), "corr launch");
}

int c64_to_c32_batch_launch(const Job *jobs_host, const Job *jobs_dev, int nj, hipStream_t s) {
    size_t maxpp = 1;
    for (int b = 0; b < nj; ++b) {
        if (!jobs_host[b].C64 || !jobs_host[b].C32) return set_error(ERR_ARG, "c64_to_c32: null matrix");
        const size_t pp = (size_t)jobs_host[b].p * jobs_host[b].p;
        maxpp = pp > maxpp ? pp : maxpp;
    }
    size_t blocks = (maxpp + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    GPUCSL_LAUNCH("corr_c64_to_c32", c64_to_c32_kernel, dim3((unsigned)blocks, 1, nj), dim3(256), 0, s, jobs_dev);
    return check_hip(hipGetLastError(), "c64_to_c32 launch");
}

// ---- single dataset: the workspace starts with one job descriptor ----------
size_t corr_workspace_bytes(int64_t n, int p) {
    return JOB_DESC_BYTES + corr_job_ws_bytes(p, corr_k_splits(n, p));
}

int corr_launch(const void *X, int dtype, int64_t n, int p, double *C64, float *C32, void *ws, size_t ws_bytes,
                hipStream_t s) {
    if (!X || !C32 || !ws || n < 2 || p < 1) return set_error(ERR_ARG, "corr: bad arguments");
    if (ws_bytes < corr_workspace_bytes(n, p))
        return set_error(ERR_WORKSPACE, "corr: workspace %zu < required %zu", ws_bytes, corr_workspace_bytes(n, p));
    Job j{};
    j.p = p;
    j.n = n;
    j.x_dtype = dtype;
    j.k_splits = corr_k_splits(n, p);
    j.X = X;
    j.C32 = C32;
    j.C64 = C64;
    j.corr_ws = (char *)ws + JOB_DESC_BYTES;
    Job *dj = (Job *)ws;
    hipLaunchKernelGGL(write_job_kernel, dim3(1), dim3(1), 0, s, dj, j);
    return corr_batch_launch(&j, dj, 1, s);
}

__global__ void c64_to_c32_flat_kernel(const double *__restrict__ C64, size_t count, float *__restrict__ C32) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) C32[i] = (float)C64[i];
}

int c64_to_c32_launch(const double *C64, int p, float *C32, hipStream_t s) {
    if (!C64 || !C32 || p < 1) return set_error(ERR_ARG, "c64_to_c32: bad arguments");
    const size_t cnt = (size_t)p * p;
    size_t blocks = (cnt + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    GPUCSL_LAUNCH("corr_c64_to_c32", c64_to_c32_flat_kernel, dim3((unsigned)blocks), dim3(256), 0, s, C64, cnt, C32);
    return check_hip(hipGetLastError(), "c64_to_c32 launch");
}

}  // namespace gpucsl
