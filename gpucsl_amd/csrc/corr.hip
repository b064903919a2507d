// Correlation matrix on the matrix cores: C = corrcoef(X) for X (n x p,
// row-major), for a table of jobs (blockIdx.z selects the job).
//
// Replaces gpucsl/pc/helpers.py:58-59 (np.corrcoef on the host).  C64 serves
// the FP64 recheck and C32 (correctly rounded from C64) feeds the FP32 CI
// tests, so the Gram matrix of the centred data is formed to FP64 accuracy:
//  - float32 data: each centred value, scaled by a power of two per column, is
//    rounded to a 42-bit integer and split into six signed base-128 digits;
//    the Gram matrix is the sum of the 26 digit-plane products whose weight is
//    at least 128^-6 of the leading one, on v_mfma_i32_16x16x64_i8 with exact
//    int32 accumulation (the splitting idea of the Ozaki scheme, Ozaki, Ogita,
//    Oishi and Rump, Numer. Algorithms 59, 2012, on integer units as in
//    Ootomo, Ozaki and Yokota, "DGEMM on integer matrix multiplication unit",
//    IJHPCA 2024; here with one fixed-point scale per column);
//  - float64 data: v_mfma_f64_16x16x4_f64 on the centred values.
// Split-K partial Gram matrices are summed in a fixed order, so results do not
// depend on scheduling.
#include "job.h"

namespace gpucsl {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int CT = 64;   // output tile edge
constexpr int KT = 16;   // k-slab per LDS stage

__global__ void write_job_kernel(Job *dst, Job src) { *dst = src; }

__host__ __device__ inline int upper_tiles(int p, int ct = CT) {
    const int t = (p + ct - 1) / ct;
    return t * (t + 1) / 2;
}

// Datasets with p <= GRAM_SMALL_P take their Gram matrix in 32 x 32 tiles of
// one wave each (at p = 200, 70 % of the MFMA work of the upper tiles is
// needed, against 49 % with 64 x 64 tiles).  Each entry accumulates the same
// k-slabs in the same order either way, so the bits do not depend on the tile.
constexpr int GRAM_SMALL_P = 256;
__host__ __device__ inline int gram_tile(int p) { return p <= GRAM_SMALL_P ? 32 : CT; }
// Float32 datasets above GRAM_SMALL_P take the integer Gram path; the others
// (and all float64 datasets) the FP64 one.
__host__ __device__ inline bool gram_int8(int dtype, int p) { return dtype == 0 && p > GRAM_SMALL_P; }

// Row splits of the column sums: a function of the job's own n, so a job's
// means are the same bits alone and in any batch.
__host__ __device__ inline int corr_row_splits(int64_t n) {
    int64_t rs = (n + 511) / 512;
    if (rs > CORR_ROW_SPLITS_MAX) rs = CORR_ROW_SPLITS_MAX;
    return rs < 1 ? 1 : (int)rs;
}

// partial[s][a] = sum of X[k][a] over the rows k of split s (cmin, cmax: the
// extremes over the same rows): 64 columns x 4 row groups per block (group g
// takes rows g, g+4, ... of the split), the four group sums added in order.
template <typename T>
__device__ __forceinline__ void colsum_body(const Job &J, int bx, int s) {
    const int p = J.p;
    const int64_t n = J.n;
    const int rs = corr_row_splits(n);
    if (s >= rs || bx * 64 >= p) return;
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int a = bx * 64 + c;
    const T *X = (const T *)J.X;
    const int64_t per = (n + rs - 1) / rs;
    const int64_t r0 = s * per;
    const int64_t r1 = r0 + per < n ? r0 + per : n;
    double acc = 0.0, mn = INFINITY, mx = -INFINITY;
    if (a < p) {
        // rows g, g+4, g+8, ... added in that order; eight loads in flight
        int64_t k = r0 + g;
        for (; k + 28 < r1; k += 32) {
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) x[u] = (double)X[(k + 4 * u) * p + a];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                acc += x[u];
                mn = fmin(mn, x[u]);
                mx = fmax(mx, x[u]);
            }
        }
        for (; k < r1; k += 4) {
            const double x = (double)X[k * p + a];
            acc += x;
            mn = fmin(mn, x);
            mx = fmax(mx, x);
        }
    }
    __shared__ double part[3][4][64];
    part[0][g][c] = acc;
    part[1][g][c] = mn;
    part[2][g][c] = mx;
    __syncthreads();
    if (g == 0 && a < p) {
        const CorrWs cw = carve_corr(J);
        const size_t at = (size_t)s * p + a;
        cw.partial[at] = ((part[0][0][c] + part[0][1][c]) + part[0][2][c]) + part[0][3][c];
        cw.cmin[at] = fmin(fmin(part[1][0][c], part[1][1][c]), fmin(part[1][2][c], part[1][3][c]));
        cw.cmax[at] = fmax(fmax(part[2][0][c], part[2][1][c]), fmax(part[2][2][c], part[2][3][c]));
    }
}

// mean of column col: the row-split partial sums added in split order, / n
__device__ __forceinline__ double column_mean(const CorrWs &cw, int64_t n, int p, int col) {
    const int rs = corr_row_splits(n);
    const double *pc = cw.partial + col;
    double m = 0.0;
    int q = 0;
    for (; q + 4 <= rs; q += 4) {
        const double x0 = pc[(size_t)q * p], x1 = pc[(size_t)(q + 1) * p];
        const double x2 = pc[(size_t)(q + 2) * p], x3 = pc[(size_t)(q + 3) * p];
        m += x0;
        m += x1;
        m += x2;
        m += x3;
    }
    for (; q < rs; ++q) m += pc[(size_t)q * p];
    return m / (double)n;
}

// A batch: job range R.job[k] has (p + 63) / 64 column blocks x its row splits.
template <typename T>
__global__ __launch_bounds__(256) void colsum_kernel(const Job *__restrict__ jobs, BlockRanges R) {
    int local;
    const Job &J = jobs[range_job(R, blockIdx.x, local)];
    const int cb = (J.p + 63) / 64;
    colsum_body<T>(J, local % cb, local / cb);
}

// One job passed by value; block (0, 0) also stores its descriptor at dst for
// the launches that follow, so a single dataset needs no separate write.
template <typename T>
__global__ __launch_bounds__(256) void colsum_one_kernel(Job one, Job *dst) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *dst = one;
    colsum_body<T>(one, blockIdx.x, blockIdx.y);
}

// gram[s] (p x p, only tiles with ta <= tb written) = Zs^T Zs over k-range s,
// with Z = X - mean.  Block: 256 threads = 4 waves in a 2 x 2 arrangement,
// each wave owns a 32 x 32 sub-tile (2 x 2 MFMA 16x16 blocks).  Raw X slabs of
// KT rows are double-buffered in LDS: the global loads of slab s+1 are in
// flight while the MFMAs of slab s run, and centring (x - mean, in FP64)
// happens when a fragment is read.
template <typename T, int CTT>
__global__ __launch_bounds__(CTT == CT ? 256 : 64) void gram_mfma_kernel(const Job *__restrict__ jobs,
                                                                          BlockRanges R) {
    constexpr int NT = CTT == CT ? 256 : 64;  // 4 waves of 32 x 32, or one
    constexpr int RPP = NT / CTT;             // slab rows per loader pass
    __shared__ T As[2][KT][CTT + 4];
    __shared__ T Bs[2][KT][CTT + 4];

    int local;
    const Job &J = jobs[range_job(R, blockIdx.x, local)];
    const int p = J.p;
    const int64_t n = J.n;
    const int tile_id = local % upper_tiles(p, CTT), split = local / upper_tiles(p, CTT);
    if (gram_int8(J.x_dtype, p) || gram_tile(p) != CTT || split >= J.k_splits) return;
    const T *__restrict__ X = (const T *)J.X;
    CorrWs cw = carve_corr(J);

    // upper-triangle tile id -> (ta, tb), ta <= tb, row-major over the triangle
    const int tiles = (p + CTT - 1) / CTT;
    int t = tile_id, ta = 0;
    while (t >= tiles - ta) {
        t -= tiles - ta;
        ++ta;
    }
    const int tb = ta + t;
    const int a0 = ta * CTT, b0 = tb * CTT;
    int64_t k_per_split = (n + J.k_splits - 1) / J.k_splits;
    k_per_split = (k_per_split + KT - 1) / KT * KT;
    const int64_t k0 = split * k_per_split;
    int64_t k1 = k0 + k_per_split;
    if (k1 > n) k1 = n;
    const int nslab = k1 > k0 ? (int)((k1 - k0 + KT - 1) / KT) : 0;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr = CTT == CT ? (wave >> 1) * 32 : 0;  // wave row offset in tile
    const int wc = CTT == CT ? (wave & 1) * 32 : 0;   // wave col offset in tile
    const int ci = lane & 15;

    f64x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

    // column means of the two tile edges: mean[a] = (sum of the row-split
    // partial sums, in split order) / n
    __shared__ double s_mean[2][CTT];
    if (tid < 2 * CTT) {
        const int col = (tid < CTT ? a0 : b0) + (tid & (CTT - 1));
        s_mean[tid / CTT][tid & (CTT - 1)] = col < p ? column_mean(cw, n, p, col) : 0.0;
    }
    __syncthreads();
    double ma[2], mb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        ma[i] = s_mean[0][wr + i * 16 + ci];
        mb[i] = s_mean[1][wc + i * 16 + ci];
    }

    // loader: NT threads cover KT rows x CTT cols as KT/RPP passes of RPP rows
    const int lcol = tid & (CTT - 1);
    const int lrow = tid / CTT;
    const int ga = a0 + lcol, gb = b0 + lcol;
    T ra[KT / RPP], rb[KT / RPP];

    auto load = [&](int s) {
        const int64_t kk = k0 + (int64_t)s * KT;
#pragma unroll
        for (int pass = 0; pass < KT / RPP; ++pass) {
            const int64_t k = kk + pass * RPP + lrow;
            const bool kv = k < k1;
            ra[pass] = (kv && ga < p) ? X[k * p + ga] : T(0);
            rb[pass] = (kv && gb < p) ? X[k * p + gb] : T(0);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int pass = 0; pass < KT / RPP; ++pass) {
            As[buf][pass * RPP + lrow][lcol] = ra[pass];
            Bs[buf][pass * RPP + lrow][lcol] = rb[pass];
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
        // the next quarter-slab's values are read from LDS while the current
        // quarter's products run (the scheduling barrier keeps them there)
        T xa[2], xb[2], na[2], nb[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) xa[i] = As[cur][lane >> 4][wr + i * 16 + ci];
#pragma unroll
        for (int j = 0; j < 2; ++j) xb[j] = Bs[cur][lane >> 4][wc + j * 16 + ci];
#pragma unroll
        for (int kq = 0; kq < KT / 4; ++kq) {
            const int kr = kq * 4 + (lane >> 4);
            const bool kv = kk + kr < k1;
            if (kq + 1 < KT / 4) {
#pragma unroll
                for (int i = 0; i < 2; ++i) na[i] = As[cur][kr + 4][wr + i * 16 + ci];
#pragma unroll
                for (int j = 0; j < 2; ++j) nb[j] = Bs[cur][kr + 4][wc + j * 16 + ci];
            }
            __builtin_amdgcn_sched_barrier(0);
            double af[2], bf[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) af[i] = kv ? (double)xa[i] - ma[i] : 0.0;
#pragma unroll
            for (int j = 0; j < 2; ++j) bf[j] = kv ? (double)xb[j] - mb[j] : 0.0;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
            if (kq + 1 < KT / 4) {
#pragma unroll
                for (int i = 0; i < 2; ++i) xa[i] = na[i];
#pragma unroll
                for (int j = 0; j < 2; ++j) xb[j] = nb[j];
            }
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

// ---- integer Gram path (float32 data) --------------------------------------
// Digits of the centred column a: z = fl(x - mean_a) is scaled by 2^(41 - e_a),
// where 2^e_a exceeds max |z| over the column (from the column extremes:
// rounding is monotone, so |fl(x - mean)| <= max(fl(max - mean), fl(mean - min))),
// rounded to an integer Z with |Z| <= 2^41, and written as
// Z = sum_a D_a 128^(5-a) with D_5..D_1 in [-64, 63] and |D_0| <= 64.  The Gram matrix of the Z is
// then sum over digit pairs (a, b) of D_a^T D_b 128^(10-a-b); the pairs with
// a + b <= 6 are kept.  A dropped pair weighs at most 128^-7 of the leading
// one, and the only dropped pairs that do not average out over the rows, the
// squares D_a^T D_a of a diagonal entry, at most 128^-8; keeping a + b <= 5
// instead leaves the square of D_3 out of G_aa, a bias of about 2^-42 n that
// shows as 1e-12 in C.  With a + b <= 6, C agrees with numpy's FP64
// corrcoef to a few 1e-15, as does the FP64 Gram matrix.  The per-column scale
// 2^(41 - e_a) enters G_ab as a factor 2^(41-e_a) 2^(41-e_b) and cancels
// exactly in G_ab / sqrt(G_aa G_bb), so the planes carry no scale.  A column
// whose mean or extremes are not finite (NaN or Inf in the data) gets zero
// digits and hence NaN correlations, as with the FP64 arithmetic.
//
// Block: 64 columns x I8_SLICE_STEPS steps of 64 rows; 256 threads.  Thread
// (column c = tid % 64, quarter q = tid / 64) converts rows 16 q .. 16 q + 15 of
// its column in each step; they are bytes 64 c + 16 q .. + 15 of the step's
// tile of every digit.  The column centre and scale come from the row
// splits' sums and extremes, the splits taken four at a time by the four waves
// and their partial sums added in wave order.
constexpr int I8_SLICE_STEPS = 2;
__global__ __launch_bounds__(256) void slice_i8_kernel(const Job *__restrict__ jobs, BlockRanges R) {
    int local;
    const Job &J = jobs[range_job(R, blockIdx.x, local)];
    if (!gram_int8(J.x_dtype, J.p)) return;
    const int p = J.p;
    const int64_t n = J.n;
    const int64_t steps = i8_rows(n) / 64;
    const int CB = i8_cols(p) / 64;
    const int cb = local % CB;
    const int64_t st0 = (int64_t)(local / CB) * I8_SLICE_STEPS;
    if (st0 >= steps || cb >= CB) return;
    const CorrWs cw = carve_corr(J);
    const int tid = threadIdx.x;
    const int c = tid & 63, q = tid >> 6;
    const int col = cb * 64 + c;
    __shared__ double s_part[3][4][64];
    __shared__ double s_mu[64], s_sc[64];
    __shared__ int4 s_tile[I8_PLANES][256 + 64];
    {
        // splits q, q + 4, q + 8, ... in that order, eight loads of each kind in flight
        const int rs = corr_row_splits(n);
        double sm = 0.0, mn = INFINITY, mx = -INFINITY;
        if (col < p)
            for (int r0 = q; r0 < rs; r0 += 32) {
                double vs[8], vn[8], vx[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int r = r0 + 4 * u;
                    const size_t at = (size_t)(r < rs ? r : q) * p + col;
                    vs[u] = cw.partial[at];
                    vn[u] = cw.cmin[at];
                    vx[u] = cw.cmax[at];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (r0 + 4 * u < rs) {
                        sm += vs[u];
                        mn = fmin(mn, vn[u]);
                        mx = fmax(mx, vx[u]);
                    }
            }
        s_part[0][q][c] = sm;
        s_part[1][q][c] = mn;
        s_part[2][q][c] = mx;
    }
    __syncthreads();
    if (tid < 64) {
        const double mu = (((s_part[0][0][c] + s_part[0][1][c]) + s_part[0][2][c]) + s_part[0][3][c]) / (double)n;
        const double mn = fmin(fmin(s_part[1][0][c], s_part[1][1][c]), fmin(s_part[1][2][c], s_part[1][3][c]));
        const double mx = fmax(fmax(s_part[2][0][c], s_part[2][1][c]), fmax(s_part[2][2][c], s_part[2][3][c]));
        const double bound = fmax(mx - mu, mu - mn);
        double sc = 0.0;
        if (col < p && isfinite(mu) && isfinite(bound) && bound > 0.0) {
            int e;
            frexp(bound, &e);
            sc = ldexp(1.0, 41 - e);
        }
        s_mu[c] = mu;
        s_sc[c] = sc;
    }
    __syncthreads();
    const float *X = (const float *)J.X;
    const double mu = s_mu[c], sc = s_sc[c];
    const bool live = sc != 0.0;
    const int64_t st1 = st0 + I8_SLICE_STEPS < steps ? st0 + I8_SLICE_STEPS : steps;
    // rows of the next step are loaded while the current step is converted
    float xn[16];
    auto load = [&](int64_t st) {
        const int64_t k0 = st * 64 + 16 * q;
#pragma unroll
        for (int j = 0; j < 16; ++j) xn[j] = live && k0 + j < n ? X[(k0 + j) * p + col] : 0.0f;
    };
    load(st0);
    for (int64_t st = st0; st < st1; ++st) {
        const int64_t k0 = st * 64 + 16 * q;
        float x[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] = xn[j];
        if (st + 1 < st1) load(st + 1);
        // Z' = Z + B with B = 64 (128^5 - 1) / 127: the five lower 7-bit fields of
        // Z' minus 64 are the lower digits and the field above them is D_0, so
        // no carries are needed.  Z' = zh 2^21 + zl with 0 <= zl < 2^21, both
        // exact in FP64.  Four rows' fields go into one word per digit and the
        // "- 64" is applied to the four bytes at once.
        int w[I8_PLANES][4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            int zh[4], zl[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int j = 4 * g + e;
                const double z = live && k0 + j < n ? rint(((double)x[j] - mu) * sc) + 17315143744.0 : 17315143744.0;
                const double hi = floor(z * (1.0 / 2097152.0));
                zh[e] = (int)hi;
                zl[e] = (int)(z - hi * 2097152.0);
            }
            auto pack = [&](const int (&v)[4], int shift, int width) {
                const unsigned m = (1u << width) - 1u;
                return (int)((((unsigned)v[0] >> shift) & m) | ((((unsigned)v[1] >> shift) & m) << 8) |
                             ((((unsigned)v[2] >> shift) & m) << 16) | ((((unsigned)v[3] >> shift) & m) << 24));
            };
            auto rebias = [](int v) {
                return (int)((((unsigned)v | 0x80808080u) - 0x40404040u) ^ 0x80808080u);
            };
            w[0][g] = pack(zh, 14, 8);
            w[1][g] = rebias(pack(zh, 7, 7));
            w[2][g] = rebias(pack(zh, 0, 7));
            w[3][g] = rebias(pack(zl, 14, 7));
            w[4][g] = rebias(pack(zl, 7, 7));
            w[5][g] = rebias(pack(zl, 0, 7));
        }
        // through LDS, so that each wave stores 1 KB of a tile contiguously: chunk
        // (c, q) sits at 16-byte slot 4 c + q + c / 4 (the pad keeps the 16 lanes
        // of a wave quarter on different banks when writing)
#pragma unroll
        for (int a = 0; a < I8_PLANES; ++a) {
            int4 v;
            v.x = w[a][0];
            v.y = w[a][1];
            v.z = w[a][2];
            v.w = w[a][3];
            s_tile[a][4 * c + q + (c >> 2)] = v;
        }
        __syncthreads();
#pragma unroll
        for (int a = 0; a < I8_PLANES; ++a)
            *(int4 *)(cw.planes + i8_tile(a, cb, st, CB, steps) + 16 * tid) = s_tile[a][tid + (tid >> 4)];
        __syncthreads();
    }
}

typedef int i32x4 __attribute__((ext_vector_type(4)));
constexpr int I8_MAXU = 6;        // digit pairs (a, b) with a + b <= I8_MAXU
// LDS panel layout: 64 bytes per column and digit, the four 16-byte chunks of
// column c in the order q ^ ((c / 2) % 4).  A fragment read (lane l: column
// l % 16 of a 16-column group, chunk l / 16) then measures free of LDS bank
// conflicts on gfx950 (SQ_LDS_BANK_CONFLICT of a probe kernel), where the
// order q ^ ((c / 4) % 4) and the plain order are two-way conflicted.
constexpr int I8_COLB = 64;

// Blocks of the integer Gram launch for one dataset: the upper 64 x 64 tiles
// (ta <= tb) taken two at a time along each tile row, (ta, ta) with
// (ta, ta + 1), then (ta, ta + 2) with (ta, ta + 3), ...; with t tiles a side
// that is sum over k = 1..t of ceil(k / 2) = (t + 1)^2 / 4 blocks.
__host__ __device__ inline int i8_pair_blocks(int p) {
    const int t = (p + CT - 1) / CT;
    return (t + 1) * (t + 1) / 4;
}

// Split-K partial Gram tiles (ta, tb0) and (ta, tb0 + 1) from the digit planes:
// the 26 digit pairs (a, b) with a + b <= 6, D_a^T D_b accumulated exactly in
// int32 per weight u = a + b (at most 6 pairs x 64 rows x 64^2 per step, so a
// split of at most I8_MAX_STEPS steps stays below 2^31), then combined as
// sum_u T_u 128^-u in FP64 from the smallest weight up.  Block: 4 waves, wave
// w computes rows 32 (w / 2) .. + 31 of tile (ta, tb0 + w % 2), i.e. 2 x 4 MFMA
// blocks of 16 x 16 per digit pair; a wave whose tile lies past the last
// column tile only loads.  A 64-row step of the panels (6 digits x the A tile
// and the two B tiles, 4 KB each) is staged in LDS by direct global-to-LDS
// loads, double-buffered: the next step's panels stream into one buffer while
// the current step is multiplied out of the other, with one barrier per step.
constexpr int I8_PANEL_SLOTS = 3;  // A tile, B tiles tb0 and tb0 + 1
typedef int8_t I8Panels[I8_PLANES][I8_PANEL_SLOTS][64][I8_COLB];
typedef __attribute__((address_space(3))) void *LdsPtr;

// The integer Gram launch covers only the blocks that have work: the jobs of
// a launch that take the integer path, up to I8_LAUNCH_JOBS of them, job[s]
// owning blocks first[s] .. first[s + 1] - 1 (pair blocks x splits, split
// major).  Passed by value, so a captured launch keeps it.
constexpr int I8_LAUNCH_JOBS = 64;
struct I8Launch {
    int count;
    int job[I8_LAUNCH_JOBS];
    int first[I8_LAUNCH_JOBS + 1];
};

// Workgroups go to the 8 XCDs in turn (workgroup w to XCD w % 8), and each
// XCD has its own L2.  A launch of one job is renumbered so that XCD k runs a
// contiguous range of (split, pair) blocks: blocks of one split that share
// panel tiles then meet in the same L2 (a bijection on the launch).  A launch
// of several jobs keeps the dispatch order, which spreads each job's blocks
// over all XCDs; its jobs come longest blocks first.
__device__ __forceinline__ void gram_block_order(const I8Launch &L, int &job, int &local) {
    const unsigned N = gridDim.x, hw = blockIdx.x;
    const unsigned q = N / 8, r = N % 8, k = hw % 8;
    const int id = L.count == 1 ? (int)(k * q + (k < r ? k : r) + hw / 8) : (int)hw;
    int lo = 0, hi = L.count - 1;  // the slot s with first[s] <= id < first[s + 1]
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (L.first[mid] <= id)
            lo = mid;
        else
            hi = mid - 1;
    }
    job = L.job[lo];
    local = id - L.first[lo];
}

__global__ __launch_bounds__(256, 1) void gram_i8_kernel(const Job *__restrict__ jobs, I8Launch launch) {
    // two separate arrays (not one indexed by step parity), so the compiler
    // sees that reads of one buffer do not wait for the loads into the other
    __shared__ __attribute__((aligned(16))) I8Panels lds0;
    __shared__ __attribute__((aligned(16))) I8Panels lds1;
    int job_id, local;
    gram_block_order(launch, job_id, local);
    const Job &J = jobs[job_id];
    const int p = J.p;
    const int64_t n = J.n;
    const int split = local / i8_pair_blocks(p);
    const CorrWs cw = carve_corr(J);
    const int tiles = (p + CT - 1) / CT;
    int x = local % i8_pair_blocks(p), ta = 0;
    for (;;) {
        const int nb = (tiles - ta + 1) >> 1;
        if (x < nb) break;
        x -= nb;
        ++ta;
    }
    const int tb0 = ta + 2 * x;
    const bool two = tb0 + 1 < tiles;  // tile (ta, tb0 + 1) exists
    const int64_t npad = i8_rows(n);
    const int ppad = i8_cols(p);
    const int64_t steps = npad / 64;
    const int64_t per = (steps + J.k_splits - 1) / J.k_splits;
    const int64_t s0 = split * per;
    const int64_t s1 = s0 + per < steps ? s0 + per : steps;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int half = wave & 1, wr = (wave >> 1) * 32;
    const int ci = lane & 15, kq = lane >> 4;
    const bool idle = half == 1 && !two;

    // LDS panel tile: 64 bytes per column, the four 16-byte chunks of column c
    // in the order q ^ ((c / 2) % 4).  A wave's direct load writes 1 KB of a
    // tile, lane l at byte 16 l, so thread tid fetches the chunk that belongs at
    // byte 16 tid: column tid / 4, chunk (tid % 4) ^ ((tid / 8) % 4).  Slot 1
    // holds B tile tb0 and slot 2 B tile tb0 + 1, where a tile that does not
    // exist is replaced by the A tile, so that the loads take no branches.
    // Tile offsets and LDS addresses are wave-uniform (M0 comes from scalars).
    const int CB = ppad / 64;
    const size_t plane_stride = (size_t)CB * steps * I8_TILE;
    const unsigned gsrc = 16u * (tid ^ ((tid >> 3) & 3));
    const int wave_u = __builtin_amdgcn_readfirstlane(wave);
    const size_t offA = i8_tile(0, ta, 0, CB, steps);
    const size_t offB0 = i8_tile(0, tb0, 0, CB, steps);
    const size_t offB1 = two ? offB0 + (size_t)steps * I8_TILE : offA;  // column tile tb0 + 1
    const int rchunk = kq ^ ((ci >> 1) & 3);                           // column 16 m + ci
    // direct load of panel tile t (digit t / 3, slot t % 3) of step st into buf
    auto issue = [&](I8Panels &buf, int64_t st, int t) {
        const int a = t / I8_PANEL_SLOTS, slot = t % I8_PANEL_SLOTS;
        const size_t off = (size_t)st * I8_TILE + a * plane_stride + (slot == 0 ? offA : slot == 1 ? offB0 : offB1);
        __builtin_amdgcn_global_load_lds(cw.planes + off + gsrc, (LdsPtr)&buf[a][slot][16 * wave_u][0], 16, 0, 0);
    };
    constexpr int I8_STEP_TILES = I8_PLANES * I8_PANEL_SLOTS;

    i32x4 acc[I8_MAXU + 1][2][4];
#pragma unroll
    for (int u = 0; u <= I8_MAXU; ++u)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[u][i][j] = i32x4{0, 0, 0, 0};

    // One step: the products of the panels in buf, and (with nxt) the direct
    // loads of step nst into the other buffer.  The fragments of every digit
    // are read once per step and held in registers.  The six digit pairs that
    // use neither digit 0 nor digit 1, (2,2), (2,3), (3,2), (3,3), (2,4) and
    // (4,2), are deferred to the start of the next step (pend), where they run
    // on the previous step's fragments of digits 2..4 while the new fragments
    // of digits 0 and 1 are read, so the matrix pipe does not wait for LDS
    // after the step's barrier.  The other pairs follow digit by digit, each
    // digit's fragments read while the previous pairs run.  The next step's
    // 18 tile loads go out two after each of the first nine of those pairs,
    // so that they have the rest of the step to land.
    i32x4 af[I8_PLANES][2], bf[I8_PLANES][4];
    auto pair = [&](int a, int b) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[a + b][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[a][i], bf[b][j], acc[a + b][i][j], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
    };
    auto deferred = [&]() {
        pair(2, 2);
        pair(2, 3);
        pair(3, 2);
        pair(3, 3);
        pair(2, 4);
        pair(4, 2);
    };
    auto step = [&](const I8Panels &buf, I8Panels *nxt, int64_t nst, bool pend) {
        if (idle) {
            if (nxt)
#pragma unroll
                for (int t = 0; t < I8_STEP_TILES; ++t) issue(*nxt, nst, t);
            return;
        }
        const int bslot = 1 + half;
        auto read = [&](int d) {
#pragma unroll
            for (int i = 0; i < 2; ++i) af[d][i] = *(const i32x4 *)&buf[d][0][wr + 16 * i + ci][16 * rchunk];
#pragma unroll
            for (int j = 0; j < 4; ++j) bf[d][j] = *(const i32x4 *)&buf[d][bslot][16 * j + ci][16 * rchunk];
        };
        int k = 0;  // this step's pairs done
        auto own = [&](int a, int b) {
            pair(a, b);
            if (nxt && 2 * k < I8_STEP_TILES) {
                issue(*nxt, nst, 2 * k);
                issue(*nxt, nst, 2 * k + 1);
            }
            ++k;
            __builtin_amdgcn_sched_barrier(0);
        };
        read(0);
        read(1);
        __builtin_amdgcn_sched_barrier(0);
        if (pend) deferred();
        own(0, 0);
        read(2);
        __builtin_amdgcn_sched_barrier(0);
        own(0, 1);
        own(1, 0);
        own(1, 1);
#pragma unroll
        for (int d = 2; d < I8_PLANES; ++d) {
            // reads of digit d + 1 go out before the pairs of digit d with
            // digits 0 and 1 (at most 12 reads in flight)
            if (d + 1 < I8_PLANES) read(d + 1);
            __builtin_amdgcn_sched_barrier(0);
            own(0, d);
            own(d, 0);
            own(1, d);
            own(d, 1);
        }
    };

    // the barrier after each step also waits for this wave's direct loads
    if (s0 < s1)
#pragma unroll
        for (int t = 0; t < I8_STEP_TILES; ++t) issue(lds0, s0, t);
    __syncthreads();
    for (int64_t st = s0; st < s1; st += 2) {
        step(lds0, st + 1 < s1 ? &lds1 : nullptr, st + 1, st > s0);
        __syncthreads();
        if (st + 1 >= s1) break;
        step(lds1, st + 2 < s1 ? &lds0 : nullptr, st + 2, true);
        __syncthreads();
    }
    if (!idle && s0 < s1) deferred();
    if (idle) return;

    // D layout of v_mfma_i32_16x16x64_i8 (checked on gfx950): accumulator item v
    // of a lane holds column (lane & 15) and row 4 * (lane >> 4) + v.
    double *out = cw.gram + (size_t)split * p * p;
    const int a0 = ta * CT + wr, b0 = (tb0 + half) * CT;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = b0 + j * 16 + ci;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = a0 + i * 16 + 4 * kq + v;
                double h = (double)acc[I8_MAXU][i][j][v];
#pragma unroll
                for (int u = I8_MAXU - 1; u >= 0; --u) h = h * (1.0 / 128.0) + (double)acc[u][i][j][v];
                if (row < p && col < p) out[(size_t)row * p + col] = h;
            }
        }
}

// C64 and C32 for a quarter (16 rows x 64 columns) of one upper tile
// (ta <= tb) of 64 x 64 and for its mirror: G_ab = sum of the split-K partial
// Gram matrices in split order, sd[a] = sqrt(G_aa / (n-1)), and
// C[a][b] = C[b][a] = clip((G_ab / (n-1)) / sd_lo / sd_hi, -1, 1) with
// lo = min(a, b): numpy corrcoef's operation order for its upper triangle,
// mirrored so that C is exactly symmetric; C32 = (float)C64.
constexpr int NQ = 4;          // row quarters per tile
constexpr int QR = CT / NQ;    // rows per quarter
__global__ __launch_bounds__(256) void normalize_tile_kernel(const Job *__restrict__ jobs, BlockRanges R) {
    int local;
    const Job &J = jobs[range_job(R, blockIdx.x, local)];
    const int p = J.p;
    const int quarter = local % NQ;
    if (local / NQ >= upper_tiles(p)) return;
    const int tiles = (p + CT - 1) / CT;
    int t = local / NQ, ta = 0;
    while (t >= tiles - ta) {
        t -= tiles - ta;
        ++ta;
    }
    const int tb = ta + t;
    const int a0 = ta * CT + quarter * QR, b0 = tb * CT;
    const CorrWs cw = carve_corr(J);
    const int ks = J.k_splits;
    const double nm1 = (double)(J.n - 1);
    const size_t pp = (size_t)p * p;
    const int tid = threadIdx.x;
    // Sums of the ks partials of R elements in split order.  The loads go out
    // four splits at a time for all R elements (clamped to split ks-1, whose
    // surplus values are not added) before any add waits on them.
    auto gsum_rows = [&](auto &g, const size_t (&off)[sizeof(g) / sizeof(double)]) {
        constexpr int R = sizeof(g) / sizeof(double);
#pragma unroll
        for (int e = 0; e < R; ++e) g[e] = 0.0;
        for (int q0 = 0; q0 < ks; q0 += 4) {
            double x[R][4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const size_t base = (size_t)(q0 + u < ks ? q0 + u : ks - 1) * pp;
#pragma unroll
                for (int e = 0; e < R; ++e) x[e][u] = cw.gram[base + off[e]];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (q0 + u < ks)
#pragma unroll
                    for (int e = 0; e < R; ++e) g[e] += x[e][u];
        }
    };
    // thread: column c, rows r0, r0 + 4, r0 + 8, r0 + 12 of the quarter, and
    // (threads below CT + QR) the diagonal entry of one of the tile's columns
    // or of the quarter's rows, for the standard deviations: all in one pass
    // of loads
    constexpr int RPT = QR / 4;
    const int c = tid & 63, r0 = tid >> 6;
    const int b = b0 + c;
    const int didx = tid < CT ? b0 + tid : a0 + (tid - CT);
    const bool dgo = tid < CT + QR && didx < p;
    size_t off[RPT + 1];
#pragma unroll
    for (int e = 0; e < RPT; ++e) {
        const int a = a0 + r0 + 4 * e;
        // lower half of a diagonal tile: the mirrored upper element
        off[e] = (a < p && b < p) ? (a <= b ? (size_t)a * p + b : (size_t)b * p + a) : 0;
    }
    off[RPT] = dgo ? (size_t)didx * p + didx : 0;
    double g[RPT + 1];
    gsum_rows(g, off);
    __shared__ double s_sdr[QR], s_sdc[CT];
    __shared__ double s_v[QR][CT + 1];
    if (tid < CT + QR) {
        const double sd = dgo ? sqrt(g[RPT] / nm1) : 1.0;
        if (tid < CT)
            s_sdc[tid] = sd;
        else
            s_sdr[tid - CT] = sd;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < RPT; ++e) {
        const int r = r0 + 4 * e, a = a0 + r;
        double v = 0.0;
        if (a < p && b < p) {
            v = g[e] / nm1;
            if (a <= b) {
                v = v / s_sdr[r];
                v = v / s_sdc[c];
            } else {
                v = v / s_sdc[c];
                v = v / s_sdr[r];
            }
            v = v > 1.0 ? 1.0 : (v < -1.0 ? -1.0 : v);
            if (J.C64) J.C64[(size_t)a * p + b] = v;
            J.C32[(size_t)a * p + b] = (float)v;
        }
        s_v[r][c] = v;
    }
    if (ta == tb) return;
    __syncthreads();
    // mirror: C[b0 + c2][a0 + r2] for the 64 x 16 transposed block
    const int r2 = tid & (QR - 1), c0 = tid / QR;
    for (int c2 = c0; c2 < CT; c2 += 256 / QR) {
        const int b = b0 + c2, a = a0 + r2;
        if (a >= p || b >= p) continue;
        const double v = s_v[r2][c2];
        if (J.C64) J.C64[(size_t)b * p + a] = v;
        J.C32[(size_t)b * p + a] = (float)v;
    }
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
// Split-K factor of one dataset's Gram matrix: ks in 1..16 (at least 256 rows
// per split; more when float32 data has more than 16 x 65536 rows) that
// minimises the busiest compute unit's work when the blocks x ks spread evenly
// over the device, ceil(blocks ks / CUs) blocks of ceil(steps / ks) row steps
// each, plus the split-K partial tiles (written once, read once; about `over`
// steps' time per block).  The integer path (blocks of two 64 x 64 tiles, one
// per compute unit, 64-row steps) and the FP64 path (64 x 64 tiles, or 32 x 32
// tiles of one wave for p <= GRAM_SMALL_P, 16-row slabs) have their own
// models.  Depends on (n, p, dtype) and the device only, so a dataset gets the
// same ks, and the same bits, alone and in any batch.
int corr_k_splits(int64_t n, int p, int dtype) {
    const bool i8 = gram_int8(dtype, p);
    // blocks per split: integer pair blocks, or FP64 work in 4-wave units
    const int64_t tu = i8 ? i8_pair_blocks(p) : gram_tile(p) == CT ? upper_tiles(p) : (upper_tiles(p, 32) + 3) / 4;
    const int64_t rows_per_step = i8 ? 64 : KT;
    const double over = 4.0;
    const int64_t cus = device_cus();
    int64_t kmax = (n + 255) / 256;
    if (kmax > 16) kmax = 16;
    // integer path: a split holds at most I8_MAX_STEPS steps
    const int64_t kmin = i8 ? (i8_rows(n) / 64 + I8_MAX_STEPS - 1) / I8_MAX_STEPS : 1;
    if (kmax < kmin) kmax = kmin;
    int best = 1;
    double best_cost = 0.0;
    for (int64_t ks = kmin; ks <= (kmax < 1 ? 1 : kmax); ++ks) {
        const int64_t rows = (n + ks - 1) / ks;
        const int64_t steps = (rows + rows_per_step - 1) / rows_per_step;
        const double cost =
            (double)((tu * ks + cus - 1) / cus) * (double)steps + over * (double)(tu * ks) / (double)cus;
        if (ks == kmin || cost < best_cost) {
            best = (int)ks;
            best_cost = cost;
        }
    }
    return best;
}

// With `one` (nj == 1), jobs_dev is where the first launch stores *one.
static int corr_jobs_launch(const Job *jobs_host, const Job *jobs_dev, int nj, const Job *one, hipStream_t s) {
    if (nj < 1) return OK;
    int maxp = 0, dtype = jobs_host[0].x_dtype;
    bool any_i8 = false;  // some job takes the integer path
    int64_t maxn = 0;
    for (int b = 0; b < nj; ++b) {
        const Job &J = jobs_host[b];
        if (!J.X || !J.C32 || !J.corr_ws || J.n < 2 || J.p < 1 || J.k_splits < 1)
            return set_error(ERR_ARG, "corr: job %d has bad arguments", b);
        if (J.x_dtype != dtype) return set_error(ERR_ARG, "corr: jobs of one launch must share x_dtype");
        if (gram_int8(dtype, J.p) && (i8_rows(J.n) / 64 + J.k_splits - 1) / J.k_splits > I8_MAX_STEPS)
            return set_error(ERR_ARG, "corr: job %d: k_splits %d leaves more than %d rows per split", b, J.k_splits,
                             (int)(64 * I8_MAX_STEPS));
        maxp = J.p > maxp ? J.p : maxp;
        maxn = J.n > maxn ? J.n : maxn;
        any_i8 = any_i8 || gram_int8(dtype, J.p);
    }
    if (dtype != 0 && dtype != 1) return set_error(ERR_ARG, "corr: dtype must be 0 (f32) or 1 (f64)");
    const int rs = corr_row_splits(maxn);
    const int cb = (maxp + 63) / 64;
    Job *dst = const_cast<Job *>(jobs_dev);
    if (one && dtype == 0)
        GPUCSL_LAUNCH("corr_colsum", colsum_one_kernel<float>, dim3(cb, rs, 1), dim3(256), 0, s, *one, dst);
    else if (one)
        GPUCSL_LAUNCH("corr_colsum", colsum_one_kernel<double>, dim3(cb, rs, 1), dim3(256), 0, s, *one, dst);
    else {
        const int rc = for_block_ranges(
            nj,
            [&](int b) { return (long long)((jobs_host[b].p + 63) / 64) * corr_row_splits(jobs_host[b].n); },
            [&](const BlockRanges &R) -> int {
                if (dtype == 0)
                    GPUCSL_LAUNCH("corr_colsum", colsum_kernel<float>, dim3(R.first[R.count]), dim3(256), 0, s,
                                  jobs_dev, R);
                else
                    GPUCSL_LAUNCH("corr_colsum", colsum_kernel<double>, dim3(R.first[R.count]), dim3(256), 0, s,
                                  jobs_dev, R);
                return check_hip(hipGetLastError(), "corr_colsum launch");
            });
        if (rc) return rc;
    }
    if (any_i8) {
        int rc = for_block_ranges(
            nj,
            [&](int b) -> long long {
                const Job &J = jobs_host[b];
                if (!gram_int8(dtype, J.p)) return 0;
                return (i8_rows(J.n) / 64 + I8_SLICE_STEPS - 1) / I8_SLICE_STEPS * (long long)(i8_cols(J.p) / 64);
            },
            [&](const BlockRanges &R) -> int {
                GPUCSL_LAUNCH("corr_slice_i8", slice_i8_kernel, dim3(R.first[R.count]), dim3(256), 0, s, jobs_dev, R);
                return check_hip(hipGetLastError(), "corr_slice_i8 launch");
            });
        if (rc) return rc;
        // the integer-path jobs, I8_LAUNCH_JOBS per launch, each launch's jobs
        // longest blocks (row steps per split) first
        auto block_steps = [&](int b) {
            return (i8_rows(jobs_host[b].n) / 64 + jobs_host[b].k_splits - 1) / jobs_host[b].k_splits;
        };
        I8Launch L{};
        for (int b = 0; b < nj; ++b) {
            if (gram_int8(dtype, jobs_host[b].p)) L.job[L.count++] = b;
            if (L.count > 0 && (L.count == I8_LAUNCH_JOBS || b == nj - 1)) {
                for (int k = 1; k < L.count; ++k)  // insertion sort, stable
                    for (int m = k; m > 0 && block_steps(L.job[m]) > block_steps(L.job[m - 1]); --m) {
                        const int t = L.job[m];
                        L.job[m] = L.job[m - 1];
                        L.job[m - 1] = t;
                    }
                for (int k = 0; k < L.count; ++k)
                    L.first[k + 1] = L.first[k] + i8_pair_blocks(jobs_host[L.job[k]].p) * jobs_host[L.job[k]].k_splits;
                GPUCSL_LAUNCH("corr_gram_i8", gram_i8_kernel, dim3(L.first[L.count]), dim3(256), 0, s, jobs_dev, L);
                L = I8Launch{};
            }
        }
    }
    // FP64 tiles: float64 data, and float32 data up to GRAM_SMALL_P columns
    // (64 x 64 tiles above GRAM_SMALL_P, 32 x 32 tiles of one wave up to it)
    for (int tile : {CT, 32}) {
        const int rc = for_block_ranges(
            nj,
            [&](int b) -> long long {
                const Job &J = jobs_host[b];
                if (gram_int8(dtype, J.p) || gram_tile(J.p) != tile) return 0;
                return (long long)upper_tiles(J.p, tile) * J.k_splits;
            },
            [&](const BlockRanges &R) -> int {
                const dim3 grid(R.first[R.count]);
                if (tile == CT)
                    GPUCSL_LAUNCH("corr_gram_mfma_f64", (gram_mfma_kernel<double, CT>), grid, dim3(256), 0, s, jobs_dev,
                                  R);
                else if (dtype == 0)
                    GPUCSL_LAUNCH("corr_gram_mfma_f64", (gram_mfma_kernel<float, 32>), grid, dim3(64), 0, s, jobs_dev, R);
                else
                    GPUCSL_LAUNCH("corr_gram_mfma_f64", (gram_mfma_kernel<double, 32>), grid, dim3(64), 0, s, jobs_dev,
                                  R);
                return check_hip(hipGetLastError(), "corr_gram_mfma_f64 launch");
            });
        if (rc) return rc;
    }
    {
        const int rc = for_block_ranges(
            nj, [&](int b) { return (long long)upper_tiles(jobs_host[b].p) * NQ; },
            [&](const BlockRanges &R) -> int {
                GPUCSL_LAUNCH("corr_normalize", normalize_tile_kernel, dim3(R.first[R.count]), dim3(256), 0, s,
                              jobs_dev, R);
                return check_hip(hipGetLastError(), "corr_normalize launch");
            });
        if (rc) return rc;
    }
    return check_hip(hipGetLastError(), "corr launch");
}

int corr_batch_launch(const Job *jobs_host, const Job *jobs_dev, int nj, hipStream_t s) {
    return corr_jobs_launch(jobs_host, jobs_dev, nj, nullptr, s);
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
    int ks = corr_k_splits(n, p, 0);
    const int ks1 = corr_k_splits(n, p, 1);
    ks = ks1 > ks ? ks1 : ks;
    return JOB_DESC_BYTES + corr_job_ws_bytes(n, p, ks);
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
    j.k_splits = corr_k_splits(n, p, dtype);
    j.X = X;
    j.C32 = C32;
    j.C64 = C64;
    j.corr_ws = (char *)ws + JOB_DESC_BYTES;
    return corr_jobs_launch(&j, (Job *)ws, 1, &j, s);
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
