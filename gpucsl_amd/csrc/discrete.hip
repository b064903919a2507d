// Discrete (Pearson chi^2) PC-stable skeleton on gfx950.
//
// Replaces gpucsl/pc/cuda/discrete_ci.cu (discrete_ci_level_0 / _level_n) and
// gpucsl/pc/cuda/helpers/discrete_helpers.cu, with the host level loop of
// gpucsl/pc/discover_skeleton_discrete.py:80-112.
//
// Data layout in HBM: uint8 category codes, column-major with a padded column
// stride (data[v * stride + k], stride a multiple of 16, padding ignored), so a
// lane reads 16 observations of a variable with one 16-byte load.
//
// One 256-thread workgroup evaluates one CI test: it counts the contingency
// table N[g][a][b] (g = stratum of S in mixed radix, first sepset variable least
// significant; a = value of v_i; b = value of v_j) with LDS atomics into one
// table replica per wave, folds the replicas, forms the integer marginals, sums
// chi^2 in FP64 in a fixed order and evaluates Perlman's pchisq.  Counts are
// exact integers.  Level 0 sums chi^2 in the reference's sequential (a, b)
// order; levels >= 1 use 64 lane-strided partial sums and a halving tree (the
// reference's order there is that of atomicAdd(double) and not defined).
//
// Every sepset of every pair adjacent at level start is tested, as in the
// reference (its early exit is a `continue`, discrete_ci.cu:240-243), so pmax is
// the max over that full enumeration.  The removing sepset is the first passing
// test in pcalg order (key = side << 62 | rank, side 0 for the row of the
// smaller endpoint), chosen with one atomicMin per passing test and written by
// a writeback kernel after the level.
#include "common.h"

namespace gpucsl {

// in gauss.hip
int compact_launch(const uint8_t *Gm, int p, int level, void *ws, unsigned long long *stats,
                   hipStream_t s);
size_t gauss_workspace_bytes(int p);
void level_list_views(void *ws, int p, int32_t **nbr, int32_t **deg);

constexpr int DT_THREADS = 256;
constexpr int DT_LDS_BYTES = 32 * 1024;
constexpr int DT_LDS_WORDS = DT_LDS_BYTES / 4;
constexpr int DT_BLOCKS_PER_CU = 4;
constexpr unsigned long long KEY_NONE = ~0ull;
enum DStat { DS_EXEC = 0, DS_TOTAL = 1, DS_MAXDEG = 5, DS_USEFUL = 6 };

// int[L] with a one-element stand-in for L = 0
template <int L>
using IArr = int[L > 0 ? L : 1];

struct DiscWs {
    unsigned long long *testoff;  // p + 1 prefix of tests per row
    unsigned long long *keys;     // p * p removal keys, upper triangle used
    int32_t *gram;                // level-0 Gram matrix of the one-hot codes (gram_w^2 words)
    int gram_w;                   // its padded width; 0: level 0 counts pair by pair
    uint32_t *gtab;               // global tables for tests too large for LDS
    size_t gtab_words;            // words per workgroup slice
    int grid;                     // workgroups of the test kernels
};

// ----------------------------------------------------------------------------
// one CI test inside a workgroup
// ----------------------------------------------------------------------------
// Counts the table of (vi, vj | S) into `tab` (cells words, zeroed here) using
// `replicas` copies at stride `cells`; returns after the replicas are folded
// into copy 0.  All threads of the block must call it.
template <int L>
__device__ void count_table(const uint8_t *__restrict__ data, int64_t n, int64_t stride, int vi,
                            int vj, const IArr<L> &S, int di, int dj, const IArr<L> &radix,
                            uint32_t *tab, int cells, int replicas) {
    const int tid = threadIdx.x;
    for (int c = tid; c < cells * replicas; c += DT_THREADS) tab[c] = 0;
    __syncthreads();
    uint32_t *mine = tab + (size_t)((tid >> 6) % replicas) * cells;
    const uint8_t *ci = data + (size_t)vi * stride;
    const uint8_t *cj = data + (size_t)vj * stride;
    const int64_t chunks = (n + 15) / 16;
    for (int64_t ch = tid; ch < chunks; ch += DT_THREADS) {
        const uint4 wa = *reinterpret_cast<const uint4 *>(ci + ch * 16);
        const uint4 wb = *reinterpret_cast<const uint4 *>(cj + ch * 16);
        uint4 ws[L > 0 ? L : 1];
#pragma unroll
        for (int c = 0; c < L; ++c) ws[c] = *reinterpret_cast<const uint4 *>(data + (size_t)S[c] * stride + ch * 16);
        const int lim = (ch * 16 + 16 <= n) ? 16 : (int)(n - ch * 16);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            if (q < lim) {
                const int word = q >> 2, sh = (q & 3) * 8;
                const unsigned a = ((&wa.x)[word] >> sh) & 0xffu;
                const unsigned b = ((&wb.x)[word] >> sh) & 0xffu;
                unsigned g = 0;
#pragma unroll
                for (int c = 0; c < L; ++c) g += (((&ws[c].x)[word] >> sh) & 0xffu) * (unsigned)radix[c];
                GPUCSL_DCHECK(a < (unsigned)di && b < (unsigned)dj && (g * di + a) * dj + b < (unsigned)cells);
                atomicAdd(&mine[(g * di + a) * dj + b], 1u);
            }
        }
    }
    __syncthreads();
    if (replicas > 1) {
        for (int c = tid; c < cells; c += DT_THREADS) {
            uint32_t v = tab[c];
            for (int r = 1; r < replicas; ++r) v += tab[(size_t)r * cells + c];
            tab[c] = v;
        }
        __syncthreads();
    }
}

// Integer marginals: Nis[g*di + a], Njs[g*dj + b], Ns[g]
// (gpucsl/pc/cuda/helpers/discrete_helpers.cu:97-133).
__device__ void marginals(const uint32_t *N, int dim_s, int di, int dj, uint32_t *Nis, uint32_t *Njs,
                          uint32_t *Ns) {
    const int tid = threadIdx.x;
    for (int r = tid; r < dim_s * di; r += DT_THREADS) {
        const int g = r / di, a = r - g * di;
        uint32_t v = 0;
        for (int b = 0; b < dj; ++b) v += N[(g * di + a) * dj + b];
        Nis[r] = v;
    }
    for (int r = tid; r < dim_s * dj; r += DT_THREADS) {
        const int g = r / dj, b = r - g * dj;
        uint32_t v = 0;
        for (int a = 0; a < di; ++a) v += N[(g * di + a) * dj + b];
        Njs[r] = v;
    }
    for (int g = tid; g < dim_s; g += DT_THREADS) {
        uint32_t v = 0;
        for (int c = 0; c < di * dj; ++c) v += N[g * di * dj + c];
        Ns[g] = v;
    }
    __syncthreads();
}

// Per-cell term (O - E)^2 / E with E = N_i+s * N_+js / N_++s, skipping
// N_++s = 0 or E = 0 (discrete_helpers.cu:140-186).
__device__ __forceinline__ double chi2_term(const uint32_t *N, const uint32_t *Nis, const uint32_t *Njs,
                                            const uint32_t *Ns, int di, int dj, int c) {
#pragma clang fp contract(off)
    const int g = c / (di * dj), a = (c / dj) % di, b = c % dj;
    const double e = Nis[g * di + a] * (double)Njs[g * dj + b] / Ns[g];
    if (Ns[g] == 0 || e == 0.0) return 0.0;
    const double d = N[c] - e;
    return d * d / e;
}

// chi^2 of the table; returns the value in thread 0 (others: unspecified).
__device__ double chi2_sum(const uint32_t *N, const uint32_t *Nis, const uint32_t *Njs, const uint32_t *Ns,
                           int dim_s, int di, int dj, bool sequential) {
#pragma clang fp contract(off)
    const int cells = dim_s * di * dj;
    double acc = 0.0;
    if (sequential) {
        if (threadIdx.x == 0)
            for (int c = 0; c < cells; ++c) {
                acc += chi2_term(N, Nis, Njs, Ns, di, dj, c);
            }
        return acc;
    }
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        for (int c = lane; c < cells; c += 64) acc += chi2_term(N, Nis, Njs, Ns, di, dj, c);
        for (int off = 32; off >= 1; off >>= 1) acc = acc + __shfl_down(acc, off);
    }
    return acc;
}

// ----------------------------------------------------------------------------
// level 0: gpucsl/pc/cuda/discrete_ci.cu:37-97
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(DT_THREADS) void disc_level0_kernel(
    const uint8_t *__restrict__ data, int64_t n, int64_t stride, const uint8_t *__restrict__ dims, int p,
    double alpha, uint8_t *__restrict__ Gm, float *__restrict__ pmax, uint32_t *__restrict__ gtab,
    size_t gtab_words, unsigned long long *__restrict__ st) {
    extern __shared__ uint32_t lds[];
    const long long pairs = (long long)p * (p - 1) / 2;
    if (blockIdx.x == 0) {
        for (int v = threadIdx.x; v < p; v += DT_THREADS) Gm[(size_t)v * p + v] = 0;
        if (threadIdx.x == 0) {
            st[DS_EXEC] = pairs;
            st[DS_TOTAL] = pairs;
            st[DS_USEFUL] = pairs;
        }
    }
    const IArr<0> dummyS = {0};
    const IArr<0> noradix = {1};
    for (long long t = blockIdx.x; t < pairs; t += gridDim.x) {
        // t -> (vi, vj) with vj < vi: vi(vi-1)/2 <= t < vi(vi+1)/2
        int vi = (int)((1.0 + sqrt(1.0 + 8.0 * (double)t)) * 0.5);
        while ((long long)vi * (vi - 1) / 2 > t) --vi;
        while ((long long)(vi + 1) * vi / 2 <= t) ++vi;
        const int vj = (int)(t - (long long)vi * (vi - 1) / 2);
        const int di = dims[vi], dj = dims[vj];
        const int cells = di * dj;
        const int marg = di + dj + 1;
        uint32_t *tab;
        int replicas;
        if (4 * cells + marg <= DT_LDS_WORDS) {
            tab = lds;
            replicas = 4;
        } else if (cells + marg <= DT_LDS_WORDS) {
            tab = lds;
            replicas = 1;
        } else {
            tab = gtab + (size_t)blockIdx.x * gtab_words;
            replicas = 1;
        }
        count_table<0>(data, n, stride, vi, vj, dummyS, di, dj, noradix, tab, cells, replicas);
        uint32_t *Nis = tab + (size_t)replicas * cells, *Njs = Nis + di, *Ns = Njs + dj;
        marginals(tab, 1, di, dj, Nis, Njs, Ns);
        const double chi2 = chi2_sum(tab, Nis, Njs, Ns, 1, di, dj, true);
        if (threadIdx.x == 0) {
            const double pv = pchisq(chi2, (di - 1) * (dj - 1));
            pmax[(size_t)vj * p + vi] = (float)pv;
            const uint8_t keep = (pv >= alpha) ? 0 : 1;
            Gm[(size_t)vi * p + vj] = keep;
            Gm[(size_t)vj * p + vi] = keep;
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------------------
// level 0 on int8 MFMA: the contingency tables of every pair are blocks of one
// Gram matrix.  With the one-hot encoding H (n x W, W = p D padded to 64,
// column v D + a = [x_v = a]), N = H^T H holds N[v D + a][u D + b] =
// #{k : x_v[k] = a, x_u[k] = b}: the level-0 table of (v, u) for v != u.
// ----------------------------------------------------------------------------
constexpr int GRAM_TILE = 64;
constexpr int GRAM_MAX_DIM = 16;
constexpr int GRAM_MAX_W = 4096;

// padded one-hot width, or 0 when level 0 counts its tables pair by pair
// (categories above GRAM_MAX_DIM or a Gram matrix above GRAM_MAX_W^2 words)
static int gram_width(int p, int max_dim) {
    if (max_dim < 1 || max_dim > GRAM_MAX_DIM) return 0;
    const size_t w = align_up((size_t)p * max_dim, GRAM_TILE);
    return w <= (size_t)GRAM_MAX_W ? (int)w : 0;
}

typedef int gram_v4i __attribute__((ext_vector_type(4)));

// 8 observations of one variable as one-hot bytes for category `a`: byte q is
// 1 iff x[q] == a (exact per-byte zero test, no carries between bytes)
__device__ __forceinline__ uint64_t onehot8(uint64_t x, unsigned a) {
    const uint64_t t = x ^ (0x0101010101010101ull * a);
    const uint64_t lo7 = 0x7f7f7f7f7f7f7f7full;
    const uint64_t z = ~(((t & lo7) + lo7) | t | lo7);
    return z >> 7;
}

// N (W x W int32, zeroed before) += the upper 64 x 64 tiles (tile row <= tile
// column) of H^T H over this block's share of the observations.  256 threads
// = 4 waves in a 2 x 2 arrangement of 32 x 32 sub-tiles, each 2 x 2
// v_mfma_i32_16x16x32_i8 blocks.  Operand layout (checked on gfx950): lane l
// gives row/column l % 16 of its block and observations 8 (l / 16) .. +7 of the
// 32-observation step; accumulator item v holds row 4 (l / 16) + v, column
// l % 16.  Split-K shares add with int32 atomics, exact in any order.
// Upper tile `tile` of share `split` of `nsplit`.
__device__ __forceinline__ void disc_gram_body(const uint8_t *__restrict__ data, int64_t n, int64_t stride, int p,
                                               int D, int W, int32_t *__restrict__ N, int tile, int split,
                                               int nsplit) {
    const int tiles = W / GRAM_TILE;
    int t = tile, tr = 0;
    while (t >= tiles - tr) {
        t -= tiles - tr;
        ++tr;
    }
    const int tc = tr + t;
    const int64_t steps = (n + 31) / 32;
    const int64_t per = (steps + nsplit - 1) / nsplit;
    const int64_t s0 = (int64_t)split * per;
    const int64_t s1 = s0 + per < steps ? s0 + per : steps;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int ci = lane & 15, kq = lane >> 4;
    const uint8_t *pa[2], *pb[2];
    unsigned ca[2], cb[2];
    bool va[2], vb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int r = tr * GRAM_TILE + wr + i * 16 + ci;
        const int c = tc * GRAM_TILE + wc + i * 16 + ci;
        va[i] = r / D < p;
        vb[i] = c / D < p;
        pa[i] = data + (size_t)(va[i] ? r / D : 0) * stride;
        pb[i] = data + (size_t)(vb[i] ? c / D : 0) * stride;
        ca[i] = (unsigned)(r % D);
        cb[i] = (unsigned)(c % D);
    }
    gram_v4i acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = gram_v4i{0, 0, 0, 0};
    for (int64_t st = s0; st < s1; ++st) {
        const int64_t k = st * 32 + 8 * kq;
        // observations k .. k+7; bytes at or past n are masked (reads stay
        // below n rounded up to 8 <= stride)
        const int64_t rem = n - k;
        const uint64_t mask = rem >= 8 ? ~0ull : (rem <= 0 ? 0ull : ((1ull << (8 * rem)) - 1ull));
        uint64_t xa[2] = {0, 0}, xb[2] = {0, 0};
        if (rem > 0) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                if (va[i]) xa[i] = *reinterpret_cast<const uint64_t *>(pa[i] + k);
                if (vb[i]) xb[i] = *reinterpret_cast<const uint64_t *>(pb[i] + k);
            }
        }
        long oa[2], ob[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            oa[i] = va[i] ? (long)(onehot8(xa[i], ca[i]) & mask) : 0;
            ob[i] = vb[i] ? (long)(onehot8(xb[i], cb[i]) & mask) : 0;
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x32_i8(oa[i], ob[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = tr * GRAM_TILE + wr + i * 16 + 4 * kq + v;
                const int c = tc * GRAM_TILE + wc + j * 16 + ci;
                if (acc[i][j][v] != 0) atomicAdd(&N[(size_t)r * W + c], acc[i][j][v]);
            }
}

__global__ __launch_bounds__(256) void disc_onehot_gram_kernel(const uint8_t *__restrict__ data, int64_t n,
                                                               int64_t stride, int p, int D, int W,
                                                               int32_t *__restrict__ N) {
    disc_gram_body(data, n, stride, p, D, W, N, blockIdx.x, blockIdx.y, gridDim.y);
}

// One thread per pair (vj < vi): the table N[a][b] (a a value of vi, b of vj)
// read from the Gram matrix, integer marginals, chi^2 summed over the cells in
// (a, b) order with the per-cell term of chi2_term, Perlman's pchisq, then pmax
// and G as disc_level0_kernel writes them.  MD >= max_dim bounds the
// marginal arrays.  Each cell is zeroed after its last read: the cells never
// read are those of categories without observations (never added to) and the
// diagonal blocks (never read), so the next level-0 run starts from zero.
// Pair t of the block of pair blocks `blk` (block 0 also clears G's diagonal
// and writes the level's statistics).
template <int MD>
__device__ __forceinline__ void disc_level0_chi2_body(const uint8_t *__restrict__ dims, int p, int D, int W,
                                                      int32_t *__restrict__ N, double alpha,
                                                      uint8_t *__restrict__ Gm, float *__restrict__ pmax,
                                                      unsigned long long *__restrict__ st, int blk) {
#pragma clang fp contract(off)
    const long long pairs = (long long)p * (p - 1) / 2;
    if (blk == 0) {
        for (int v = threadIdx.x; v < p; v += blockDim.x) Gm[(size_t)v * p + v] = 0;
        if (threadIdx.x == 0) {
            st[DS_EXEC] = pairs;
            st[DS_TOTAL] = pairs;
            st[DS_USEFUL] = pairs;
        }
    }
    const long long t = (long long)blk * blockDim.x + threadIdx.x;
    if (t >= pairs) return;
    int vi = (int)((1.0 + sqrt(1.0 + 8.0 * (double)t)) * 0.5);
    while ((long long)vi * (vi - 1) / 2 > t) --vi;
    while ((long long)(vi + 1) * vi / 2 <= t) ++vi;
    const int vj = (int)(t - (long long)vi * (vi - 1) / 2);
    const int di = dims[vi], dj = dims[vj];
    // N[a][b] = Gram[vj D + b][vi D + a] (upper triangle: vj D + b < vi D)
    int32_t *T = N + (size_t)vj * D * W + (size_t)vi * D;
    uint32_t Ni[MD], Nj[MD], Ns = 0;
#pragma unroll
    for (int a = 0; a < MD; ++a) Ni[a] = Nj[a] = 0;
#pragma unroll
    for (int a = 0; a < MD; ++a)
#pragma unroll
        for (int b = 0; b < MD; ++b)
            if (a < di && b < dj) {
                const uint32_t v = (uint32_t)T[(size_t)b * W + a];
                Ni[a] += v;
                Nj[b] += v;
                Ns += v;
            }
    double chi2 = 0.0;
#pragma unroll
    for (int a = 0; a < MD; ++a)
#pragma unroll
        for (int b = 0; b < MD; ++b)
            if (a < di && b < dj) {
                int32_t *cell = T + (size_t)b * W + a;
                const uint32_t o = (uint32_t)*cell;
                *cell = 0;  // leave the pair's table zero for the next level-0 run
                const double e = Ni[a] * (double)Nj[b] / Ns;
                double term = 0.0;
                if (!(Ns == 0 || e == 0.0)) {
                    const double d = o - e;
                    term = d * d / e;
                }
                chi2 += term;
            }
    const double pv = pchisq(chi2, (di - 1) * (dj - 1));
    pmax[(size_t)vj * p + vi] = (float)pv;
    const uint8_t keep = (pv >= alpha) ? 0 : 1;
    Gm[(size_t)vi * p + vj] = keep;
    Gm[(size_t)vj * p + vi] = keep;
}

template <int MD>
__global__ __launch_bounds__(256) void disc_level0_chi2_kernel(const uint8_t *__restrict__ dims, int p, int D, int W,
                                                               int32_t *__restrict__ N, double alpha,
                                                               uint8_t *__restrict__ Gm, float *__restrict__ pmax,
                                                               unsigned long long *__restrict__ st) {
    disc_level0_chi2_body<MD>(dims, p, D, W, N, alpha, Gm, pmax, st, blockIdx.x);
}

// Buffer set-up of a discrete skeleton in one launch: G = 1, sepsets = -1,
// pmax = 0, removal keys = none, the level-0 Gram matrix and the statistics 0.
__device__ __forceinline__ void disc_init_body(uint8_t *__restrict__ Gm, int32_t *__restrict__ sep,
                                               float *__restrict__ pmax, unsigned long long *__restrict__ keys,
                                               int32_t *__restrict__ gram, unsigned long long *__restrict__ st,
                                               size_t pp, size_t ns, size_t ng, int nst, size_t t0, size_t stride) {
    for (size_t t = t0; t < pp; t += stride) {
        Gm[t] = 1;
        pmax[t] = 0.0f;
        keys[t] = KEY_NONE;
    }
    if (sep)
        for (size_t t = t0; t < ns; t += stride) sep[t] = -1;
    if (gram)
        for (size_t t = t0; t < ng; t += stride) gram[t] = 0;
    for (size_t t = t0; st && t < (size_t)nst; t += stride) st[t] = 0;
}

__global__ void disc_init_kernel(uint8_t *__restrict__ Gm, int32_t *__restrict__ sep, float *__restrict__ pmax,
                                 unsigned long long *__restrict__ keys, int32_t *__restrict__ gram,
                                 unsigned long long *__restrict__ st, size_t pp, size_t ns, size_t ng, int nst) {
    disc_init_body(Gm, sep, pmax, keys, gram, st, pp, ns, ng, nst, (size_t)blockIdx.x * blockDim.x + threadIdx.x,
                   (size_t)gridDim.x * blockDim.x);
}

// ----------------------------------------------------------------------------
// level l >= 1: gpucsl/pc/cuda/discrete_ci.cu:113-246
// ----------------------------------------------------------------------------
// testoff[r] = sum over rows < r of deg * C(deg - 1, L); also T_l, max degree.
// By one block of 1024 threads.
__device__ __forceinline__ void disc_testoff_body(const int32_t *__restrict__ deg, int p, int L,
                                                  unsigned long long *__restrict__ testoff,
                                                  unsigned long long *__restrict__ st) {
    __shared__ unsigned long long sT[1024];
    __shared__ int sM[1024];
    const int t = threadIdx.x;
    const int per = (p + 1023) / 1024;
    const int r0 = t * per;
    const int r1 = min(p, r0 + per);
    unsigned long long loc = 0;
    int md = 0;
    for (int r = r0; r < r1; ++r) {
        const int d = deg[r];
        loc += (unsigned long long)d * binom(d - 1, L);
        md = max(md, d);
    }
    sT[t] = loc;
    sM[t] = md;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        unsigned long long v = t >= off ? sT[t - off] : 0;
        __syncthreads();
        sT[t] += v;
        __syncthreads();
    }
    unsigned long long e = sT[t] - loc;
    for (int r = r0; r < r1; ++r) {
        testoff[r] = e;
        e += (unsigned long long)deg[r] * binom(deg[r] - 1, L);
    }
    if (t == 1023) {
        testoff[p] = sT[1023];
        st[DS_EXEC] = sT[1023];
        st[DS_TOTAL] = sT[1023];
        st[DS_USEFUL] = sT[1023];
    }
    __syncthreads();
    for (int off = 512; off >= 1; off >>= 1) {
        if (t < off) sM[t] = max(sM[t], sM[t + off]);
        __syncthreads();
    }
    if (t == 0) st[DS_MAXDEG] = (unsigned long long)sM[0];
}

__global__ __launch_bounds__(1024) void disc_testoff_kernel(const int32_t *__restrict__ deg, int p, int L,
                                                            unsigned long long *__restrict__ testoff,
                                                            unsigned long long *__restrict__ st) {
    disc_testoff_body(deg, p, L, testoff, st);
}

// Test t (< testoff[p]) of a level: the row i whose tests hold t, the
// neighbour j and the sepset of rank `rank` among the C(deg - 1, L) subsets of
// i's other neighbours; the table in LDS (gslice: this workgroup's global
// table when it does not fit).  All threads of the block call it.
template <int L>
__device__ __forceinline__ void disc_test(const uint8_t *__restrict__ data, int64_t n, int64_t stride,
                                          const uint8_t *__restrict__ dims, int p, double alpha,
                                          const int32_t *__restrict__ nbr, const int32_t *__restrict__ deg,
                                          const unsigned long long *__restrict__ testoff, float *__restrict__ pmax,
                                          unsigned long long *__restrict__ keys, uint32_t *gslice,
                                          unsigned long long t, uint32_t *lds) {
    {
        int lo = 0, hi = p;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (testoff[mid] <= t) lo = mid; else hi = mid;
        }
        const int i = lo;
        const int d = deg[i];
        const unsigned long long per_pair = binom(d - 1, L);
        const unsigned long long o = t - testoff[i];
        const int jpos = (int)(o / per_pair);
        const unsigned long long rank = o - (unsigned long long)jpos * per_pair;
        const int32_t *row = nbr + (size_t)i * p;
        GPUCSL_DCHECK(i >= 0 && i < p && jpos >= 0 && jpos < d);
        const int j = row[jpos];
        int pos[L];
        IArr<L> S, radix;
        unrank_lex<L>(d - 1, rank, pos);
        int dim_s = 1;
#pragma unroll
        for (int c = 0; c < L; ++c) {
            const int idx = pos[c] - 1;
            S[c] = row[idx + (idx >= jpos)];
            radix[c] = dim_s;
            dim_s *= dims[S[c]];
        }
        const int di = dims[i], dj = dims[j];
        const int cells = dim_s * di * dj;
        const int marg = dim_s * (di + dj + 1);
        uint32_t *tab;
        int replicas;
        if (4 * cells + marg <= DT_LDS_WORDS) {
            tab = lds;
            replicas = 4;
        } else if (cells + marg <= DT_LDS_WORDS) {
            tab = lds;
            replicas = 1;
        } else {
            tab = gslice;
            replicas = 1;
        }
        count_table<L>(data, n, stride, i, j, S, di, dj, radix, tab, cells, replicas);
        uint32_t *Nis = tab + (size_t)replicas * cells, *Njs = Nis + dim_s * di, *Ns = Njs + dim_s * dj;
        marginals(tab, dim_s, di, dj, Nis, Njs, Ns);
        double chi2 = chi2_sum(tab, Nis, Njs, Ns, dim_s, di, dj, false);
        if (threadIdx.x < 64) {
            // the sum sits in lane 0; wave 0 evaluates the p-value together
            chi2 = __shfl(chi2, 0);
            const float pf = (float)pchisq_wave(chi2, (di - 1) * (dj - 1) * dim_s);
            const int a = i < j ? i : j, b = i < j ? j : i;
            if (threadIdx.x == 0) atomic_max_nonneg(&pmax[(size_t)a * p + b], pf);
            if (threadIdx.x == 0 && (double)pf >= alpha) {
                const unsigned long long key = ((unsigned long long)(i < j ? 0 : 1) << 62) | rank;
                atomicMin(&keys[(size_t)a * p + b], key);
            }
        }
        __syncthreads();
    }
}

template <int L>
__global__ __launch_bounds__(DT_THREADS) void disc_level_kernel(
    const uint8_t *__restrict__ data, int64_t n, int64_t stride, const uint8_t *__restrict__ dims, int p,
    double alpha, const int32_t *__restrict__ nbr, const int32_t *__restrict__ deg,
    const unsigned long long *__restrict__ testoff, float *__restrict__ pmax,
    unsigned long long *__restrict__ keys, uint32_t *__restrict__ gtab, size_t gtab_words) {
    extern __shared__ uint32_t lds[];
    const unsigned long long T = testoff[p];
    for (unsigned long long t = blockIdx.x; t < T; t += gridDim.x)
        disc_test<L>(data, n, stride, dims, p, alpha, nbr, deg, testoff, pmax, keys,
                     gtab + (size_t)blockIdx.x * gtab_words, t, lds);
}

// Removes the edges whose key was set, writes their sepsets in both
// directions and resets the keys for the next level.
template <int L>
__device__ __forceinline__ void disc_writeback_pair(int p, const int32_t *__restrict__ nbr,
                                                    const int32_t *__restrict__ deg,
                                                    unsigned long long *__restrict__ keys, uint8_t *__restrict__ Gm,
                                                    int32_t *__restrict__ sep, int maxL, int a, int b) {
    if (b >= p || b <= a) return;
    const unsigned long long key = keys[(size_t)a * p + b];
    if (key == KEY_NONE) return;
    keys[(size_t)a * p + b] = KEY_NONE;
    const int side = (int)(key >> 62);
    const unsigned long long rank = key & ((1ull << 62) - 1);
    const int i = side ? b : a, j = side ? a : b;
    const int32_t *row = nbr + (size_t)i * p;
    const int d = deg[i];
    int lo = 0, hi = d;  // position of j in the ascending neighbour list
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (row[mid] < j) lo = mid + 1; else hi = mid;
    }
    const int jpos = lo;
    int pos[L];
    unrank_lex<L>(d - 1, rank, pos);
    for (int c = 0; c < L; ++c) {
        const int idx = pos[c] - 1;
        const int v = row[idx + (idx >= jpos)];
        sep[((size_t)i * p + j) * maxL + c] = v;
        sep[((size_t)j * p + i) * maxL + c] = v;
    }
    Gm[(size_t)a * p + b] = 0;
    Gm[(size_t)b * p + a] = 0;
}

template <int L>
__global__ void disc_writeback_kernel(int p, const int32_t *__restrict__ nbr, const int32_t *__restrict__ deg,
                                      unsigned long long *__restrict__ keys, uint8_t *__restrict__ Gm,
                                      int32_t *__restrict__ sep, int maxL) {
    disc_writeback_pair<L>(p, nbr, deg, keys, Gm, sep, maxL, blockIdx.y, blockIdx.x * blockDim.x + threadIdx.x);
}

// ----------------------------------------------------------------------------
// levels above the templated range (9..GPUCSL_MAX_LEVEL), the level taken at
// run time: the same test per workgroup, with the conditioning set and its
// radices in LDS and the stratum index of 16 observations accumulated
// conditioning variable by conditioning variable (one 16-byte load each).
// ----------------------------------------------------------------------------
__device__ void count_table_rt(const uint8_t *__restrict__ data, int64_t n, int64_t stride, int vi, int vj,
                               const int *S, const int *radix, int L, int di, int dj, uint32_t *tab, int cells,
                               int replicas) {
    const int tid = threadIdx.x;
    for (int c = tid; c < cells * replicas; c += DT_THREADS) tab[c] = 0;
    __syncthreads();
    uint32_t *mine = tab + (size_t)((tid >> 6) % replicas) * cells;
    const uint8_t *ci = data + (size_t)vi * stride;
    const uint8_t *cj = data + (size_t)vj * stride;
    const int64_t chunks = (n + 15) / 16;
    for (int64_t ch = tid; ch < chunks; ch += DT_THREADS) {
        unsigned g[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) g[q] = 0;
        for (int c = 0; c < L; ++c) {
            const uint4 w = *reinterpret_cast<const uint4 *>(data + (size_t)S[c] * stride + ch * 16);
            const unsigned rad = (unsigned)radix[c];
#pragma unroll
            for (int q = 0; q < 16; ++q) g[q] += (((&w.x)[q >> 2] >> ((q & 3) * 8)) & 0xffu) * rad;
        }
        const uint4 wa = *reinterpret_cast<const uint4 *>(ci + ch * 16);
        const uint4 wb = *reinterpret_cast<const uint4 *>(cj + ch * 16);
        const int lim = (ch * 16 + 16 <= n) ? 16 : (int)(n - ch * 16);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            if (q < lim) {
                const int word = q >> 2, sh = (q & 3) * 8;
                const unsigned a = ((&wa.x)[word] >> sh) & 0xffu;
                const unsigned b = ((&wb.x)[word] >> sh) & 0xffu;
                GPUCSL_DCHECK(a < (unsigned)di && b < (unsigned)dj && (g[q] * di + a) * dj + b < (unsigned)cells);
                atomicAdd(&mine[(g[q] * di + a) * dj + b], 1u);
            }
        }
    }
    __syncthreads();
    if (replicas > 1) {
        for (int c = tid; c < cells; c += DT_THREADS) {
            uint32_t v = tab[c];
            for (int r = 1; r < replicas; ++r) v += tab[(size_t)r * cells + c];
            tab[c] = v;
        }
        __syncthreads();
    }
}

// disc_test with the level taken at run time; sS, sRadix (GPUCSL_MAX_LEVEL
// ints each) and sDimS are LDS scratch of the block.
__device__ __forceinline__ void disc_test_rt(const uint8_t *__restrict__ data, int64_t n, int64_t stride,
                                             const uint8_t *__restrict__ dims, int p, int L, double alpha,
                                             const int32_t *__restrict__ nbr, const int32_t *__restrict__ deg,
                                             const unsigned long long *__restrict__ testoff,
                                             float *__restrict__ pmax, unsigned long long *__restrict__ keys,
                                             uint32_t *gslice, unsigned long long t, uint32_t *lds, int *sS,
                                             int *sRadix, int *sDimS) {
    {
        int lo = 0, hi = p;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (testoff[mid] <= t) lo = mid; else hi = mid;
        }
        const int i = lo;
        const int d = deg[i];
        const unsigned long long per_pair = binom(d - 1, L);
        const unsigned long long o = t - testoff[i];
        const int jpos = (int)(o / per_pair);
        const unsigned long long rank = o - (unsigned long long)jpos * per_pair;
        const int32_t *row = nbr + (size_t)i * p;
        const int j = row[jpos];
        if (threadIdx.x == 0) {
            int pos[GPUCSL_MAX_LEVEL];
            unrank_lex_rt(d - 1, L, rank, pos);
            int dim_s = 1;
            for (int c = 0; c < L; ++c) {
                const int idx = pos[c] - 1;
                sS[c] = row[idx + (idx >= jpos)];
                sRadix[c] = dim_s;
                dim_s *= dims[sS[c]];
            }
            *sDimS = dim_s;
        }
        __syncthreads();
        const int dim_s = *sDimS;
        const int di = dims[i], dj = dims[j];
        const int cells = dim_s * di * dj;
        const int marg = dim_s * (di + dj + 1);
        uint32_t *tab;
        int replicas;
        if (4 * cells + marg <= DT_LDS_WORDS) {
            tab = lds;
            replicas = 4;
        } else if (cells + marg <= DT_LDS_WORDS) {
            tab = lds;
            replicas = 1;
        } else {
            tab = gslice;
            replicas = 1;
        }
        count_table_rt(data, n, stride, i, j, sS, sRadix, L, di, dj, tab, cells, replicas);
        uint32_t *Nis = tab + (size_t)replicas * cells, *Njs = Nis + dim_s * di, *Ns = Njs + dim_s * dj;
        marginals(tab, dim_s, di, dj, Nis, Njs, Ns);
        double chi2 = chi2_sum(tab, Nis, Njs, Ns, dim_s, di, dj, false);
        if (threadIdx.x < 64) {
            chi2 = __shfl(chi2, 0);
            const float pf = (float)pchisq_wave(chi2, (di - 1) * (dj - 1) * dim_s);
            const int a = i < j ? i : j, b = i < j ? j : i;
            if (threadIdx.x == 0) atomic_max_nonneg(&pmax[(size_t)a * p + b], pf);
            if (threadIdx.x == 0 && (double)pf >= alpha) {
                const unsigned long long key = ((unsigned long long)(i < j ? 0 : 1) << 62) | rank;
                atomicMin(&keys[(size_t)a * p + b], key);
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(DT_THREADS) void disc_level_rt_kernel(
    const uint8_t *__restrict__ data, int64_t n, int64_t stride, const uint8_t *__restrict__ dims, int p, int L,
    double alpha, const int32_t *__restrict__ nbr, const int32_t *__restrict__ deg,
    const unsigned long long *__restrict__ testoff, float *__restrict__ pmax,
    unsigned long long *__restrict__ keys, uint32_t *__restrict__ gtab, size_t gtab_words) {
    extern __shared__ uint32_t lds[];
    __shared__ int sS[GPUCSL_MAX_LEVEL], sRadix[GPUCSL_MAX_LEVEL];
    __shared__ int sDimS;
    const unsigned long long T = testoff[p];
    for (unsigned long long t = blockIdx.x; t < T; t += gridDim.x)
        disc_test_rt(data, n, stride, dims, p, L, alpha, nbr, deg, testoff, pmax, keys,
                     gtab + (size_t)blockIdx.x * gtab_words, t, lds, sS, sRadix, &sDimS);
}

__device__ __forceinline__ void disc_writeback_pair_rt(int p, int L, const int32_t *__restrict__ nbr,
                                                       const int32_t *__restrict__ deg,
                                                       unsigned long long *__restrict__ keys,
                                                       uint8_t *__restrict__ Gm, int32_t *__restrict__ sep, int maxL,
                                                       int a, int b) {
    if (b >= p || b <= a) return;
    const unsigned long long key = keys[(size_t)a * p + b];
    if (key == KEY_NONE) return;
    keys[(size_t)a * p + b] = KEY_NONE;
    const int side = (int)(key >> 62);
    const unsigned long long rank = key & ((1ull << 62) - 1);
    const int i = side ? b : a, j = side ? a : b;
    const int32_t *row = nbr + (size_t)i * p;
    const int d = deg[i];
    int lo = 0, hi = d;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (row[mid] < j) lo = mid + 1; else hi = mid;
    }
    const int jpos = lo;
    int pos[GPUCSL_MAX_LEVEL];
    unrank_lex_rt(d - 1, L, rank, pos);
    for (int c = 0; c < L; ++c) {
        const int idx = pos[c] - 1;
        const int v = row[idx + (idx >= jpos)];
        sep[((size_t)i * p + j) * maxL + c] = v;
        sep[((size_t)j * p + i) * maxL + c] = v;
    }
    Gm[(size_t)a * p + b] = 0;
    Gm[(size_t)b * p + a] = 0;
}

__global__ void disc_writeback_rt_kernel(int p, int L, const int32_t *__restrict__ nbr,
                                         const int32_t *__restrict__ deg, unsigned long long *__restrict__ keys,
                                         uint8_t *__restrict__ Gm, int32_t *__restrict__ sep, int maxL) {
    disc_writeback_pair_rt(p, L, nbr, deg, keys, Gm, sep, maxL, blockIdx.y, blockIdx.x * blockDim.x + threadIdx.x);
}

// Single test for known-answer checks: table (cells words), chi^2 and p.
template <int L>
__global__ __launch_bounds__(DT_THREADS) void disc_single_test_kernel(
    const uint8_t *__restrict__ data, int64_t n, int64_t stride, const uint8_t *__restrict__ dims, int vi,
    int vj, const int32_t *__restrict__ Sin, int sequential, uint32_t *__restrict__ table_out,
    double *__restrict__ out) {
    extern __shared__ uint32_t lds[];
    IArr<L> S, radix;
    int dim_s = 1;
    for (int c = 0; c < L; ++c) {
        S[c] = Sin[c];
        radix[c] = dim_s;
        dim_s *= dims[S[c]];
    }
    const int di = dims[vi], dj = dims[vj];
    const int cells = dim_s * di * dj;
    if (cells + dim_s * (di + dj + 1) > DT_LDS_WORDS) {
        // table and marginals must fit the workgroup's LDS: report, do not count
        if (threadIdx.x == 0) {
            out[0] = __builtin_nan("");
            out[1] = __builtin_nan("");
            out[2] = -1.0;
        }
        return;
    }
    count_table<L>(data, n, stride, vi, vj, S, di, dj, radix, lds, cells, 1);
    uint32_t *Nis = lds + cells, *Njs = Nis + dim_s * di, *Ns = Njs + dim_s * dj;
    marginals(lds, dim_s, di, dj, Nis, Njs, Ns);
    double chi2 = chi2_sum(lds, Nis, Njs, Ns, dim_s, di, dj, sequential != 0);
    for (int c = threadIdx.x; c < cells; c += DT_THREADS) table_out[c] = lds[c];
    if (threadIdx.x < 64) {
        chi2 = __shfl(chi2, 0);
        const double pv = pchisq_wave(chi2, (di - 1) * (dj - 1) * dim_s);
        if (threadIdx.x == 0) {
            out[0] = chi2;
            out[1] = pv;
            out[2] = (double)((di - 1) * (dj - 1) * dim_s);
        }
    }
}

// -1 on and below the diagonal (gpucsl/pc/helpers.py:96-101).
__global__ void lower_minus_one_kernel(float *__restrict__ pmax, int p) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (b < p && b <= a) pmax[(size_t)a * p + b] = -1.0f;
}

// ----------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------
int lower_minus_one_launch(float *pmax, int p, hipStream_t s) {
    if (!pmax || p < 1) return set_error(ERR_ARG, "pmax finalize: bad arguments");
    GPUCSL_LAUNCH("pmax_lower", lower_minus_one_kernel, dim3((p + 255) / 256, p), dim3(256), 0, s, pmax, p);
    return check_hip(hipGetLastError(), "pmax finalize launch");
}

static int disc_grid() { return device_cus() * DT_BLOCKS_PER_CU; }

// Largest table (+ marginals) one test can need, in words.
// Words of the largest contingency table plus marginals; saturates at
// TABLE_WORDS_LIMIT (tables that large can never be held).
constexpr size_t TABLE_WORDS_LIMIT = (size_t)1 << 48;
static size_t max_table_words(int max_dim, int level) {
    size_t ds = 1;
    for (int t = 0; t < level; ++t) {
        ds *= (size_t)max_dim;
        if (ds >= TABLE_WORDS_LIMIT) return TABLE_WORDS_LIMIT;
    }
    const size_t cells = (size_t)max_dim * max_dim * ds;
    if (cells >= TABLE_WORDS_LIMIT) return TABLE_WORDS_LIMIT;
    const size_t w = cells + ds * (2 * (size_t)max_dim + 1);
    return w >= TABLE_WORDS_LIMIT ? TABLE_WORDS_LIMIT : w;
}

static size_t gtab_words_for(int max_dim, int max_level) {
    size_t need = max_table_words(max_dim, max_level);
    return need <= (size_t)DT_LDS_WORDS ? 0 : align_up(need, 64);
}

// Workspace = a fixed part (level lists, test offsets, removal keys) plus one
// global table slice per concurrently running block when the largest table
// exceeds the LDS budget.  A smaller workspace runs fewer blocks at once.
static size_t disc_fixed_bytes(int p, int max_dim) {
    size_t off = align_up(gauss_workspace_bytes(p), 256);
    off = align_up(off + sizeof(unsigned long long) * (size_t)(p + 1), 256);
    off = align_up(off + sizeof(unsigned long long) * (size_t)p * p, 256);
    const size_t W = (size_t)gram_width(p, max_dim);
    off = align_up(off + sizeof(int32_t) * W * W, 256);
    return off;
}

size_t disc_workspace_bytes(int p, int max_level, int max_dim) {
    if (max_table_words(max_dim, max_level) >= TABLE_WORDS_LIMIT) return SIZE_MAX;
    return align_up(disc_fixed_bytes(p, max_dim) +
                        sizeof(uint32_t) * gtab_words_for(max_dim, max_level) * (size_t)disc_grid(),
                    256);
}

size_t disc_min_workspace_bytes(int p, int max_level, int max_dim) {
    if (max_table_words(max_dim, max_level) >= TABLE_WORDS_LIMIT) return SIZE_MAX;
    return align_up(disc_fixed_bytes(p, max_dim) + sizeof(uint32_t) * gtab_words_for(max_dim, max_level), 256);
}

// blocks that the workspace has table slices for (all of the grid when the
// tables fit in LDS)
static int disc_slices(size_t ws_bytes, int p, int max_level, int max_dim) {
    const size_t words = gtab_words_for(max_dim, max_level);
    if (words == 0) return disc_grid();
    const size_t fixed = disc_fixed_bytes(p, max_dim);
    const size_t n = ws_bytes > fixed ? (ws_bytes - fixed) / (sizeof(uint32_t) * words) : 0;
    return n < (size_t)disc_grid() ? (int)n : disc_grid();
}

static DiscWs disc_carve(void *ws, int p, int max_level, int max_dim, size_t ws_bytes) {
    DiscWs w;
    char *b = (char *)ws;
    size_t off = align_up(gauss_workspace_bytes(p), 256);
    w.testoff = (unsigned long long *)(b + off);
    off = align_up(off + sizeof(unsigned long long) * (size_t)(p + 1), 256);
    w.keys = (unsigned long long *)(b + off);
    off = align_up(off + sizeof(unsigned long long) * (size_t)p * p, 256);
    w.gram = (int32_t *)(b + off);
    w.gram_w = gram_width(p, max_dim);
    off = align_up(off + sizeof(int32_t) * (size_t)w.gram_w * w.gram_w, 256);
    w.gtab = (uint32_t *)(b + off);
    w.gtab_words = gtab_words_for(max_dim, max_level);
    w.grid = disc_slices(ws_bytes, p, max_level, max_dim);
    return w;
}

int disc_init_launch(int p, int max_level, int max_dim, uint8_t *Gm, int32_t *sep, float *pmax, void *ws,
                     unsigned long long *st, int nst, hipStream_t s) {
    if (p < 1 || max_level < 0 || !Gm || !pmax || !ws) return set_error(ERR_ARG, "disc_init: bad arguments");
    DiscWs w = disc_carve(ws, p, max_level, max_dim, disc_min_workspace_bytes(p, max_level, max_dim));
    const size_t pp = (size_t)p * p, ns = max_level > 0 && sep ? pp * max_level : 0;
    const size_t ng = (size_t)w.gram_w * w.gram_w;
    size_t blocks = ((ns > pp ? (ns > ng ? ns : ng) : (pp > ng ? pp : ng)) + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks * 256 < (size_t)nst) blocks = ((size_t)nst + 255) / 256;
    GPUCSL_LAUNCH("disc_init", disc_init_kernel, dim3((unsigned)blocks), dim3(256), 0, s, Gm,
                  max_level > 0 ? sep : nullptr, pmax, w.keys, ng ? w.gram : nullptr, st, pp, ns, ng, nst);
    return check_hip(hipGetLastError(), "disc_init launch");
}

int disc_level0_launch(const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int p,
                       int max_level, int max_dim, double alpha, uint8_t *Gm, float *pmax,
                       unsigned long long *st, void *ws, size_t ws_bytes, hipStream_t s) {
    if (!data || !dims || !Gm || !pmax || !st || !ws || p < 1 || n < 1)
        return set_error(ERR_ARG, "disc_level0: bad arguments");
    if (stride < (n + 15) / 16 * 16 || stride % 16 != 0)
        return set_error(ERR_ARG, "disc_level0: stride must be a multiple of 16 and >= n rounded up to 16");
    if (((uintptr_t)data & 15) != 0) return set_error(ERR_ARG, "disc_level0: data must be 16-byte aligned");
    if (ws_bytes < disc_min_workspace_bytes(p, max_level, max_dim))
        return set_error(ERR_WORKSPACE, "disc_level0: workspace too small");
    if (max_table_words(max_dim, 0) > (size_t)DT_LDS_WORDS && gtab_words_for(max_dim, max_level) < max_table_words(max_dim, 0))
        return set_error(ERR_WORKSPACE, "disc_level0: table workspace too small");
    DiscWs w = disc_carve(ws, p, max_level, max_dim, ws_bytes);
    if (w.gram_w) {
        // the tables of every pair from one int8 Gram matrix (zeroed by disc_init)
        const int W = w.gram_w, tiles = W / GRAM_TILE, ut = tiles * (tiles + 1) / 2;
        const int64_t steps = (n + 31) / 32;
        int64_t ks = (4 * (int64_t)device_cus() + ut - 1) / ut;
        const int64_t kmax = (steps + 63) / 64;  // at least 64 steps (2048 observations) per share
        if (ks > kmax) ks = kmax;
        if (ks < 1) ks = 1;
        GPUCSL_LAUNCH("disc_level0_gram", disc_onehot_gram_kernel, dim3(ut, (unsigned)ks), dim3(256), 0, s, data, n,
                      stride, p, max_dim, W, w.gram);
        const long long pairs = (long long)p * (p - 1) / 2;
        const unsigned blocks = (unsigned)((pairs + 255) / 256 > 0 ? (pairs + 255) / 256 : 1);
        if (max_dim <= 4)
            GPUCSL_LAUNCH("disc_level0", disc_level0_chi2_kernel<4>, dim3(blocks), dim3(256), 0, s, dims, p, max_dim,
                          W, w.gram, alpha, Gm, pmax, st);
        else if (max_dim <= 8)
            GPUCSL_LAUNCH("disc_level0", disc_level0_chi2_kernel<8>, dim3(blocks), dim3(256), 0, s, dims, p, max_dim,
                          W, w.gram, alpha, Gm, pmax, st);
        else
            GPUCSL_LAUNCH("disc_level0", disc_level0_chi2_kernel<16>, dim3(blocks), dim3(256), 0, s, dims, p,
                          max_dim, W, w.gram, alpha, Gm, pmax, st);
        return check_hip(hipGetLastError(), "disc_level0 launch");
    }
    GPUCSL_LAUNCH("disc_level0", disc_level0_kernel, dim3(w.grid), dim3(DT_THREADS), DT_LDS_BYTES, s, data, n,
                  stride, dims, p, alpha, Gm, pmax, w.gtab, w.gtab_words, st);
    return check_hip(hipGetLastError(), "disc_level0 launch");
}

static const char *const kDiscName[] = {"", "disc_level1", "disc_level2", "disc_level3", "disc_level4",
                                        "disc_level5", "disc_level6", "disc_level7", "disc_level8"};

template <int L>
static void disc_level_launch_t(const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int p,
                                int maxL, double alpha, uint8_t *Gm, int32_t *sep, float *pmax,
                                const DiscWs &w, int32_t *nbr, int32_t *deg, hipStream_t s) {
    GPUCSL_LAUNCH(kDiscName[L], disc_level_kernel<L>, dim3(w.grid), dim3(DT_THREADS), DT_LDS_BYTES, s, data, n,
                  stride, dims, p, alpha, nbr, deg, w.testoff, pmax, w.keys, w.gtab, w.gtab_words);
    GPUCSL_LAUNCH("disc_writeback", disc_writeback_kernel<L>, dim3((p + 255) / 256, p), dim3(256), 0, s, p, nbr,
                  deg, w.keys, Gm, sep, maxL);
}

int disc_level_launch(int level, const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int p,
                      int max_level, int max_dim, double alpha, uint8_t *Gm, int32_t *sep, float *pmax,
                      unsigned long long *st, void *ws, size_t ws_bytes, hipStream_t s) {
    if (level < 1 || level > max_level) return set_error(ERR_ARG, "disc_level: level %d outside [1, %d]", level, max_level);
    if (level > GPUCSL_MAX_LEVEL)
        return set_error(ERR_UNSUPPORTED, "disc_level: levels above %d are not built", GPUCSL_MAX_LEVEL);
    if (!data || !dims || !Gm || !sep || !pmax || !st || !ws) return set_error(ERR_ARG, "disc_level: null pointer");
    if (stride < (n + 15) / 16 * 16 || stride % 16 != 0)
        return set_error(ERR_ARG, "disc_level: stride must be a multiple of 16 and >= n rounded up to 16");
    if (ws_bytes < disc_min_workspace_bytes(p, max_level, max_dim))
        return set_error(ERR_WORKSPACE, "disc_level: workspace too small");
    int rc = compact_launch(Gm, p, level, ws, nullptr, s);
    if (rc) return rc;
    int32_t *nbr, *deg;
    level_list_views(ws, p, &nbr, &deg);
    DiscWs w = disc_carve(ws, p, max_level, max_dim, ws_bytes);
    GPUCSL_LAUNCH("disc_testoff", disc_testoff_kernel, dim3(1), dim3(1024), 0, s, deg, p, level, w.testoff, st);
    switch (level) {
#define GPUCSL_D(LV)                                                                                   \
    case LV:                                                                                           \
        disc_level_launch_t<LV>(data, n, stride, dims, p, max_level, alpha, Gm, sep, pmax, w, nbr, deg, s); \
        break;
        GPUCSL_D(1) GPUCSL_D(2) GPUCSL_D(3) GPUCSL_D(4) GPUCSL_D(5) GPUCSL_D(6) GPUCSL_D(7) GPUCSL_D(8)
#undef GPUCSL_D
    default:
        GPUCSL_LAUNCH("disc_level_rt", disc_level_rt_kernel, dim3(w.grid), dim3(DT_THREADS), DT_LDS_BYTES, s, data,
                      n, stride, dims, p, level, alpha, nbr, deg, w.testoff, pmax, w.keys, w.gtab, w.gtab_words);
        GPUCSL_LAUNCH("disc_writeback", disc_writeback_rt_kernel, dim3((p + 255) / 256, p), dim3(256), 0, s, p,
                      level, nbr, deg, w.keys, Gm, sep, max_level);
        break;
    }
    return check_hip(hipGetLastError(), "disc_level launch");
}

// ----------------------------------------------------------------------------
// batched jobs: the jobs of a launch travel in its arguments (a captured
// launch keeps them); per job the same kernels' arithmetic as the
// single-dataset path, on the job's own buffers and workspace, so the results
// are bit-identical.
// ----------------------------------------------------------------------------
struct DiscEntry {
    const uint8_t *data, *dims;
    int64_t n, stride;
    int p, max_level, max_dim, gram_w;
    int ks, grid;  // level-0 Gram shares; workgroups the job's table slices allow
    double alpha;
    uint8_t *G;
    int32_t *sep;
    float *pmax;
    unsigned long long *stats;
    uint16_t *skel;
    int32_t *nbr, *deg;
    unsigned long long *testoff, *keys;
    int32_t *gram;
    uint32_t *gtab;
    size_t gtab_words;
};
struct DiscLaunch {
    int count;
    DiscEntry e[GPUCSL_DISC_BATCH_JOBS];
};

__global__ void disc_batch_init_kernel(DiscLaunch B) {
    const DiscEntry &e = B.e[blockIdx.y];
    const size_t pp = (size_t)e.p * e.p, ns = e.max_level > 0 && e.sep ? pp * e.max_level : 0;
    const size_t ng = (size_t)e.gram_w * e.gram_w;
    disc_init_body(e.G, e.max_level > 0 ? e.sep : nullptr, e.pmax, e.keys, ng ? e.gram : nullptr, e.stats, pp, ns,
                   ng, (e.max_level + 1) * GPUCSL_STAT_SLOTS, (size_t)blockIdx.x * blockDim.x + threadIdx.x,
                   (size_t)gridDim.x * blockDim.x);
}

__global__ __launch_bounds__(256) void disc_batch_gram_kernel(DiscLaunch B) {
    const DiscEntry &e = B.e[blockIdx.z];
    const int tiles = e.gram_w / GRAM_TILE;
    if (e.gram_w == 0 || (int)blockIdx.x >= tiles * (tiles + 1) / 2 || (int)blockIdx.y >= e.ks) return;
    disc_gram_body(e.data, e.n, e.stride, e.p, e.max_dim, e.gram_w, e.gram, blockIdx.x, blockIdx.y, e.ks);
}

template <int MD>
__global__ __launch_bounds__(256) void disc_batch_level0_chi2_kernel(DiscLaunch B) {
    const DiscEntry &e = B.e[blockIdx.z];
    const long long pairs = (long long)e.p * (e.p - 1) / 2;
    if (e.gram_w == 0 || (long long)blockIdx.x * 256 >= (pairs > 0 ? pairs : 1)) return;
    disc_level0_chi2_body<MD>(e.dims, e.p, e.max_dim, e.gram_w, e.gram, e.alpha, e.G, e.pmax, e.stats, blockIdx.x);
}

// Level-start neighbour lists of every job at `level`: one wave per row,
// ascending neighbours (diagonal excluded) and the degree, as the
// single-dataset compaction leaves them.
__global__ __launch_bounds__(256) void disc_batch_compact_kernel(DiscLaunch B, int level) {
    const DiscEntry &e = B.e[blockIdx.y];
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (level > e.max_level || row >= e.p) return;
    const int p = e.p, lane = __lane_id();
    const uint8_t *g = e.G + (size_t)row * p;
    int32_t *out = e.nbr + (size_t)row * p;
    int base = 0;
    for (int c0 = 0; c0 < p; c0 += 64) {
        const int c = c0 + lane;
        const bool on = c < p && c != row && g[c] != 0;
        const uint64_t m = __ballot(on);
        if (on) out[base + __popcll(m & lanemask_lt())] = c;
        base += __popcll(m);
    }
    if (lane == 0) e.deg[row] = base;
}

__global__ __launch_bounds__(1024) void disc_batch_testoff_kernel(DiscLaunch B, int level) {
    const DiscEntry &e = B.e[blockIdx.x];
    if (level > e.max_level) return;
    disc_testoff_body(e.deg, e.p, level, e.testoff, e.stats + (size_t)level * GPUCSL_STAT_SLOTS);
}

// The tests of every job at this level, numbered one after the other in job
// order; workgroup w runs the tests whose number is w modulo the grid.
template <int L>
__global__ __launch_bounds__(DT_THREADS) void disc_batch_level_kernel(DiscLaunch B) {
    extern __shared__ uint32_t lds[];
    const unsigned long long grid = gridDim.x;
    unsigned long long base = 0;
    for (int k = 0; k < B.count; ++k) {
        const DiscEntry &e = B.e[k];
        if (L > e.max_level) continue;
        const unsigned long long T = e.testoff[e.p];
        for (unsigned long long t = (blockIdx.x + grid - base % grid) % grid; t < T; t += grid)
            disc_test<L>(e.data, e.n, e.stride, e.dims, e.p, e.alpha, e.nbr, e.deg, e.testoff, e.pmax, e.keys,
                         e.gtab + (size_t)blockIdx.x * e.gtab_words, t, lds);
        base += T;
    }
}

__global__ __launch_bounds__(DT_THREADS) void disc_batch_level_rt_kernel(DiscLaunch B, int L) {
    extern __shared__ uint32_t lds[];
    __shared__ int sS[GPUCSL_MAX_LEVEL], sRadix[GPUCSL_MAX_LEVEL];
    __shared__ int sDimS;
    const unsigned long long grid = gridDim.x;
    unsigned long long base = 0;
    for (int k = 0; k < B.count; ++k) {
        const DiscEntry &e = B.e[k];
        if (L > e.max_level) continue;
        const unsigned long long T = e.testoff[e.p];
        for (unsigned long long t = (blockIdx.x + grid - base % grid) % grid; t < T; t += grid)
            disc_test_rt(e.data, e.n, e.stride, e.dims, e.p, L, e.alpha, e.nbr, e.deg, e.testoff, e.pmax, e.keys,
                         e.gtab + (size_t)blockIdx.x * e.gtab_words, t, lds, sS, sRadix, &sDimS);
        base += T;
    }
}

template <int L>
__global__ void disc_batch_writeback_kernel(DiscLaunch B) {
    const DiscEntry &e = B.e[blockIdx.z];
    if (L > e.max_level || (int)blockIdx.y >= e.p) return;
    disc_writeback_pair<L>(e.p, e.nbr, e.deg, e.keys, e.G, e.sep, e.max_level, blockIdx.y,
                           blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void disc_batch_writeback_rt_kernel(DiscLaunch B, int L) {
    const DiscEntry &e = B.e[blockIdx.z];
    if (L > e.max_level || (int)blockIdx.y >= e.p) return;
    disc_writeback_pair_rt(e.p, L, e.nbr, e.deg, e.keys, e.G, e.sep, e.max_level, blockIdx.y,
                           blockIdx.x * blockDim.x + threadIdx.x);
}

// pmax = -1 on and below the diagonal, and the uint16 skeleton when asked for.
__global__ void disc_batch_finish_kernel(DiscLaunch B) {
    const DiscEntry &e = B.e[blockIdx.z];
    const int a = blockIdx.y, b = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= e.p || b >= e.p) return;
    const size_t at = (size_t)a * e.p + b;
    if (b <= a) e.pmax[at] = -1.0f;
    if (e.skel) e.skel[at] = e.G[at];
}

// Entry of one job, checked as the single-dataset calls check their arguments.
static int disc_entry(const gpucsl_disc_job &J, int idx, DiscEntry &e) {
    if (!J.data || !J.dims || !J.G || !J.pmax || !J.stats || !J.ws || J.p < 1 || J.n < 1 || J.max_level < 0 ||
        (J.max_level > 0 && !J.sepsets))
        return set_error(ERR_ARG, "disc_batch: job %d has bad arguments", idx);
    if (J.stride < (J.n + 15) / 16 * 16 || J.stride % 16 != 0)
        return set_error(ERR_ARG, "disc_batch: job %d: stride must be a multiple of 16 and >= n rounded up to 16",
                         idx);
    if (((uintptr_t)J.data & 15) != 0) return set_error(ERR_ARG, "disc_batch: job %d: data must be 16-byte aligned", idx);
    const int md = J.max_dim < 1 ? 1 : J.max_dim;
    if (J.ws_bytes < disc_min_workspace_bytes(J.p, J.max_level, md))
        return set_error(ERR_WORKSPACE, "disc_batch: job %d: workspace too small", idx);
    DiscWs w = disc_carve(J.ws, J.p, J.max_level, md, J.ws_bytes);
    e = DiscEntry{};
    e.data = J.data;
    e.dims = J.dims;
    e.n = J.n;
    e.stride = J.stride;
    e.p = J.p;
    e.max_level = J.max_level;
    e.max_dim = md;
    e.gram_w = w.gram_w;
    e.grid = w.grid < 1 ? 1 : w.grid;
    e.alpha = J.alpha;
    e.G = J.G;
    e.sep = J.sepsets;
    e.pmax = J.pmax;
    e.stats = reinterpret_cast<unsigned long long *>(J.stats);
    e.skel = J.skeleton;
    level_list_views(J.ws, J.p, &e.nbr, &e.deg);
    e.testoff = w.testoff;
    e.keys = w.keys;
    e.gram = w.gram;
    e.gtab = w.gtab;
    e.gtab_words = w.gtab_words;
    if (w.gram_w) {
        const int tiles = w.gram_w / GRAM_TILE, ut = tiles * (tiles + 1) / 2;
        const int64_t steps = (J.n + 31) / 32;
        int64_t ks = (4 * (int64_t)device_cus() + ut - 1) / ut;
        const int64_t kmax = (steps + 63) / 64;
        if (ks > kmax) ks = kmax;
        e.ks = ks < 1 ? 1 : (int)ks;
    }
    return OK;
}

// Runs f(launch, first job index) for the jobs in launches of at most
// GPUCSL_DISC_BATCH_JOBS.
template <class F>
static int disc_batch_chunks(const gpucsl_disc_job *jobs, int nj, F f) {
    if (nj < 0 || (nj > 0 && !jobs)) return set_error(ERR_ARG, "disc_batch: bad job table");
    for (int c0 = 0; c0 < nj; c0 += GPUCSL_DISC_BATCH_JOBS) {
        DiscLaunch B{};
        B.count = nj - c0 < GPUCSL_DISC_BATCH_JOBS ? nj - c0 : GPUCSL_DISC_BATCH_JOBS;
        for (int k = 0; k < B.count; ++k) {
            const int rc = disc_entry(jobs[c0 + k], c0 + k, B.e[k]);
            if (rc) return rc;
        }
        const int rc = f(B, c0);
        if (rc) return rc;
    }
    return OK;
}

int disc_batch_init_launch(const gpucsl_disc_job *jobs, int nj, hipStream_t s) {
    return disc_batch_chunks(jobs, nj, [&](const DiscLaunch &B, int) -> int {
        size_t most = 1;
        for (int k = 0; k < B.count; ++k) {
            const DiscEntry &e = B.e[k];
            const size_t pp = (size_t)e.p * e.p, ns = e.max_level > 0 ? pp * e.max_level : 0;
            const size_t ng = (size_t)e.gram_w * e.gram_w, nst = (size_t)(e.max_level + 1) * GPUCSL_STAT_SLOTS;
            for (size_t v : {pp, ns, ng, nst}) most = v > most ? v : most;
        }
        size_t blocks = (most + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        GPUCSL_LAUNCH("disc_init", disc_batch_init_kernel, dim3((unsigned)blocks, B.count), dim3(256), 0, s, B);
        return check_hip(hipGetLastError(), "disc_batch_init launch");
    });
}

int disc_batch_level0_launch(const gpucsl_disc_job *jobs, int nj, hipStream_t s) {
    return disc_batch_chunks(jobs, nj, [&](const DiscLaunch &B, int c0) -> int {
        int max_ut = 0, max_ks = 0, max_md = 1;
        long long max_pairs = 1;
        for (int k = 0; k < B.count; ++k) {
            const DiscEntry &e = B.e[k];
            if (!e.gram_w) {
                // tables counted pair by pair, as the single-dataset path does
                const gpucsl_disc_job &J = jobs[c0 + k];
                const int rc = disc_level0_launch(J.data, J.n, J.stride, J.dims, J.p, J.max_level, e.max_dim,
                                                  J.alpha, J.G, J.pmax, e.stats, J.ws, J.ws_bytes, s);
                if (rc) return rc;
                continue;
            }
            const int tiles = e.gram_w / GRAM_TILE, ut = tiles * (tiles + 1) / 2;
            max_ut = ut > max_ut ? ut : max_ut;
            max_ks = e.ks > max_ks ? e.ks : max_ks;
            max_md = e.max_dim > max_md ? e.max_dim : max_md;
            const long long pairs = (long long)e.p * (e.p - 1) / 2;
            max_pairs = pairs > max_pairs ? pairs : max_pairs;
        }
        if (max_ut == 0) return OK;
        GPUCSL_LAUNCH("disc_level0_gram", disc_batch_gram_kernel, dim3(max_ut, max_ks, B.count), dim3(256), 0, s, B);
        const dim3 grid((unsigned)((max_pairs + 255) / 256), 1, B.count);
        if (max_md <= 4)
            GPUCSL_LAUNCH("disc_level0", disc_batch_level0_chi2_kernel<4>, grid, dim3(256), 0, s, B);
        else if (max_md <= 8)
            GPUCSL_LAUNCH("disc_level0", disc_batch_level0_chi2_kernel<8>, grid, dim3(256), 0, s, B);
        else
            GPUCSL_LAUNCH("disc_level0", disc_batch_level0_chi2_kernel<16>, grid, dim3(256), 0, s, B);
        return check_hip(hipGetLastError(), "disc_batch_level0 launch");
    });
}

int disc_batch_level_launch(int level, const gpucsl_disc_job *jobs, int nj, hipStream_t s) {
    if (level < 1) return set_error(ERR_ARG, "disc_batch_level: level %d < 1", level);
    return disc_batch_chunks(jobs, nj, [&](const DiscLaunch &B, int) -> int {
        int maxp = 0, grid = disc_grid();
        for (int k = 0; k < B.count; ++k) {
            const DiscEntry &e = B.e[k];
            if (level > e.max_level) continue;
            maxp = e.p > maxp ? e.p : maxp;
            grid = e.grid < grid ? e.grid : grid;
        }
        if (maxp == 0) return OK;  // no job of this launch runs the level
        if (level > GPUCSL_MAX_LEVEL)
            return set_error(ERR_UNSUPPORTED, "disc_batch_level: levels above %d are not built", GPUCSL_MAX_LEVEL);
        GPUCSL_LAUNCH("compact", disc_batch_compact_kernel, dim3((maxp + 3) / 4, B.count), dim3(256), 0, s, B, level);
        GPUCSL_LAUNCH("disc_testoff", disc_batch_testoff_kernel, dim3(B.count), dim3(1024), 0, s, B, level);
        const dim3 wb((maxp + 255) / 256, maxp, B.count);
        switch (level) {
#define GPUCSL_DB(LV)                                                                                             \
    case LV:                                                                                                      \
        GPUCSL_LAUNCH(kDiscName[LV], disc_batch_level_kernel<LV>, dim3(grid), dim3(DT_THREADS), DT_LDS_BYTES, s, B); \
        GPUCSL_LAUNCH("disc_writeback", disc_batch_writeback_kernel<LV>, wb, dim3(256), 0, s, B);                 \
        break;
            GPUCSL_DB(1) GPUCSL_DB(2) GPUCSL_DB(3) GPUCSL_DB(4) GPUCSL_DB(5) GPUCSL_DB(6) GPUCSL_DB(7) GPUCSL_DB(8)
#undef GPUCSL_DB
        default:
            GPUCSL_LAUNCH("disc_level_rt", disc_batch_level_rt_kernel, dim3(grid), dim3(DT_THREADS), DT_LDS_BYTES, s,
                          B, level);
            GPUCSL_LAUNCH("disc_writeback", disc_batch_writeback_rt_kernel, wb, dim3(256), 0, s, B, level);
            break;
        }
        return check_hip(hipGetLastError(), "disc_batch_level launch");
    });
}

int disc_batch_finish_launch(const gpucsl_disc_job *jobs, int nj, hipStream_t s) {
    return disc_batch_chunks(jobs, nj, [&](const DiscLaunch &B, int) -> int {
        int maxp = 1;
        for (int k = 0; k < B.count; ++k) maxp = B.e[k].p > maxp ? B.e[k].p : maxp;
        GPUCSL_LAUNCH("disc_finish", disc_batch_finish_kernel, dim3((maxp + 255) / 256, maxp, B.count), dim3(256), 0,
                      s, B);
        return check_hip(hipGetLastError(), "disc_batch_finish launch");
    });
}

// ----------------------------------------------------------------------------
// ordinal encoding on the device (gpucsl/pc/discover_skeleton_discrete.py:38-39,
// sklearn's OrdinalEncoder): a value's code is its rank among the column's
// sorted distinct values.  One block per column: the distinct values are found
// in increasing order, each as the block-wide minimum of the values above the
// previous one (one pass over the column per category), kept in LDS, and every
// value is then coded by bisection.  At most 255 categories; NaN is reported.
// ----------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void disc_encode_kernel(const T *__restrict__ values, int64_t n, int64_t ld,
                                                          int row_major, uint8_t *__restrict__ codes, int64_t stride,
                                                          uint8_t *__restrict__ dims, int32_t *__restrict__ status) {
    const int v = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ T s_val[256];
    __shared__ T s_m[4];
    __shared__ int s_h[4], s_nan[4];
    __shared__ int s_more;
    auto at = [&](int64_t k) -> T { return row_major ? values[k * ld + v] : values[(size_t)v * ld + k]; };
    int cnt = 0, any_nan = 0;
    T prev = T(0);
    for (;;) {
        T m = T(0);
        int has = 0;
        for (int64_t k = tid; k < n; k += 256) {
            const T x = at(k);
            if (!(x == x)) {
                any_nan = 1;
                continue;
            }
            if ((cnt == 0 || x > prev) && (!has || x < m)) {
                m = x;
                has = 1;
            }
        }
        for (int off = 32; off >= 1; off >>= 1) {
            const T om = __shfl_xor(m, off);
            const int oh = __shfl_xor(has, off);
            if (oh && (!has || om < m)) {
                m = om;
                has = 1;
            }
        }
        const int wn = __any(any_nan) ? 1 : 0;
        if (lane == 0) {
            s_m[wave] = m;
            s_h[wave] = has;
            s_nan[wave] = wn;
        }
        __syncthreads();
        if (tid == 0) {
            T bm = T(0);
            int bh = 0;
            for (int w = 0; w < 4; ++w)
                if (s_h[w] && (!bh || s_m[w] < bm)) {
                    bm = s_m[w];
                    bh = 1;
                }
            s_more = bh;
            if (bh && cnt < 255) s_val[cnt] = bm;
            s_nan[0] = s_nan[0] | s_nan[1] | s_nan[2] | s_nan[3];
        }
        __syncthreads();
        any_nan = s_nan[0];
        if (!s_more) break;
        if (cnt == 255) {  // a 256th distinct value
            cnt = 256;
            break;
        }
        prev = s_val[cnt];
        ++cnt;
        __syncthreads();  // s_m / s_h / s_more are reused
    }
    if (tid == 0) {
        dims[v] = cnt <= 255 ? (uint8_t)cnt : 0;
        status[v] = cnt > 255 ? 1 : (any_nan ? 2 : 0);
    }
    if (cnt > 255 || any_nan) return;
    for (int64_t k = tid; k < stride; k += 256) {
        uint8_t code = 0;
        if (k < n) {
            const T x = at(k);
            int lo = 0, hi = cnt;  // s_val[lo] <= x < s_val[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_val[mid] <= x) lo = mid; else hi = mid;
            }
            code = (uint8_t)lo;
        }
        codes[(size_t)v * stride + k] = code;
    }
}

int disc_encode_launch(const void *values, int dtype, int64_t n, int p, int64_t ld, int row_major, uint8_t *codes,
                       int64_t stride, uint8_t *dims, int32_t *status, hipStream_t s) {
    if (!values || !codes || !dims || !status || n < 1 || p < 1 || stride < n || stride % 16 != 0)
        return set_error(ERR_ARG, "disc_encode: bad arguments");
    if (ld < (row_major ? p : n)) return set_error(ERR_ARG, "disc_encode: leading dimension too small");
    switch (dtype) {
    case 0:
        GPUCSL_LAUNCH("disc_encode", disc_encode_kernel<int32_t>, dim3(p), dim3(256), 0, s,
                      (const int32_t *)values, n, ld, row_major, codes, stride, dims, status);
        break;
    case 1:
        GPUCSL_LAUNCH("disc_encode", disc_encode_kernel<long long>, dim3(p), dim3(256), 0, s,
                      (const long long *)values, n, ld, row_major, codes, stride, dims, status);
        break;
    case 2:
        GPUCSL_LAUNCH("disc_encode", disc_encode_kernel<float>, dim3(p), dim3(256), 0, s, (const float *)values,
                      n, ld, row_major, codes, stride, dims, status);
        break;
    case 3:
        GPUCSL_LAUNCH("disc_encode", disc_encode_kernel<double>, dim3(p), dim3(256), 0, s, (const double *)values,
                      n, ld, row_major, codes, stride, dims, status);
        break;
    default:
        return set_error(ERR_ARG, "disc_encode: dtype must be 0 (int32), 1 (int64), 2 (float32) or 3 (float64)");
    }
    return check_hip(hipGetLastError(), "disc_encode launch");
}

int disc_single_test_launch(const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int vi, int vj,
                            const int32_t *S, int l, int sequential, uint32_t *table_out, double *out,
                            hipStream_t s) {
    if (l < 0 || l > 4) return set_error(ERR_ARG, "disc_ci_test: level must be in 0..4");
    switch (l) {
    case 0: hipLaunchKernelGGL(disc_single_test_kernel<0>, dim3(1), dim3(DT_THREADS), DT_LDS_BYTES, s, data, n, stride, dims, vi, vj, S, sequential, table_out, out); break;
    case 1: hipLaunchKernelGGL(disc_single_test_kernel<1>, dim3(1), dim3(DT_THREADS), DT_LDS_BYTES, s, data, n, stride, dims, vi, vj, S, sequential, table_out, out); break;
    case 2: hipLaunchKernelGGL(disc_single_test_kernel<2>, dim3(1), dim3(DT_THREADS), DT_LDS_BYTES, s, data, n, stride, dims, vi, vj, S, sequential, table_out, out); break;
    case 3: hipLaunchKernelGGL(disc_single_test_kernel<3>, dim3(1), dim3(DT_THREADS), DT_LDS_BYTES, s, data, n, stride, dims, vi, vj, S, sequential, table_out, out); break;
    case 4: hipLaunchKernelGGL(disc_single_test_kernel<4>, dim3(1), dim3(DT_THREADS), DT_LDS_BYTES, s, data, n, stride, dims, vi, vj, S, sequential, table_out, out); break;
    }
    return check_hip(hipGetLastError(), "disc_ci_test launch");
}

}  // namespace gpucsl
