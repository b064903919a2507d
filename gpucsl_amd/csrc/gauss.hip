// Gaussian PC-stable skeleton on gfx950.
//
// Replaces gpucsl/pc/cuda/gaussian_ci.cu (level 0 and level n), the compaction
// of gpucsl/pc/cuda/helpers/graph_helpers.cu:35-56 and the host level loop of
// gpucsl/pc/discover_skeleton_gaussian.py:30-113.
//
// Layout in HBM (p variables, L = max level):
//   C32  p*p float   correlation, FP32 (read by the CI tests)
//   C64  p*p double  correlation, FP64 (read only by rechecks)
//   G    p*p uint8   adjacency (1 edge, 0 none), kept symmetric
//   sep  p*p*L int32 separation sets, -1 padded, [(i*p+j)*L + k]
//   zmin p*p float   min Fisher-z per edge, upper triangle [min*p+max]
//   workspace: nbr p*p int32 (level-start neighbour lists), deg/cntA/cntB p
//              int32, rowoffA/rowoffB (p+1) int32, two work counters.
//
// A level l >= 1 runs as compaction -> scan -> phase A -> phase B.  Phase A
// tests every level-start edge (a, b), a < b, with S drawn from adj(a)\{b};
// phase B tests (b, a) with S from adj(b)\{a} for the edges phase A kept.
// Lane groups walk a pair's subset ranks in increasing order and stop at the
// first passing test, so the removing sepset is the lexicographically first
// one in pcalg order and zmin covers exactly the tests up to it.  No host sync
// happens between levels: a level whose work list is empty costs a few
// near-empty launches.
#include "common.h"

namespace gpucsl {

constexpr float EPS32 = 5.9604645e-8f;  // 2^-24
constexpr int PHASE_BLOCKS_PER_CU = 8;
constexpr int MAX_PHASE_BLOCKS = 4096;
constexpr int STRIPES = 8;
// counters: 8 stripes for each of FAST A, SLOW A, FAST B, SLOW B, then the two
// queue lengths.
constexpr int NUM_COUNTERS = 64;
constexpr int CNT_QUEUE = 32;
// statistics of the level-l kernels: 64 striped accumulators of 8 words each
// (block b adds into stripe b % 64), zeroed by scan_kernel, summed by
// level_stats_kernel
constexpr int BSTAT_WORDS = 8;
constexpr int BSTAT_STRIPES = 64;
enum BStat { BS_EXEC = 0, BS_USEFUL = 1, BS_RECHECK = 2, BS_FLIP = 3, BS_MAXRATIO = 4 };
// row-blocked FAST kernel: one wave per block, a fixed LDS budget, and work
// items of up to PAIRS_PER_ITEM ordered pairs of one row
constexpr int RB_BLOCKS_PER_CU = 16;
constexpr int RB_LDS_BYTES = 16384;
constexpr int PAIRS_PER_ITEM = 4;

// per-level statistics slots (uint64 each)
enum StatSlot { ST_EXEC = 0, ST_TOTAL = 1, ST_RECHECK = 2, ST_FLIP = 3, ST_MAXRATIO = 4, ST_MAXDEG = 5, ST_USEFUL = 6, ST_SLOTS = 8 };

// ----------------------------------------------------------------------------
// level 0: gpucsl/pc/cuda/gaussian_ci.cu:18-51
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void gauss_level0_kernel(const float *__restrict__ C32,
                                                           const double *__restrict__ C64, int p,
                                                           float thr32, double thr, float margin_k,
                                                           uint8_t *__restrict__ Gm,
                                                           float *__restrict__ zmin,
                                                           unsigned long long *__restrict__ st) {
    const int a = blockIdx.y;
    const int b = blockIdx.x * 256 + threadIdx.x;
    unsigned long long n_re = 0, n_flip = 0;
    float mr = 0.0f;
    if (b < p) {
        if (b == a) {
            Gm[(size_t)a * p + a] = 0;
        } else if (b > a) {
            const float r = C32[(size_t)a * p + b];
            const float z = fabsf(atanhf(r));
            const float err = EPS32 * (fabsf(r) + z + 1e-30f) / fmaxf(1.0f - r * r, 1e-30f);
            const bool bad = !(z <= 3.0e38f);
            float zf;
            bool pass;
            if (bad || fabsf(z - thr32) <= margin_k * err) {
                const double z64 = ref64::gauss_z0(C64[(size_t)a * p + b]);
                pass = z64 < thr;
                zf = (float)z64;
                ++n_re;
                if (!bad) {
                    n_flip += (pass != (z < thr32));
                    mr = fmaxf(mr, fabsf(z - zf) / err);
                }
            } else {
                pass = z < thr32;
                zf = z;
            }
            const uint8_t keep = pass ? 0 : 1;
            Gm[(size_t)a * p + b] = keep;
            Gm[(size_t)b * p + a] = keep;
            if (!(zf != zf)) zmin[(size_t)a * p + b] = zf;
        }
    }
    // executed = total = useful = p(p-1)/2 is known; only the rare recheck
    // counters need atomics.
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        const unsigned long long pairs = (unsigned long long)p * (p - 1) / 2;
        st[ST_EXEC] = pairs;
        st[ST_TOTAL] = pairs;
        st[ST_USEFUL] = pairs;
    }
    const unsigned long long any = __ballot(n_re != 0);
    if (any) {
        for (int off = 32; off >= 1; off >>= 1) {
            n_re += __shfl_xor(n_re, off);
            n_flip += __shfl_xor(n_flip, off);
            mr = fmaxf(mr, __shfl_xor(mr, off));
        }
        if (__lane_id() == 0) {
            atomicAdd(&st[ST_RECHECK], n_re);
            if (n_flip) atomicAdd(&st[ST_FLIP], n_flip);
            if (mr > 0.0f) atomicMax(&st[ST_MAXRATIO], (unsigned long long)__float_as_uint(mr));
        }
    }
}

// ----------------------------------------------------------------------------
// compaction: one wave per row (gpucsl/pc/cuda/helpers/graph_helpers.cu:35-56)
// nbr[row] = ascending neighbours (diagonal excluded), deg, and the split into
// neighbours below (cntB) and above (cntA) the row index.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void compact_kernel(const uint8_t *__restrict__ Gm, int p,
                                                      int32_t *__restrict__ nbr,
                                                      int32_t *__restrict__ deg,
                                                      int32_t *__restrict__ cntA,
                                                      int32_t *__restrict__ cntB) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= p) return;
    const int lane = __lane_id();
    const uint8_t *g = Gm + (size_t)row * p;
    int32_t *out = nbr + (size_t)row * p;
    int base = 0, below = 0;
    for (int c0 = 0; c0 < p; c0 += 64) {
        const int c = c0 + lane;
        const bool e = c < p && c != row && g[c] != 0;
        const uint64_t m = __ballot(e);
        if (e) out[base + __popcll(m & lanemask_lt())] = c;
        base += __popcll(m);
        if (c0 + 64 <= row)
            below += __popcll(m);
        else if (c0 < row)
            below += __popcll(m & ((row - c0) >= 64 ? ~0ull : ((1ull << (row - c0)) - 1)));
    }
    if (lane == 0) {
        deg[row] = base;
        cntB[row] = below;
        cntA[row] = base - below;
    }
}

// Exclusive scans of cntA/cntB into rowoffA/rowoffB (p+1 entries each), the
// level's full enumeration size T_l and max degree, and work-counter reset.
// One block of 1024 threads.
__global__ __launch_bounds__(1024) void scan_kernel(const int32_t *__restrict__ deg,
                                                    const int32_t *__restrict__ cntA,
                                                    const int32_t *__restrict__ cntB, int p,
                                                    int level, int32_t *__restrict__ rowoffA,
                                                    int32_t *__restrict__ rowoffB,
                                                    int32_t *__restrict__ itemoffA,
                                                    int32_t *__restrict__ itemoffB,
                                                    unsigned int *__restrict__ counters,
                                                    unsigned long long *__restrict__ bstats,
                                                    unsigned long long *__restrict__ st) {
    __shared__ int sA[1024], sB[1024], sIA[1024], sIB[1024];
    if (bstats && threadIdx.x < BSTAT_STRIPES * BSTAT_WORDS) bstats[threadIdx.x] = 0;
    __shared__ unsigned long long sT[1024];
    __shared__ int sM[1024];
    const int t = threadIdx.x;
    const int per = (p + 1023) / 1024;
    const int r0 = t * per;
    int r1 = r0 + per;
    if (r1 > p) r1 = p;
    int la = 0, lb = 0, lia = 0, lib = 0, md = 0;
    unsigned long long tt = 0;
    for (int r = r0; r < r1; ++r) {
        la += cntA[r];
        lb += cntB[r];
        lia += (cntA[r] + PAIRS_PER_ITEM - 1) / PAIRS_PER_ITEM;
        lib += (cntB[r] + PAIRS_PER_ITEM - 1) / PAIRS_PER_ITEM;
        const int d = deg[r];
        md = d > md ? d : md;
        tt += (unsigned long long)d * binom(d - 1, level);
    }
    sA[t] = la;
    sB[t] = lb;
    sIA[t] = lia;
    sIB[t] = lib;
    sT[t] = tt;
    sM[t] = md;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int va = 0, vb = 0, via = 0, vib = 0;
        if (t >= off) {
            va = sA[t - off];
            vb = sB[t - off];
            via = sIA[t - off];
            vib = sIB[t - off];
        }
        __syncthreads();
        sA[t] += va;
        sB[t] += vb;
        sIA[t] += via;
        sIB[t] += vib;
        __syncthreads();
    }
    int ea = sA[t] - la, eb = sB[t] - lb, eia = sIA[t] - lia, eib = sIB[t] - lib;
    for (int r = r0; r < r1; ++r) {
        rowoffA[r] = ea;
        rowoffB[r] = eb;
        itemoffA[r] = eia;
        itemoffB[r] = eib;
        ea += cntA[r];
        eb += cntB[r];
        eia += (cntA[r] + PAIRS_PER_ITEM - 1) / PAIRS_PER_ITEM;
        eib += (cntB[r] + PAIRS_PER_ITEM - 1) / PAIRS_PER_ITEM;
    }
    if (t == 1023) {
        rowoffA[p] = sA[1023];
        rowoffB[p] = sB[1023];
        itemoffA[p] = sIA[1023];
        itemoffB[p] = sIB[1023];
        for (int c = 0; c < NUM_COUNTERS; ++c) counters[c] = 0;
    }
    for (int off = 512; off >= 1; off >>= 1) {
        __syncthreads();
        if (t < off) {
            sT[t] += sT[t + off];
            sM[t] = sM[t] > sM[t + off] ? sM[t] : sM[t + off];
        }
    }
    if (t == 0 && st) {
        st[ST_TOTAL] = sT[0];
        st[ST_MAXDEG] = (unsigned long long)sM[0];
    }
}

// ----------------------------------------------------------------------------
// FP32 partial-correlation test on the (l+2) x (l+2) submatrix ordered
// (S..., i, j): Cholesky of C_SS, forward solves for C_Si and C_Sj, then the
// 2x2 Schur complement H = M0 - M1 M2^-1 M1^T of gaussian_helpers.cu:78-89 and
// the Fisher transform of :220-227.  dmin is the smallest pivot (residual
// variance), which scales the rounding error of H.
// ----------------------------------------------------------------------------
template <int L>
__device__ __forceinline__ float z32_from_values(float cij, const float (&ci)[L], const float (&cj)[L],
                                                 float (&c)[L][L], float &dmin, float &rho) {
    float rinv[L];
    dmin = 3.0e38f;
#pragma unroll
    for (int a = 0; a < L; ++a) {
#pragma unroll
        for (int b = 0; b < a; ++b) {
            float s = c[a][b];
#pragma unroll
            for (int k = 0; k < b; ++k) s -= c[a][k] * c[b][k];
            c[a][b] = s * rinv[b];
        }
        float d = c[a][a];
#pragma unroll
        for (int k = 0; k < a; ++k) d -= c[a][k] * c[a][k];
        dmin = fminf(dmin, d);
        rinv[a] = rsqrtf(d);
    }
    float wi[L], wj[L];
    float h00 = 1.0f, h11 = 1.0f, h01 = cij;
#pragma unroll
    for (int a = 0; a < L; ++a) {
        float si = ci[a], sj = cj[a];
#pragma unroll
        for (int k = 0; k < a; ++k) {
            si -= c[a][k] * wi[k];
            sj -= c[a][k] * wj[k];
        }
        wi[a] = si * rinv[a];
        wj[a] = sj * rinv[a];
        h00 -= wi[a] * wi[a];
        h11 -= wj[a] * wj[a];
        h01 -= wi[a] * wj[a];
    }
    dmin = fminf(dmin, fminf(h00, h11));
    rho = h01 * rsqrtf(fabsf(h00) * fabsf(h11));
    return fabsf(atanhf(rho));
}

template <int L>
__device__ __forceinline__ float z32_test(const float *__restrict__ C, int p, int i, int j,
                                          const int (&S)[L], float &dmin, float &rho) {
    float c[L][L];
    float ci[L], cj[L];
    const float cij = C[(size_t)i * p + j];
#pragma unroll
    for (int a = 0; a < L; ++a) {
        ci[a] = C[(size_t)i * p + S[a]];
        cj[a] = C[(size_t)j * p + S[a]];
#pragma unroll
        for (int b = 0; b <= a; ++b) c[a][b] = C[(size_t)S[a] * p + S[b]];
    }
    return z32_from_values<L>(cij, ci, cj, c, dmin, rho);
}

// error estimate of the FP32 Fisher-z: rounding of C32 and of the (L+2)-step
// elimination, amplified by the smallest pivot and by dz/drho = 1/(1-rho^2)
template <int L>
__device__ __forceinline__ float z32_error(float dmin, float rho) {
    return EPS32 * (float)(L + 2) * (1.0f + 1.0f / fmaxf(dmin, 1e-30f)) / fmaxf(1.0f - rho * rho, 1e-30f);
}

__device__ __forceinline__ bool z32_bad(float z, float dmin, float rho) {
    return !(dmin > 1e-5f) || !(z <= 3.0e38f) || !(fabsf(rho) < 1.0f);
}

// ----------------------------------------------------------------------------
// level l >= 1, one phase.  GS lanes form a group that owns one ordered pair
// (i, j) at a time; groups fetch pairs from a global counter.
// gpucsl/pc/cuda/gaussian_ci.cu:66-120, set_helpers.cu:25-71,
// graph_helpers.cu:81-133.
//
// FAST (SLOW = false): FP32 tests only.  When a test of the current iteration
// lies within the error margin of the threshold (or is ill-conditioned), the
// group stops before deciding anything from that iteration and queues
// (pair, iteration, zmin so far); the pair is finished by the SLOW variant,
// which reads the queue and rechecks such tests in FP64 with the reference's
// operation order.  Keeping the FP64 code out of the FAST kernel keeps its
// register count low.
// ----------------------------------------------------------------------------
struct SlowItem {
    int w;
    unsigned int iter;
    float zl;
    int pad;
};

template <int L, int GS, bool SLOW>
__global__ __launch_bounds__(256) void gauss_phase_kernel(
    const float *__restrict__ C32, const double *__restrict__ C64, int p,
    const int32_t *__restrict__ nbr, const int32_t *__restrict__ deg,
    const int32_t *__restrict__ rowoff, int side, float thr32, double thr, float margin_k,
    uint8_t *__restrict__ Gm, int32_t *__restrict__ sep, int maxL, float *__restrict__ zmin,
    unsigned int *__restrict__ stripe_counters, SlowItem *__restrict__ queue,
    unsigned int *__restrict__ qcount, unsigned long long *__restrict__ bstats) {
    const int lane = __lane_id();
    const int gbase = lane & ~(GS - 1);
    const int gl = lane & (GS - 1);
    const uint64_t gmask = (GS == 64) ? ~0ull : (((1ull << GS) - 1ull) << gbase);
    const int total = SLOW ? (int)*qcount : rowoff[p];

    // Work distribution without hot atomics: each group first takes
    // S = total / groups pairs statically (interleaved, so neighbouring groups
    // share row i), then the remainder through one of 8 striped counters
    // chosen by block index.
    constexpr int GPB = 256 / GS;
    const int groups = gridDim.x * GPB;
    const int gid = blockIdx.x * GPB + (threadIdx.x / GS);
    // at least one static pair per group, so a level with fewer pairs than
    // groups runs without any atomics
    const int S_static = total <= groups ? 1 : total / groups;
    const int dyn_base = total <= groups ? total : S_static * groups;
    const int nstripes = gridDim.x < STRIPES ? gridDim.x : STRIPES;
    const int stripe = blockIdx.x % nstripes;
    const int rem = total - dyn_base;
    const int stripe_len = rem > stripe ? (rem - stripe + nstripes - 1) / nstripes : 0;
    unsigned int *counter = stripe_counters + stripe;
    int k_static = 0;

    int w = 0, i = 0, j = 0, jpos = 0, Nn = 0;
    uint64_t count = 0, iter = 0;
    float zl = __builtin_inff();
    bool have = false, done = false;
    unsigned long long n_exec = 0, n_re = 0, n_flip = 0, n_useful = 0;
    float mr = 0.0f;

    while (true) {
        if (!have && !done) {
            int t = total;
            if (k_static < S_static) {
                t = gid + k_static * groups;
                ++k_static;
                if (t >= dyn_base) t = total;
            } else if (stripe_len > 0) {
                if (gl == 0) {
                    const unsigned int cur =
                        __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if ((int)cur < stripe_len) {
                        const unsigned int k = atomicAdd(counter, 1u);
                        if ((int)k < stripe_len) t = dyn_base + stripe + nstripes * (int)k;
                    }
                }
                t = __shfl(t, gbase);
            }
            if (t >= total) {
                done = true;
            } else {
                unsigned int it0 = 0;
                float z0 = __builtin_inff();
                if (SLOW) {
                    const SlowItem q = queue[t];
                    w = q.w;
                    it0 = q.iter;
                    z0 = q.zl;
                } else {
                    w = t;
                }
                int lo = 0, hi = p;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (rowoff[mid] <= w) lo = mid; else hi = mid;
                }
                i = lo;
                const int k = w - rowoff[i];
                const int d = deg[i];
                jpos = side == 0 ? d - (rowoff[i + 1] - rowoff[i]) + k : k;
                j = nbr[(size_t)i * p + jpos];
                Nn = d - 1;
                count = binom(Nn, L);
                iter = it0;
                zl = gl == 0 ? z0 : __builtin_inff();
                have = count > 0 && (SLOW || side == 0 || Gm[(size_t)i * p + j] != 0);
            }
        }
        if (__ballot(!done) == 0) break;
        if (!have) continue;

        const uint64_t rank = iter * GS + gl;
        const bool act = rank < count;
        bool pass = false, near = false;
        float zf = __builtin_inff();
        int S[L];
        if (act) {
            int pos[L];
            unrank_lex<L>(Nn, rank, pos);
#pragma unroll
            for (int c = 0; c < L; ++c) {
                int idx = pos[c] - 1;
                idx += (idx >= jpos);
                S[c] = nbr[(size_t)i * p + idx];
            }
            float dmin, rho;
            const float z = z32_test<L>(C32, p, i, j, S, dmin, rho);
            const float err = EPS32 * (float)(L + 2) * (1.0f + 1.0f / fmaxf(dmin, 1e-30f)) /
                              fmaxf(1.0f - rho * rho, 1e-30f);
            const bool bad = !(dmin > 1e-5f) || !(z <= 3.0e38f) || !(fabsf(rho) < 1.0f);
            near = bad || fabsf(z - thr32) <= margin_k * err;
            if (SLOW && near) {
                const double z64 = ref64::gauss_z<L>(C64, p, i, j, S);
                pass = z64 < thr;
                zf = (float)z64;
                ++n_re;
                if (!bad) {
                    n_flip += (pass != (z < thr32));
                    mr = fmaxf(mr, fabsf(z - zf) / err);
                }
            } else {
                pass = z < thr32;
                zf = z;
            }
        }
        if (!SLOW) {
            if (__ballot(near) & gmask) {
                // hand the pair to the FP64-capable kernel from this iteration on
#pragma unroll
                for (int off = GS / 2; off >= 1; off >>= 1) zl = fminf(zl, __shfl_xor(zl, off));
                if (gl == 0) {
                    const unsigned int slot = atomicAdd(qcount, 1u);
                    queue[slot] = SlowItem{w, (unsigned int)iter, zl, 0};
                }
                have = false;
                continue;
            }
        }
        if (act) ++n_exec;
        const uint64_t pm = __ballot(pass) & gmask;
        bool finish = false;
        if (pm) {
            const int P = __ffsll((unsigned long long)pm) - 1 - gbase;
            if (act && gl <= P) {
                ++n_useful;
                if (zf < zl) zl = zf;
            }
            if (gl == P) {
                Gm[(size_t)i * p + j] = 0;
                Gm[(size_t)j * p + i] = 0;
#pragma unroll
                for (int c = 0; c < L; ++c) {
                    sep[((size_t)i * p + j) * maxL + c] = S[c];
                    sep[((size_t)j * p + i) * maxL + c] = S[c];
                }
            }
            finish = true;
        } else {
            if (act) {
                ++n_useful;
                if (zf < zl) zl = zf;
            }
            ++iter;
            finish = iter * GS >= count;
        }
        if (finish) {
#pragma unroll
            for (int off = GS / 2; off >= 1; off >>= 1) zl = fminf(zl, __shfl_xor(zl, off));
            if (gl == 0) {
                const int a = i < j ? i : j, b = i < j ? j : i;
                float *zp = zmin + (size_t)a * p + b;
                if (zl < *zp) *zp = zl;
            }
            have = false;
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        n_exec += __shfl_xor(n_exec, off);
        n_re += __shfl_xor(n_re, off);
        n_flip += __shfl_xor(n_flip, off);
        n_useful += __shfl_xor(n_useful, off);
        mr = fmaxf(mr, __shfl_xor(mr, off));
    }
    __shared__ unsigned long long wstat[4][5];
    const int wave = threadIdx.x >> 6;
    if (lane == 0) {
        wstat[wave][BS_EXEC] = n_exec;
        wstat[wave][BS_USEFUL] = n_useful;
        wstat[wave][BS_RECHECK] = n_re;
        wstat[wave][BS_FLIP] = n_flip;
        wstat[wave][BS_MAXRATIO] = __float_as_uint(mr);
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        unsigned long long v = 0;
        for (int w2 = 0; w2 < 4; ++w2) {
            const unsigned long long x = wstat[w2][threadIdx.x];
            v = threadIdx.x == BS_MAXRATIO ? (x > v ? x : v) : v + x;
        }
        unsigned long long *dst = bstats + (size_t)(blockIdx.x % BSTAT_STRIPES) * BSTAT_WORDS + threadIdx.x;
        if (v) {
            if (threadIdx.x == BS_MAXRATIO)
                atomicMax(dst, v);
            else
                atomicAdd(dst, v);
        }
    }
}

// ----------------------------------------------------------------------------
// level l >= 1, FAST phase, row-blocked.  One wave per block; a work item is
// (row i, up to PAIRS_PER_ITEM ordered pairs (i, j) of this phase).  The wave
// stages what every test of row i reads into LDS:
//   mode FULL   (deg + 1)^2 floats: C[A, A] with A = adj(i) + [i], so every CI
//               test reads LDS only;
//   mode ROW    neighbour ids, C[i, adj(i)] and diag C[s, s]; C[j, S] and the
//               off-diagonal C[S, S] come from global memory (L2/MALL);
//   mode GLOBAL nothing staged (only for degrees beyond the LDS budget).
// Lane groups of GS lanes walk a pair's subset ranks in order as in the
// unblocked kernel; near-threshold tests go to the FP64 queue.
// ----------------------------------------------------------------------------
template <int L, int GS>
__global__ __launch_bounds__(64) void gauss_rowblock_kernel(
    const float *__restrict__ C32, int p, const int32_t *__restrict__ nbr,
    const int32_t *__restrict__ deg, const int32_t *__restrict__ rowoff,
    const int32_t *__restrict__ itemoff, int side, float thr32, float margin_k,
    uint8_t *__restrict__ Gm, int32_t *__restrict__ sep, int maxL, float *__restrict__ zmin,
    SlowItem *__restrict__ queue, unsigned int *__restrict__ qcount,
    unsigned long long *__restrict__ bstats) {
    extern __shared__ int lds_words[];
    const int lane = threadIdx.x;
    const int gbase = lane & ~(GS - 1);
    const int gl = lane & (GS - 1);
    const int grp = lane / GS;
    constexpr int NG = 64 / GS;
    const uint64_t gmask = (GS == 64) ? ~0ull : (((1ull << GS) - 1ull) << gbase);
    const int nitems = itemoff[p];
    constexpr int WORDS = RB_LDS_BYTES / 4;

    unsigned long long n_exec = 0, n_useful = 0;

    for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
        int lo = 0, hi = p;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (itemoff[mid] <= item) lo = mid; else hi = mid;
        }
        const int i = lo;
        const int chunk = item - itemoff[i];
        const int d = deg[i];
        const int cnt = rowoff[i + 1] - rowoff[i];
        const int q0 = chunk * PAIRS_PER_ITEM;
        const int q1 = min(cnt, q0 + PAIRS_PER_ITEM);
        const int jbase = side == 0 ? d - cnt : 0;
        const int D = d + 1;
        const int mode = (D * D + d <= WORDS) ? 0 : (3 * d <= WORDS ? 1 : 2);
        const int32_t *gnb = nbr + (size_t)i * p;

        __syncthreads();  // the previous item is done with LDS
        int *snb = lds_words;
        float *sf = reinterpret_cast<float *>(lds_words + d);
        if (mode != 2) {
            for (int k0 = 0; k0 < d; k0 += 64 * 4) {
                int v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = k0 + u * 64 + lane;
                    v[u] = k < d ? gnb[k] : 0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = k0 + u * 64 + lane;
                    if (k < d) snb[k] = v[u];
                }
            }
        }
        __syncthreads();
        // Loads are issued in batches before their LDS stores so that each lane
        // keeps several gathers in flight.
        if (mode == 0) {
            // sub[a * D + b] = C[A[a], A[b]], A = adj(i) + [i]; D <= 63 < 64
            const bool colok = lane < D;
            const int vb = lane < d ? snb[colok ? lane : 0] : i;
            for (int a0 = 0; a0 < D; a0 += 8) {
                float v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int a = a0 + u;
                    const int va = a < d ? snb[a] : i;
                    v[u] = (a < D && colok) ? C32[(size_t)va * p + vb] : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int a = a0 + u;
                    if (a < D && colok) sf[a * D + lane] = v[u];
                }
            }
        } else if (mode == 1) {
            for (int k0 = 0; k0 < d; k0 += 64 * 4) {
                float v[4], w[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = k0 + u * 64 + lane;
                    const int vv = k < d ? snb[k] : 0;
                    v[u] = k < d ? C32[(size_t)i * p + vv] : 0.0f;
                    w[u] = k < d ? C32[(size_t)vv * p + vv] : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = k0 + u * 64 + lane;
                    if (k < d) {
                        sf[k] = v[u];
                        sf[d + k] = w[u];
                    }
                }
            }
        }
        __syncthreads();

        const int Nn = d - 1;
        const uint64_t count = binom(Nn, L);
        for (int q = q0 + grp; q < q1; q += NG) {
            const int jpos = jbase + q;
            const int j = mode == 2 ? gnb[jpos] : snb[jpos];
            if (count == 0) continue;
            if (side == 1 && Gm[(size_t)i * p + j] == 0) continue;
            float zl = __builtin_inff();
            for (uint64_t iter = 0;; ++iter) {
                const uint64_t rank = iter * GS + gl;
                const bool act = rank < count;
                bool pass = false, near = false;
                float zf = __builtin_inff();
                int sp[L];
                if (act) {
                    int pos[L];
                    unrank_lex<L>(Nn, rank, pos);
#pragma unroll
                    for (int c = 0; c < L; ++c) {
                        int idx = pos[c] - 1;
                        sp[c] = idx + (idx >= jpos);
                    }
                    float cij, ci[L], cj[L], cm[L][L];
                    if (mode == 0) {
                        cij = sf[d * D + jpos];
#pragma unroll
                        for (int a = 0; a < L; ++a) {
                            ci[a] = sf[d * D + sp[a]];
                            cj[a] = sf[jpos * D + sp[a]];
#pragma unroll
                            for (int b = 0; b <= a; ++b) cm[a][b] = sf[sp[a] * D + sp[b]];
                        }
                    } else if (mode == 1) {
                        int S[L];
#pragma unroll
                        for (int a = 0; a < L; ++a) S[a] = snb[sp[a]];
                        cij = sf[jpos];
#pragma unroll
                        for (int a = 0; a < L; ++a) {
                            ci[a] = sf[sp[a]];
                            cj[a] = C32[(size_t)j * p + S[a]];
#pragma unroll
                            for (int b = 0; b < a; ++b) cm[a][b] = C32[(size_t)S[a] * p + S[b]];
                            cm[a][a] = sf[d + sp[a]];
                        }
                    } else {
                        int S[L];
#pragma unroll
                        for (int a = 0; a < L; ++a) S[a] = gnb[sp[a]];
                        cij = C32[(size_t)i * p + j];
#pragma unroll
                        for (int a = 0; a < L; ++a) {
                            ci[a] = C32[(size_t)i * p + S[a]];
                            cj[a] = C32[(size_t)j * p + S[a]];
#pragma unroll
                            for (int b = 0; b <= a; ++b) cm[a][b] = C32[(size_t)S[a] * p + S[b]];
                        }
                    }
                    float dmin, rho;
                    const float z = z32_from_values<L>(cij, ci, cj, cm, dmin, rho);
                    near = z32_bad(z, dmin, rho) || fabsf(z - thr32) <= margin_k * z32_error<L>(dmin, rho);
                    pass = z < thr32;
                    zf = z;
                }
                if (__ballot(near) & gmask) {
#pragma unroll
                    for (int off = GS / 2; off >= 1; off >>= 1) zl = fminf(zl, __shfl_xor(zl, off));
                    if (gl == 0) {
                        const unsigned int slot = atomicAdd(qcount, 1u);
                        queue[slot] = SlowItem{rowoff[i] + q, (unsigned int)iter, zl, 0};
                    }
                    break;
                }
                if (act) ++n_exec;
                const uint64_t pm = __ballot(pass) & gmask;
                bool finish;
                if (pm) {
                    const int P = __ffsll((unsigned long long)pm) - 1 - gbase;
                    if (act && gl <= P) {
                        ++n_useful;
                        if (zf < zl) zl = zf;
                    }
                    if (gl == P) {
                        Gm[(size_t)i * p + j] = 0;
                        Gm[(size_t)j * p + i] = 0;
#pragma unroll
                        for (int c = 0; c < L; ++c) {
                            const int v = mode == 2 ? gnb[sp[c]] : snb[sp[c]];
                            sep[((size_t)i * p + j) * maxL + c] = v;
                            sep[((size_t)j * p + i) * maxL + c] = v;
                        }
                    }
                    finish = true;
                } else {
                    if (act) {
                        ++n_useful;
                        if (zf < zl) zl = zf;
                    }
                    finish = (iter + 1) * GS >= count;
                }
                if (finish) {
#pragma unroll
                    for (int off = GS / 2; off >= 1; off >>= 1) zl = fminf(zl, __shfl_xor(zl, off));
                    if (gl == 0) {
                        const int a = i < j ? i : j, b = i < j ? j : i;
                        float *zp = zmin + (size_t)a * p + b;
                        if (zl < *zp) *zp = zl;
                    }
                    break;
                }
            }
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        n_exec += __shfl_xor(n_exec, off);
        n_useful += __shfl_xor(n_useful, off);
    }
    if (lane == 0 && n_exec) {
        unsigned long long *dst = bstats + (size_t)(blockIdx.x % BSTAT_STRIPES) * BSTAT_WORDS;
        atomicAdd(dst + BS_EXEC, n_exec);
        atomicAdd(dst + BS_USEFUL, n_useful);
    }
}

// Sums the striped accumulators of one level into its statistics slots.
__global__ __launch_bounds__(64) void level_stats_kernel(const unsigned long long *__restrict__ bstats,
                                                         unsigned long long *__restrict__ st) {
    const int t = threadIdx.x;
    unsigned long long v[5];
    for (int k = 0; k < 5; ++k) v[k] = bstats[(size_t)t * BSTAT_WORDS + k];
    for (int off = 32; off >= 1; off >>= 1) {
        for (int k = 0; k < 4; ++k) v[k] += __shfl_xor(v[k], off);
        const unsigned long long o = __shfl_xor(v[4], off);
        v[4] = o > v[4] ? o : v[4];
    }
    if (t == 0) {
        st[ST_EXEC] = v[BS_EXEC];
        st[ST_USEFUL] = v[BS_USEFUL];
        st[ST_RECHECK] = v[BS_RECHECK];
        st[ST_FLIP] = v[BS_FLIP];
        st[ST_MAXRATIO] = v[BS_MAXRATIO];
    }
}

// ----------------------------------------------------------------------------
// pmax, init and output conversion
// ----------------------------------------------------------------------------
// pmax = 2 (1 - Phi(sqrt(n-3) |zmin|)) = erfc(sqrt(n-3) |zmin| / sqrt 2); -1 on
// and below the diagonal.  gpucsl/pc/helpers.py:62-82.
__global__ void pmax_kernel(const float *__restrict__ zmin, int p, double sqrt_nm3,
                            float *__restrict__ pmax) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (b >= p) return;
    const size_t t = (size_t)a * p + b;
    if (b <= a) {
        pmax[t] = -1.0f;
        return;
    }
    const double x = fabs(sqrt_nm3 * (double)zmin[t]);
    pmax[t] = (float)erfc(x * 0.70710678118654752440);
}

__global__ void to_u16_kernel(const uint8_t *__restrict__ g, size_t count, uint16_t *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) out[i] = g[i];
}

// ----------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------
// counters: see NUM_COUNTERS; all reset by scan_kernel.  bstats: per-block
// statistics partials of the four phase kernels of a level.
struct GaussWs {
    int32_t *nbr, *deg, *cntA, *cntB, *rowoffA, *rowoffB, *itemoffA, *itemoffB;
    unsigned int *counters;
    SlowItem *queueA, *queueB;
    unsigned long long *bstats;
};

// A phase has at most one queue entry per ordered pair of its side, and a side
// has at most p(p-1)/2 pairs.
static size_t queue_capacity(int p) { return (size_t)p * (p - 1) / 2 + 1; }

size_t gauss_workspace_bytes(int p) {
    size_t off = 0;
    off = align_up(off + sizeof(int32_t) * (size_t)p * p, 256);
    off = align_up(off + sizeof(int32_t) * (size_t)p * 3, 256);
    off = align_up(off + sizeof(int32_t) * (size_t)(p + 1) * 4, 256);
    off = align_up(off + 256, 256);
    off = align_up(off + sizeof(SlowItem) * queue_capacity(p) * 2, 256);
    off = align_up(off + sizeof(unsigned long long) * BSTAT_STRIPES * BSTAT_WORDS, 256);
    return off;
}

static GaussWs carve(void *ws, int p) {
    GaussWs w;
    char *b = (char *)ws;
    size_t off = 0;
    w.nbr = (int32_t *)(b + off);
    off = align_up(off + sizeof(int32_t) * (size_t)p * p, 256);
    w.deg = (int32_t *)(b + off);
    w.cntA = w.deg + p;
    w.cntB = w.deg + 2 * p;
    off = align_up(off + sizeof(int32_t) * (size_t)p * 3, 256);
    w.rowoffA = (int32_t *)(b + off);
    w.rowoffB = w.rowoffA + (p + 1);
    w.itemoffA = w.rowoffB + (p + 1);
    w.itemoffB = w.itemoffA + (p + 1);
    off = align_up(off + sizeof(int32_t) * (size_t)(p + 1) * 4, 256);
    w.counters = (unsigned int *)(b + off);
    off = align_up(off + 256, 256);
    w.queueA = (SlowItem *)(b + off);
    w.queueB = w.queueA + queue_capacity(p);
    off = align_up(off + sizeof(SlowItem) * queue_capacity(p) * 2, 256);
    w.bstats = (unsigned long long *)(b + off);
    return w;
}

static int phase_grid() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        int n = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cus = n;
    }
    int g = cus * PHASE_BLOCKS_PER_CU;
    return g > MAX_PHASE_BLOCKS ? MAX_PHASE_BLOCKS : (g < 4 * STRIPES ? 4 * STRIPES : g);
}
static int slow_grid() { return phase_grid() / 4; }
static int rowblock_grid() {
    int g = phase_grid() / PHASE_BLOCKS_PER_CU * RB_BLOCKS_PER_CU;
    return g > MAX_PHASE_BLOCKS ? MAX_PHASE_BLOCKS : g;
}

int gauss_init_launch(int p, int max_level, uint8_t *Gm, int32_t *sep, float *zmin, hipStream_t s) {
    if (p < 1 || max_level < 0 || !Gm || !zmin) return set_error(ERR_ARG, "gauss_init: bad arguments");
    launch_fill<uint8_t>(Gm, (uint8_t)1, (size_t)p * p, s);
    if (max_level > 0 && sep) launch_fill<int32_t>(sep, -1, (size_t)p * p * max_level, s);
    launch_fill<float>(zmin, __builtin_inff(), (size_t)p * p, s);
    return check_hip(hipGetLastError(), "gauss_init launch");
}

int gauss_level0_launch(const float *C32, const double *C64, int p, double thr, float margin_k,
                        uint8_t *Gm, float *zmin, unsigned long long *stats, hipStream_t s) {
    if (!C32 || !C64 || !Gm || !zmin || !stats || p < 1) return set_error(ERR_ARG, "gauss_level0: bad arguments");
    dim3 grid((p + 255) / 256, p);
    GPUCSL_LAUNCH("gauss_level0", gauss_level0_kernel, grid, dim3(256), 0, s, C32, C64, p, (float)thr, thr,
                       margin_k, Gm, zmin, stats);
    return check_hip(hipGetLastError(), "gauss_level0 launch");
}

// Views of the level-start neighbour lists inside a Gaussian/discrete workspace.
void level_list_views(void *ws, int p, int32_t **nbr, int32_t **deg) {
    GaussWs w = carve(ws, p);
    *nbr = w.nbr;
    *deg = w.deg;
}

int compact_launch(const uint8_t *Gm, int p, int level, void *ws, unsigned long long *stats,
                   hipStream_t s) {
    GaussWs w = carve(ws, p);
    GPUCSL_LAUNCH("compact", compact_kernel, dim3((p + 3) / 4), dim3(256), 0, s, Gm, p, w.nbr, w.deg, w.cntA,
                       w.cntB);
    GPUCSL_LAUNCH("compact_scan", scan_kernel, dim3(1), dim3(1024), 0, s, w.deg, w.cntA, w.cntB, p, level,
                       w.rowoffA, w.rowoffB, w.itemoffA, w.itemoffB, w.counters, w.bstats, stats);
    return check_hip(hipGetLastError(), "compact launch");
}

static const char *const kPhaseFastName[] = {
    "", "gauss_level1_fast", "gauss_level2_fast", "gauss_level3_fast", "gauss_level4_fast",
    "gauss_level5_fast", "gauss_level6_fast", "gauss_level7_fast", "gauss_level8_fast"};
static const char *const kPhaseSlowName[] = {
    "", "gauss_level1_slow", "gauss_level2_slow", "gauss_level3_slow", "gauss_level4_slow",
    "gauss_level5_slow", "gauss_level6_slow", "gauss_level7_slow", "gauss_level8_slow"};

template <int L, int GS>
static void launch_phase(const float *C32, const double *C64, int p, const GaussWs &w, int side,
                         double thr, float margin_k, uint8_t *Gm, int32_t *sep, int maxL,
                         float *zmin, unsigned long long *st, hipStream_t s) {
    (void)st;
    const int32_t *rowoff = side == 0 ? w.rowoffA : w.rowoffB;
    SlowItem *queue = side == 0 ? w.queueA : w.queueB;
    unsigned int *qcount = w.counters + CNT_QUEUE + side;
    unsigned int *fast_ctr = w.counters + (2 * side) * STRIPES;
    unsigned int *slow_ctr = w.counters + (2 * side + 1) * STRIPES;
    unsigned long long *fast_bs = w.bstats;
    unsigned long long *slow_bs = w.bstats;
    (void)fast_ctr;
    GPUCSL_LAUNCH(kPhaseFastName[L], (gauss_rowblock_kernel<L, GS>), dim3(rowblock_grid()), dim3(64),
                  RB_LDS_BYTES, s, C32, p, w.nbr, w.deg, rowoff, side == 0 ? w.itemoffA : w.itemoffB,
                  side, (float)thr, margin_k, Gm, sep, maxL, zmin, queue, qcount, fast_bs);
    GPUCSL_LAUNCH(kPhaseSlowName[L], (gauss_phase_kernel<L, GS, true>), dim3(slow_grid()), dim3(256), 0, s, C32,
                       C64, p, w.nbr, w.deg, rowoff, side, (float)thr, thr, margin_k, Gm, sep, maxL,
                       zmin, slow_ctr, queue, qcount, slow_bs);
}

template <int L>
static int phase_dispatch_gs(int gs, const float *C32, const double *C64, int p, const GaussWs &w,
                             int side, double thr, float margin_k, uint8_t *Gm, int32_t *sep,
                             int maxL, float *zmin, unsigned long long *st, hipStream_t s) {
    switch (gs) {
    case 8: launch_phase<L, 8>(C32, C64, p, w, side, thr, margin_k, Gm, sep, maxL, zmin, st, s); break;
    case 16: launch_phase<L, 16>(C32, C64, p, w, side, thr, margin_k, Gm, sep, maxL, zmin, st, s); break;
    case 32: launch_phase<L, 32>(C32, C64, p, w, side, thr, margin_k, Gm, sep, maxL, zmin, st, s); break;
    case 64: launch_phase<L, 64>(C32, C64, p, w, side, thr, margin_k, Gm, sep, maxL, zmin, st, s); break;
    default: return set_error(ERR_ARG, "group size must be 8, 16, 32 or 64 (got %d)", gs);
    }
    return OK;
}

int gauss_level_launch(int level, const float *C32, const double *C64, int p, int max_level,
                       double thr, float margin_k, int group_size, uint8_t *Gm, int32_t *sep,
                       float *zmin, unsigned long long *stats, void *ws, size_t ws_bytes,
                       hipStream_t s) {
    if (level < 1 || level > max_level) return set_error(ERR_ARG, "gauss_level: level %d outside [1, %d]", level, max_level);
    if (level > GPUCSL_MAX_LEVEL_TEMPLATE)
        return set_error(ERR_UNSUPPORTED, "gauss_level: levels above %d are not built", GPUCSL_MAX_LEVEL_TEMPLATE);
    if (!C32 || !C64 || !Gm || !sep || !zmin || !stats || !ws) return set_error(ERR_ARG, "gauss_level: null pointer");
    if (ws_bytes < gauss_workspace_bytes(p)) return set_error(ERR_WORKSPACE, "gauss_level: workspace too small");
    int rc = compact_launch(Gm, p, level, ws, stats, s);
    if (rc) return rc;
    GaussWs w = carve(ws, p);
    for (int side = 0; side < 2; ++side) {
        switch (level) {
#define GPUCSL_CASE(LV)                                                                              \
    case LV:                                                                                         \
        rc = phase_dispatch_gs<LV>(group_size, C32, C64, p, w, side, thr, margin_k, Gm, sep, max_level, \
                                   zmin, stats, s);                                                  \
        break;
            GPUCSL_CASE(1) GPUCSL_CASE(2) GPUCSL_CASE(3) GPUCSL_CASE(4)
            GPUCSL_CASE(5) GPUCSL_CASE(6) GPUCSL_CASE(7) GPUCSL_CASE(8)
#undef GPUCSL_CASE
        }
        if (rc) return rc;
    }
    GPUCSL_LAUNCH("level_stats", level_stats_kernel, dim3(1), dim3(64), 0, s, w.bstats, stats);
    return check_hip(hipGetLastError(), "gauss_level launch");
}

int pmax_launch(const float *zmin, int p, int64_t n, float *pmax, hipStream_t s) {
    if (!zmin || !pmax || p < 1 || n < 4) return set_error(ERR_ARG, "pmax: bad arguments");
    GPUCSL_LAUNCH("pmax", pmax_kernel, dim3((p + 255) / 256, p), dim3(256), 0, s, zmin, p,
                       sqrt((double)(n - 3)), pmax);
    return check_hip(hipGetLastError(), "pmax launch");
}

int to_u16_launch(const uint8_t *Gm, int p, uint16_t *out, hipStream_t s) {
    const size_t cnt = (size_t)p * p;
    size_t blocks = (cnt + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    GPUCSL_LAUNCH("to_u16", to_u16_kernel, dim3((unsigned)blocks), dim3(256), 0, s, Gm, cnt, out);
    return check_hip(hipGetLastError(), "to_u16 launch");
}

}  // namespace gpucsl
