// Gaussian PC-stable skeleton on gfx950, for a table of jobs (datasets).
//
// Replaces gpucsl/pc/cuda/gaussian_ci.cu (level 0 and level n), the compaction
// of gpucsl/pc/cuda/helpers/graph_helpers.cu:35-56 and the host level loop of
// gpucsl/pc/discover_skeleton_gaussian.py:30-113.  Every kernel takes a device
// array of job descriptors (gpucsl_gauss_job): a single-dataset call is a
// table of one, a batch runs all its datasets in the same launches.
//
// Layout in HBM per job (p variables, L = max level):
//   C32  p*p float   correlation, FP32 (read by the CI tests)
//   C64  p*p double  correlation, FP64 (read only by rechecks)
//   G    p*p uint8   adjacency (1 edge, 0 none), kept symmetric
//   sep  p*p*L int32 separation sets, -1 padded, [(i*p+j)*L + k]
//   zmin p*p float   min Fisher-z per edge, upper triangle [min*p+max]
//   ws   nbr p*p int32 (level-start neighbour lists), deg/cntA/cntB p int32,
//        rowoffA/rowoffB/itemoffA/itemoffB (p+1) int32
// Per launch: 64 counters, 64x8 statistics stripes, per-job work-item bases,
// and two FP64-recheck queues.
//
// A level l >= 1 runs as compaction -> per-job scan -> job prefix -> phase A ->
// phase B.  Phase A tests every level-start edge (a, b), a < b, with S drawn
// from adj(a)\{b}; phase B tests (b, a) with S from adj(b)\{a} for the edges
// phase A kept.  Lane groups walk a pair's subset ranks in increasing order and
// stop at the first passing test, so the removing sepset is the
// lexicographically first one in pcalg order and zmin covers exactly the tests
// up to it.  No host sync happens between levels.
#include "job.h"

#include <mutex>

namespace gpucsl {

constexpr float EPS32 = 5.9604645e-8f;  // 2^-24
// FAST kernel: 256-thread blocks with an LDS staging area of at most 64 KB
constexpr int FB_MAX_LDS_BYTES = 65536;

// Lanes that test one pair together, from the pair's subset count: gs_min
// lanes (gs_min in {8, 16, 32, 64}) while that takes at most small_iters
// iterations, a 64-lane wave while that takes at most wave_iters, else the
// whole 256-thread block.
struct Team {
    int gs_min, small_max, wave_max;  // count limits of the gs_min and wave teams
    int item_tests;                   // subset tests a work item aims at
    int wave_items;                   // 1: every wave takes whole items on its own (teams of <= 64 lanes)
};

__host__ __device__ __forceinline__ int team_lanes(uint64_t count, const Team &t) {
    if (count <= (uint64_t)t.small_max) return t.gs_min;
    if (count <= (uint64_t)t.wave_max) return 64;
    return 256;
}

// Pairs per work item: enough for every team of the block, and about
// item_tests subset tests in all.
__host__ __device__ __forceinline__ int items_pairs(uint64_t count, const Team &t) {
    const int teams = (t.wave_items ? 64 : 256) / team_lanes(count, t);
    // ceil(item_tests / count); item_tests < 2^25, so a 32-bit division
    const uint32_t it = (uint32_t)t.item_tests;
    const uint32_t w = count == 0 ? 1u : (count >= it ? 1u : (it + (uint32_t)count - 1u) / (uint32_t)count);
    return (int)w > teams ? (int)w : teams;
}

// C(m, level) with the level's divisions by constants
__host__ __device__ __forceinline__ uint64_t binom_level(int m, int level) {
    switch (level) {
        case 1: return m >= 1 ? (uint64_t)m : 0;
        case 2: return binom_k<2>(m);
        case 3: return binom_k<3>(m);
        case 4: return binom_k<4>(m);
        case 5: return binom_k<5>(m);
        case 6: return binom_k<6>(m);
        case 7: return binom_k<7>(m);
        case 8: return binom_k<8>(m);
        default: return binom(m, level);
    }
}

// Launch shapes of the FAST kernel per (launch bucket, level), settable at run
// time from an autotuning table.  The bucket is chosen on the host from the
// launch: one dataset with p <= 256, one dataset with p > 256, or a batch of
// several; the degree of a row is handled on the device, where each row's
// subset count C(deg - 1, l) picks its team size against the shape's
// thresholds.  Level index 0 is unused.  Process-wide, defaults on first
// use; reads and writes under one mutex.
struct TuningTable {
    gpucsl_level_tuning lv[GPUCSL_TUNING_BUCKETS][GPUCSL_MAX_LEVEL_TEMPLATE + 1];
};
static TuningTable &tuning_table() {
    static TuningTable t = [] {
        TuningTable x;
        for (int b = 0; b < GPUCSL_TUNING_BUCKETS; ++b)
            for (int l = 0; l <= GPUCSL_MAX_LEVEL_TEMPLATE; ++l)
                x.lv[b][l] = gpucsl_level_tuning{8, 8, 4, 32768, 4096, 0};
        return x;
    }();
    return t;
}
static std::mutex g_tuning_mu;

int tuning_bucket(int nj, int max_p) {
    if (nj > 1) return GPUCSL_TUNING_BATCH;
    return max_p <= 256 ? GPUCSL_TUNING_SMALL : GPUCSL_TUNING_LARGE;
}

static gpucsl_level_tuning tuning_of(int bucket, int level) {
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    return tuning_table().lv[bucket][level];
}

static Team team_for(int bucket, int level, int gs_min) {
    const gpucsl_level_tuning t = tuning_of(bucket, level);
    if (t.wave_items) return Team{gs_min, gs_min * t.small_team_iters, 0x7fffffff, t.item_tests, 1};
    return Team{gs_min, gs_min * t.small_team_iters, 64 * t.wave_team_iters, t.item_tests, 0};
}

// bucket -1: every bucket; level 0: every level
int set_level_tuning(int bucket, int level, const gpucsl_level_tuning *t) {
    if (!t) return set_error(ERR_ARG, "tuning: null");
    if (t->small_team_iters < 1 || t->wave_team_iters < 1 || t->fast_blocks_per_cu < 1 || t->fast_blocks_per_cu > 32 ||
        t->fast_lds_bytes < 4096 || t->fast_lds_bytes > FB_MAX_LDS_BYTES || t->fast_lds_bytes % 1024 ||
        t->item_tests < 64 || t->item_tests > (1 << 24) || t->wave_items < 0 || t->wave_items > 1)
        return set_error(ERR_ARG, "tuning: a field is out of range");
    if (bucket < -1 || bucket >= GPUCSL_TUNING_BUCKETS)
        return set_error(ERR_ARG, "tuning: bucket %d outside -1..%d", bucket, GPUCSL_TUNING_BUCKETS - 1);
    if (level < 0 || level > GPUCSL_MAX_LEVEL_TEMPLATE)
        return set_error(ERR_ARG, "tuning: level %d outside 0..%d", level, GPUCSL_MAX_LEVEL_TEMPLATE);
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    for (int b = 0; b < GPUCSL_TUNING_BUCKETS; ++b) {
        if (bucket >= 0 && b != bucket) continue;
        for (int l = 1; l <= GPUCSL_MAX_LEVEL_TEMPLATE; ++l)
            if (level == 0 || l == level) tuning_table().lv[b][l] = *t;
    }
    return OK;
}

int get_level_tuning(int bucket, int level, gpucsl_level_tuning *t) {
    if (!t || level < 1 || level > GPUCSL_MAX_LEVEL_TEMPLATE || bucket < 0 || bucket >= GPUCSL_TUNING_BUCKETS)
        return set_error(ERR_ARG, "tuning: bad bucket or level, or null");
    *t = tuning_of(bucket, level);
    return OK;
}

enum BStat { BS_EXEC = 0, BS_USEFUL = 1, BS_RECHECK = 2, BS_FLIP = 3, BS_MAXRATIO = 4 };
enum StatSlot { ST_EXEC = 0, ST_TOTAL = 1, ST_RECHECK = 2, ST_FLIP = 3, ST_MAXRATIO = 4, ST_MAXDEG = 5, ST_USEFUL = 6, ST_SLOTS = 8 };

// zmin scale of a level's tests: 1, or sqrt(n - l - 3) in pcalg mode (so
// that zmin orders tests by their p-value across levels)
__device__ __forceinline__ float level_zscale(const Job &J, int level) {
    if (!(J.flags & GPUCSL_JOB_PCALG_COMPAT)) return 1.0f;
    const double m = (double)(J.n - level - 3);
    return m > 0.0 ? (float)sqrt(m) : 0.0f;
}

__device__ __forceinline__ bool sepset_one_way(const Job &J) { return (J.flags & GPUCSL_JOB_PCALG_COMPAT) != 0; }

__device__ __forceinline__ int job_of(const int32_t *__restrict__ base, int nj, long long t) {
    int lo = 0, hi = nj;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (base[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// ----------------------------------------------------------------------------
// init: G = 1, sepsets = -1, zmin = +inf (gaussian_device_manager.py:106-133)
// ----------------------------------------------------------------------------
// Job R.job[k] takes init_blocks(p, max_level) workgroups.
__host__ __device__ inline int init_blocks(int p, int max_level) {
    const size_t work = (size_t)p * p * (max_level > 1 ? max_level : 1);
    const size_t b = (work + 256 * 16 - 1) / (256 * 16);
    return b < 1 ? 1 : (b > 1024 ? 1024 : (int)b);
}

__global__ void job_init_kernel(const Job *__restrict__ jobs, BlockRanges R) {
    int local;
    const Job &J = jobs[range_job(R, blockIdx.x, local)];
    const size_t pp = (size_t)J.p * J.p;
    const size_t stride = (size_t)init_blocks(J.p, J.max_level) * blockDim.x;
    const size_t t0 = (size_t)local * blockDim.x + threadIdx.x;
    for (size_t t = t0; t < pp; t += stride) {
        J.G[t] = 1;
        J.zmin[t] = __builtin_inff();
    }
    const size_t ns = pp * (size_t)J.max_level;
    for (size_t t = t0; t < ns; t += stride) J.sepsets[t] = -1;
}

// ----------------------------------------------------------------------------
// level 0: gpucsl/pc/cuda/gaussian_ci.cu:18-51
// ----------------------------------------------------------------------------
// Job R.job[k]: one workgroup per row and 256 columns, p * ceil(p / 256).
__global__ __launch_bounds__(256) void gauss_level0_kernel(const Job *__restrict__ jobs, float margin_k,
                                                           unsigned long long *__restrict__ st, BlockRanges R) {
    int local;
    const Job &J = jobs[range_job(R, blockIdx.x, local)];
    const int p = J.p;
    const int cbk = (p + 255) / 256;
    const int a = local / cbk;
    if (a >= p) return;
    const int b = (local % cbk) * 256 + threadIdx.x;
    const double thr = J.thr[0];
    const float thr32 = (float)thr;
    const float zscale = level_zscale(J, 0);
    uint8_t *__restrict__ Gm = J.G;
    unsigned long long n_re = 0, n_flip = 0;
    float mr = 0.0f;
    if (b < p) {
        if (b == a) {
            Gm[(size_t)a * p + a] = 0;
        } else if (b > a) {
            const float r = J.C32[(size_t)a * p + b];
            const float z = fabsf(atanhf(r));
            const float err = EPS32 * (fabsf(r) + z + 1e-30f) / fmaxf(1.0f - r * r, 1e-30f);
            const bool bad = !(z <= 3.0e38f);
            float zf;
            bool pass;
            if (bad || fabsf(z - thr32) <= margin_k * err) {
                const double z64 = ref64::gauss_z0(J.C64[(size_t)a * p + b]);
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
            if (!(zf != zf)) J.zmin[(size_t)a * p + b] = zf * zscale;
        }
    }
    // executed = total = useful = p(p-1)/2 per job; only the rare recheck
    // counters need per-wave atomics.
    if (local == 0 && threadIdx.x == 0) {
        const unsigned long long pairs = (unsigned long long)p * (p - 1) / 2;
        atomicAdd(&st[ST_EXEC], pairs);
        atomicAdd(&st[ST_TOTAL], pairs);
        atomicAdd(&st[ST_USEFUL], pairs);
    }
    if (__ballot(n_re != 0)) {
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

struct PlanArgs {
    unsigned int *done;            // blocks finished; null: scan only
    int nj;
    int32_t *baseA, *baseB;
    unsigned int *counters;
    unsigned long long *bstats;
    WorkItem *itemsA, *itemsB;
};
__device__ void jobprefix_body(const Job *__restrict__ jobs, int nj, int32_t *__restrict__ baseA,
                               int32_t *__restrict__ baseB, unsigned int *__restrict__ counters,
                               unsigned long long *__restrict__ bstats, int *sA, int *sB);

__device__ __forceinline__ int wave_incl_scan(int x) {
    const int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    return x;
}

// One job's level plan, by one block of NT threads (NT = 256 or 1024):
// exclusive scans of cntA/cntB into rowoffA/rowoffB and of the rows' work
// items into itemoffA/itemoffB; adds T_l and the max degree to the launch
// statistics.  A job whose max level is below `level` gets no work.  With
// `single` (a launch of one job) it also writes the work-item lists, the item
// bases and resets the launch counters, so the level needs no other planning.
//
// With `pre`, the compaction launch has already left each row's item counts
// in itemoffA/itemoffB and its pairs per item in rowoffA; the scans read them
// before overwriting them with the offsets.
template <int NT>
__device__ void scan_job(const Job &J, int level, const Team &team, unsigned long long *__restrict__ st,
                         const PlanArgs &plan, bool single, bool pre) {
    constexpr int NW = NT / 64;
    const int p = J.p;
    const bool active = level <= J.max_level;
    JobWs w = carve_job(J.ws, p);
    __shared__ int s_wsum[4][NW], s_woff[4][NW], s_tot[4];
    __shared__ unsigned long long s_T[NW];
    __shared__ int s_M[NW];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int per = (p + NT - 1) / NT;
    const int r0 = t * per;
    int r1 = r0 + per;
    if (r1 > p) r1 = p;
    int la = 0, lb = 0, lia = 0, lib = 0, md = 0;
    unsigned long long tt = 0;
    for (int r = r0; r < r1; ++r) {
        const int ca = active ? w.cntA[r] : 0, cb = active ? w.cntB[r] : 0;
        la += ca;
        lb += cb;
        const int d = w.deg[r];
        const uint64_t count = active ? binom_level(d - 1, level) : 0;
        if (pre) {
            lia += w.itemoffA[r];
            lib += w.itemoffB[r];
        } else if (count) {
            const int ppi = items_pairs(count, team);
            lia += (ca + ppi - 1) / ppi;
            lib += (cb + ppi - 1) / ppi;
        }
        if (active) {
            md = d > md ? d : md;
            tt += (unsigned long long)d * count;
        }
    }
    // block-wide exclusive scans of the four counts: wave scans, then a scan
    // of the NW wave totals
    const int ia = wave_incl_scan(la), ib = wave_incl_scan(lb), iia = wave_incl_scan(lia), iib = wave_incl_scan(lib);
    for (int off = 32; off >= 1; off >>= 1) {
        tt += __shfl_xor(tt, off);
        const int o = __shfl_xor(md, off);
        md = o > md ? o : md;
    }
    if (lane == 63) {
        s_wsum[0][wv] = ia;
        s_wsum[1][wv] = ib;
        s_wsum[2][wv] = iia;
        s_wsum[3][wv] = iib;
    }
    if (lane == 0) {
        s_T[wv] = tt;
        s_M[wv] = md;
    }
    __syncthreads();
    if (t < NW) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int x = s_wsum[k][t];
            const int own = x;
#pragma unroll
            for (int off = 1; off < NW; off <<= 1) {
                const int y = __shfl_up(x, off);
                if (t >= off) x += y;
            }
            s_woff[k][t] = x - own;
            if (t == NW - 1) s_tot[k] = x;
        }
    }
    __syncthreads();
    int ea = s_woff[0][wv] + ia - la, eb = s_woff[1][wv] + ib - lb;
    int eia = s_woff[2][wv] + iia - lia, eib = s_woff[3][wv] + iib - lib;
    for (int r = r0; r < r1; ++r) {
        const int ca = active ? w.cntA[r] : 0, cb = active ? w.cntB[r] : 0;
        int na = 0, nb = 0, ppi = 1;
        if (pre) {
            na = w.itemoffA[r];
            nb = w.itemoffB[r];
            ppi = w.rowoffA[r];
        } else {
            const uint64_t count = active ? binom_level(w.deg[r] - 1, level) : 0;
            if (count) {
                ppi = items_pairs(count, team);
                na = (ca + ppi - 1) / ppi;
                nb = (cb + ppi - 1) / ppi;
            }
        }
        w.rowoffA[r] = ea;
        w.rowoffB[r] = eb;
        w.itemoffA[r] = eia;
        w.itemoffB[r] = eib;
        ea += ca;
        eb += cb;
        if (na + nb) {
            if (single) {
                // one job: its item lists start at 0
                for (int k = 0; k < na; ++k) plan.itemsA[eia + k] = WorkItem{0, r, k * ppi, 0};
                for (int k = 0; k < nb; ++k) plan.itemsB[eib + k] = WorkItem{0, r, k * ppi, 0};
            }
            eia += na;
            eib += nb;
        }
    }
    if (t == 0) {
        w.rowoffA[p] = s_tot[0];
        w.rowoffB[p] = s_tot[1];
        w.itemoffA[p] = s_tot[2];
        w.itemoffB[p] = s_tot[3];
        // the level kernels read the |rho| threshold instead of evaluating
        // the FP64 tanh once per work item
        if (active && level < GPUCSL_MAX_LEVELS) w.rthr[level] = (float)tanh(J.thr[level]);
        if (st) {
            unsigned long long T = 0;
            int M = 0;
            for (int k = 0; k < NW; ++k) {
                T += s_T[k];
                M = s_M[k] > M ? s_M[k] : M;
            }
            if (T) atomicAdd(&st[ST_TOTAL], T);
            atomicMax(&st[ST_MAXDEG], (unsigned long long)M);
        }
    }
    if (!single) return;
    for (int k = t; k < NUM_COUNTERS; k += NT) plan.counters[k] = 0;
    for (int k = t; k < BSTAT_STRIPES * BSTAT_WORDS; k += NT) plan.bstats[k] = 0;
    if (t == 0) {
        plan.baseA[0] = 0;
        plan.baseA[1] = s_tot[2];
        plan.baseB[0] = 0;
        plan.baseB[1] = s_tot[3];
    }
}

// ----------------------------------------------------------------------------
// compaction: one wave per row (gpucsl/pc/cuda/helpers/graph_helpers.cu:35-56)
// nbr[row] = ascending neighbours (diagonal excluded), deg, and the split into
// neighbours below (cntB) and above (cntA) the row index.  In a planning
// launch (plan.done set) each row's wave also leaves its share of the plan for
// the scan launch: item counts in itemoffA/itemoffB and pairs per item in
// rowoffA, until the scans replace them with offsets.
// ----------------------------------------------------------------------------
__device__ __forceinline__ void compact_rows(const Job &J, int level, const Team &team, bool pre, int rblk) {
    const int p = J.p;
    const int row = rblk * 4 + (threadIdx.x >> 6);
    JobWs w = carve_job(J.ws, p);
    const bool active = level <= J.max_level;
    if (row < p) {
        const int lane = __lane_id();
        const uint8_t *g = J.G + (size_t)row * p;
        int32_t *out = w.nbr + (size_t)row * p;
        int base = 0, below = 0;
        // 1024 columns per group: the group's 16 loads go out before its first
        // store (the byte loads may alias the stores, so they are not reordered)
        constexpr int U = 16;
        for (int g0 = 0; g0 < p; g0 += 64 * U) {
            uint8_t v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int c = g0 + u * 64 + lane;
                v[u] = c < p ? g[c] : 0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int c0 = g0 + u * 64, c = c0 + lane;
                const bool e = v[u] != 0 && c != row;
                const uint64_t m = __ballot(e);
                if (e) out[base + __popcll(m & lanemask_lt())] = c;
                base += __popcll(m);
                if (c0 + 64 <= row)
                    below += __popcll(m);
                else if (c0 < row)
                    below += __popcll(m & ((1ull << (row - c0)) - 1));
            }
        }
        if (lane == 0) {
            w.deg[row] = base;
            w.cntB[row] = below;
            w.cntA[row] = base - below;
            if (pre) {
                const uint64_t count = active ? binom_level(base - 1, level) : 0;
                int ppi = 1, na = 0, nb = 0;
                if (count) {
                    ppi = items_pairs(count, team);
                    na = (base - below + ppi - 1) / ppi;
                    nb = (below + ppi - 1) / ppi;
                }
                w.itemoffA[row] = na;
                w.itemoffB[row] = nb;
                w.rowoffA[row] = ppi;
            }
        }
    }
}

// Job R.job[k]: (p + 3) / 4 workgroups of four rows.
__global__ __launch_bounds__(256) void compact_kernel(const Job *__restrict__ jobs, int level, Team team,
                                                      unsigned long long *__restrict__ st, PlanArgs plan,
                                                      BlockRanges R) {
    const bool pre = plan.done != nullptr;
    if (pre && blockIdx.x == 0 && threadIdx.x == 0) *plan.done = 0;
    int local;
    const Job &J = jobs[range_job(R, blockIdx.x, local)];
    compact_rows(J, level, team, pre, local);
}

// One dataset passed by value: a compaction that no scan follows needs no
// descriptor in device memory.
__global__ __launch_bounds__(256) void compact_one_kernel(Job one, int level, Team team) {
    compact_rows(one, level, team, false, blockIdx.x);
}

// Per job (one block each) scan_job; in a batch with a plan, the block that
// finishes last computes the job bases and resets the launch counters (the
// work-item lists follow in items_kernel, one block per job).
__global__ __launch_bounds__(1024) void scan_kernel(const Job *__restrict__ jobs, int level, Team team,
                                                    unsigned long long *__restrict__ st, PlanArgs plan) {
    const Job &J = jobs[blockIdx.x];
    __shared__ int sA[1024], sB[1024];  // job prefix of a batch (last block)
    const bool single = plan.done && plan.nj == 1;
    scan_job<1024>(J, level, team, st, plan, single, plan.done != nullptr);
    if (!plan.done || single) return;
    __shared__ int s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(plan.done, 1u) == (unsigned int)(plan.nj - 1);
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    jobprefix_body(jobs, plan.nj, plan.baseA, plan.baseB, plan.counters, plan.bstats, sA, sB);
}

// Work-item lists of a batch: block b writes job b's items {job, row, first
// pair} at the job's base plus the row's item offset.
__global__ __launch_bounds__(256) void items_kernel(const Job *__restrict__ jobs, int level, Team team, PlanArgs plan) {
    const int b = blockIdx.x;
    const Job &J = jobs[b];
    if (level > J.max_level) return;
    const int p = J.p;
    JobWs w = carve_job(J.ws, p);
    const int32_t ba = plan.baseA[b], bb = plan.baseB[b];
    for (int r = threadIdx.x; r < p; r += blockDim.x) {
        const uint64_t count = binom_level(w.deg[r] - 1, level);
        if (!count) continue;
        const int ppi = items_pairs(count, team);
        const int ia = w.itemoffA[r], ib = w.itemoffB[r];
        const int na = w.itemoffA[r + 1] - ia, nb = w.itemoffB[r + 1] - ib;
        for (int k = 0; k < na; ++k) plan.itemsA[ba + ia + k] = WorkItem{b, r, k * ppi, 0};
        for (int k = 0; k < nb; ++k) plan.itemsB[bb + ib + k] = WorkItem{b, r, k * ppi, 0};
    }
}

// Work-item bases of every job of the launch (exclusive prefix over jobs of
// itemoff[p]); resets the counters and the statistics stripes.  Run by one
// block of 1024 threads; sA, sB are 1024-int shared arrays.
__device__ void jobprefix_body(const Job *__restrict__ jobs, int nj, int32_t *__restrict__ baseA,
                               int32_t *__restrict__ baseB, unsigned int *__restrict__ counters,
                               unsigned long long *__restrict__ bstats, int *sA, int *sB) {
    const int t = threadIdx.x;
    if (t < NUM_COUNTERS) counters[t] = 0;
    if (t < BSTAT_STRIPES * BSTAT_WORDS) bstats[t] = 0;
    const int per = (nj + 1023) / 1024;
    const int j0 = t * per;
    const int j1 = min(nj, j0 + per);
    int la = 0, lb = 0;
    for (int b = j0; b < j1; ++b) {
        JobWs w = carve_job(jobs[b].ws, jobs[b].p);
        la += w.itemoffA[jobs[b].p];
        lb += w.itemoffB[jobs[b].p];
    }
    __syncthreads();
    sA[t] = la;
    sB[t] = lb;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int va = 0, vb = 0;
        if (t >= off) {
            va = sA[t - off];
            vb = sB[t - off];
        }
        __syncthreads();
        sA[t] += va;
        sB[t] += vb;
        __syncthreads();
    }
    int ea = sA[t] - la, eb = sB[t] - lb;
    for (int b = j0; b < j1; ++b) {
        baseA[b] = ea;
        baseB[b] = eb;
        JobWs w = carve_job(jobs[b].ws, jobs[b].p);
        ea += w.itemoffA[jobs[b].p];
        eb += w.itemoffB[jobs[b].p];
    }
    if (t == 1023) {
        baseA[nj] = sA[1023];
        baseB[nj] = sB[1023];
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
__device__ __forceinline__ void rho_from_values(float cij, const float (&ci)[L], const float (&cj)[L],
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
}

template <int L>
__device__ __forceinline__ float z32_from_values(float cij, const float (&ci)[L], const float (&cj)[L],
                                                 float (&c)[L][L], float &dmin, float &rho) {
    rho_from_values<L>(cij, ci, cj, c, dmin, rho);
    return fabsf(atanhf(rho));
}

// error bound of the FP32 partial correlation: rounding of C32 and of the
// (L+2)-step elimination, amplified by the smallest pivot
template <int L>
__device__ __forceinline__ float rho_error(float dmin) {
    // v_rcp_f32 (about 1 ulp) is ample for a bound scaled by the margin
    return EPS32 * (float)(L + 2) * (1.0f + __builtin_amdgcn_rcpf(fmaxf(dmin, 1e-30f)));
}

__device__ __forceinline__ bool rho_bad(float dmin, float ar) { return !(dmin > 1e-5f) || !(ar < 1.0f); }

// Fisher z of a non-negative |rho| (+inf for none counted)
__device__ __forceinline__ float z_of_rho(float ar) { return ar < 1.0f ? atanhf(ar) : __builtin_inff(); }

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
// level l >= 1, FAST phase.  A block of 256 threads takes one work item: pairs
// q0..q1-1 of row i on this side (j > i on side A, j < i on side B).  The
// pairs of a row share the subset candidates adj(i)\{j}, C(deg-1, l) per pair,
// so the lanes that test one pair together (a "team") are sized by that count:
// teams of tm.gs_min lanes or of one wave pull the item's pairs one by one from
// an LDS counter; beyond that the whole block works through the pairs in turn.
// A team walks the pair's subset ranks in increasing order and stops at the
// first passing test, so the removing sepset is the first in pcalg order and
// zmin covers exactly the tests up to it.  What the tests read is staged in LDS
// by the whole block (MODE, a template parameter so that LDS reads compile to
// ds_read and global reads to global_load):
//   0 FULL  (l >= 2) C[A, A] with A = adj(i) + [i]: the tests read LDS only;
//   1 ROW   neighbour ids, C[i, adj(i)] and diag C[s, s]; C[j, S] and the
//           off-diagonal C[S, S] come from global memory (L2);
//   2 none  (rows beyond the LDS budget) everything from global memory.
// Decisions are taken on the partial correlation rho itself: a test passes
// when |rho| < tanh(threshold), and one within margin * (its FP32 error bound)
// of that value sends the pair, from the start of that iteration, to the FP64
// queue.  The smallest |rho| a team counts becomes zmin = atanh(min |rho|).
// gpucsl/pc/cuda/gaussian_ci.cu:66-120, set_helpers.cu:25-71,
// graph_helpers.cu:81-133.
// ----------------------------------------------------------------------------
// Loads through the global address space (global_load instead of flat_load:
// no LDS counter to wait on).  Only for pointers into device memory.
typedef const __attribute__((address_space(1))) float gfloat_t;
typedef const __attribute__((address_space(1))) int gint_t;
__device__ __forceinline__ float ldg(const float *ptr) { return *(gfloat_t *)(ptr); }
__device__ __forceinline__ int ldg(const int *ptr) { return *(gint_t *)(ptr); }
// Byte loads and the stores of the per-pair results likewise: flat accesses
// also count against the LDS wait counter, so the next pair's LDS claim would
// wait for them.
typedef __attribute__((address_space(1))) uint8_t gu8_t;
typedef __attribute__((address_space(1))) int32_t gi32_t;
__device__ __forceinline__ uint8_t ldg(const uint8_t *ptr) { return *(const gu8_t *)(ptr); }
__device__ __forceinline__ void stg(uint8_t *ptr, uint8_t v) { *(gu8_t *)(ptr) = v; }
__device__ __forceinline__ void stg(int32_t *ptr, int32_t v) { *(gi32_t *)(ptr) = v; }

struct ItemCtx {
    const Job *J;
    int p, i, d, D, Nn, cnt, jbase, q0, q1, job, side;
    uint64_t count;
    float rthr, margin_k, zscale;
    const int32_t *gnb, *rowoff;
    const float *C32;
    const int *snb;    // LDS
    const float *sf;   // LDS
    const uint64_t *tab;  // LDS binomial table C(m, k), k = 1..L, m = 0..Nn; null: computed
    float *prow;          // LDS, ROW mode: C[j, adj(i)] of the pair each team works on (d floats per team)
    int prow_slot0;       // team slot (threadIdx.x / team) of prow's first row
};

// Levels 1 and 2, ROW mode: a pair's C[j, k] is read from global memory (L2)
// by the lane whose test needs it instead of staging the whole row
// C[j, adj(i)] in LDS first; most pairs end within their first team
// iterations, so staging would gather up to deg values to use a few dozen.
template <int L>
constexpr bool lazy_pair_row() { return L <= 2; }

template <int L, int MODE>
__device__ __forceinline__ float rho_test(const ItemCtx &c, int jpos, int j, const int (&sp)[L], bool &near,
                                          const float *rowJ_lds) {
    float cij, ci[L], cj[L], cm[L][L];
    if constexpr (MODE == 0) {
        const float *sf = c.sf;
        const int D = c.D, d = c.d;
        cij = sf[d * D + jpos];
#pragma unroll
        for (int a = 0; a < L; ++a) {
            ci[a] = sf[d * D + sp[a]];
            cj[a] = sf[jpos * D + sp[a]];
#pragma unroll
            for (int b = 0; b <= a; ++b) cm[a][b] = sf[sp[a] * D + sp[b]];
        }
    } else if constexpr (MODE == 1) {
        const float *rowI = c.sf, *diag = c.sf + c.d;
        int S[L];
#pragma unroll
        for (int a = 0; a < L; ++a) S[a] = c.snb[sp[a]];
        cij = rowI[jpos];
#pragma unroll
        for (int a = 0; a < L; ++a) {
            ci[a] = rowI[sp[a]];
            cj[a] = lazy_pair_row<L>() ? ldg(c.C32 + (size_t)j * c.p + S[a]) : rowJ_lds[sp[a]];
#pragma unroll
            for (int b = 0; b < a; ++b) cm[a][b] = ldg(c.C32 + (size_t)S[a] * c.p + S[b]);
            cm[a][a] = diag[sp[a]];
        }
    } else {
        const float *__restrict__ C = c.C32;
        const int p = c.p;
        int S[L];
#pragma unroll
        for (int a = 0; a < L; ++a) S[a] = ldg(c.gnb + sp[a]);
        cij = ldg(C + (size_t)c.i * p + j);
#pragma unroll
        for (int a = 0; a < L; ++a) {
            ci[a] = ldg(C + (size_t)c.i * p + S[a]);
            cj[a] = ldg(C + (size_t)j * p + S[a]);
#pragma unroll
            for (int b = 0; b <= a; ++b) cm[a][b] = ldg(C + (size_t)S[a] * p + S[b]);
        }
    }
    float dmin, rho;
    rho_from_values<L>(cij, ci, cj, cm, dmin, rho);
    const float ar = fabsf(rho);
    near = rho_bad(dmin, ar) || fabsf(ar - c.rthr) <= c.margin_k * rho_error<L>(dmin);
    return ar;
}

template <int L, int MODE>
__device__ __forceinline__ int nbr_at(const ItemCtx &c, int k) {
    if constexpr (MODE == 2)
        return ldg(c.gnb + k);
    else
        return c.snb[k];
}

template <int L, int MODE>
__device__ __forceinline__ void write_sepset(const ItemCtx &c, int j, const int (&sp)[L]) {
    const Job &J = *c.J;
    const int p = c.p, i = c.i;
    stg(J.G + (size_t)i * p + j, (uint8_t)0);
    stg(J.G + (size_t)j * p + i, (uint8_t)0);
    int32_t *sep = J.sepsets;
    const int maxL = J.max_level;
    const bool both = !sepset_one_way(J);
#pragma unroll
    for (int k = 0; k < L; ++k) {
        const int v = nbr_at<L, MODE>(c, sp[k]);
        stg(sep + ((size_t)i * p + j) * maxL + k, v);
        if (both) stg(sep + ((size_t)j * p + i) * maxL + k, v);
    }
}

// zmin[a, b] = min(zmin[a, b], z) for a non-negative z (NaN leaves it as it
// is): non-negative floats order like their bits as unsigned integers, so one
// atomic without a return value does it and the wave does not wait for a load
// (through the global address space: a flat atomic would also count against
// the LDS wait counter)
__device__ __forceinline__ void zmin_update(float *zp, float z) {
    typedef __attribute__((address_space(1))) unsigned int guint_t;
    __hip_atomic_fetch_min((guint_t *)(zp), __float_as_uint(z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void update_zmin(const Job &J, int p, int i, int j, float z) {
    const int a = i < j ? i : j, b = i < j ? j : i;
    zmin_update(J.zmin + (size_t)a * p + b, z);
}

// C[j, adj(i)] into LDS by `lanes` lanes (lane index `l`)
__device__ __forceinline__ void stage_pair_row(const ItemCtx &c, int j, float *dst, int l, int lanes) {
    const float *__restrict__ rowJ = c.C32 + (size_t)j * c.p;
    for (int k0 = 0; k0 < c.d; k0 += lanes * 4) {
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + u * lanes + l;
            v[u] = k < c.d ? ldg(rowJ + c.snb[k]) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + u * lanes + l;
            if (k < c.d) dst[k] = v[u];
        }
    }
}

// Minimum over each team's lanes (team a power of two, 1..64, lanes aligned
// to the team size) of non-negative floats, valid in every lane of the team.
// Non-negative floats (+inf included) order like their bit patterns as
// unsigned integers.  Within 16 lanes the steps are DPP moves (quad
// permutes, half-row and row mirrors), which stay inside the team; wider
// teams add lane swaps through ds_bpermute.
__device__ __forceinline__ unsigned int umin_u32(unsigned int a, unsigned int b) { return a < b ? a : b; }

template <int CTRL>
__device__ __forceinline__ unsigned int dpp_u32(unsigned int u) {
    return (unsigned int)__builtin_amdgcn_update_dpp((int)u, (int)u, CTRL, 0xF, 0xF, false);
}

__device__ __forceinline__ float team_min(float x, int team) {
    unsigned int v = __float_as_uint(x);
    if (team >= 2) v = umin_u32(v, dpp_u32<0xB1>(v));   // quad_perm [1,0,3,2]
    if (team >= 4) v = umin_u32(v, dpp_u32<0x4E>(v));   // quad_perm [2,3,0,1]
    if (team >= 8) v = umin_u32(v, dpp_u32<0x141>(v));  // row_half_mirror
    if (team >= 16) v = umin_u32(v, dpp_u32<0x140>(v)); // row_mirror
    if (team >= 32) v = umin_u32(v, (unsigned int)__shfl_xor((int)v, 16));
    if (team >= 64) v = umin_u32(v, (unsigned int)__shfl_xor((int)v, 32));
    return __uint_as_float(v);
}

// Level 1 has C(d - 1, 1) = d - 1 < 2^31 subsets per pair, so its loops
// count in 32-bit arithmetic; higher levels in 64 bits.
template <int L>
using rank_t = typename std::conditional<(L == 1), uint32_t, uint64_t>::type;

template <int L>
__device__ __forceinline__ void subset_of(const ItemCtx &c, uint64_t rank, int jpos, int (&sp)[L]) {
    if constexpr (L == 1) {
        // rank r < C(d - 1, 1): the r-th of the row's neighbours other than j
        const int r = (int)rank;
        sp[0] = r + (r >= jpos);
        GPUCSL_DCHECK(sp[0] >= 0 && sp[0] < c.d && sp[0] != jpos);
        return;
    }
    int pos[L];
    if (L >= 2 && c.tab)
        unrank_lex_tab<L>(c.Nn, rank, pos, c.tab, c.Nn + 1);
    else
        unrank_lex<L>(c.Nn, rank, pos);
#pragma unroll
    for (int k = 0; k < L; ++k) {
        const int idx = pos[k] - 1;
        sp[k] = idx + (idx >= jpos);
        GPUCSL_DCHECK(sp[k] >= 0 && sp[k] < c.d && sp[k] != jpos);
    }
}

// the whole block works through the item's pairs in turn
// Hands a pair with a near test to the FP64 queue.  A side-A pair is marked
// pending (G = 2 both ways) so that side B leaves it to the SLOW kernel, which
// finishes side A and then, if no test removed the pair, runs its side B and
// sets G back to 1.
__device__ __forceinline__ void enqueue_slow(const ItemCtx &c, int j, const SlowItem &item,
                                             SlowItem *__restrict__ queue, unsigned int *__restrict__ qcount) {
    const unsigned int slot = atomicAdd(qcount, 1u);
    queue[slot] = item;
    if (c.side == 0) {
        stg(c.J->G + (size_t)c.i * c.p + j, (uint8_t)2);
        stg(c.J->G + (size_t)j * c.p + c.i, (uint8_t)2);
    }
}

template <int L, int MODE>
__device__ void block_pairs(const ItemCtx &c, SlowItem *__restrict__ queue, unsigned int *__restrict__ qcount,
                            unsigned int &n_exec, unsigned int &n_useful,
                            unsigned long long (&s_pm)[4], unsigned long long (&s_nm)[4], float (&s_r)[4]) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint8_t *__restrict__ Gm = c.J->G;
    for (int q = c.q0; q < c.q1; ++q) {
        const int jpos = c.jbase + q;
        const int j = nbr_at<L, MODE>(c, jpos);
        if (c.side == 1 && ldg(Gm + (size_t)c.i * c.p + j) != 1) continue;  // removed, or pending in the SLOW kernel
        if constexpr (MODE == 1 && !lazy_pair_row<L>()) {
            __syncthreads();  // the previous pair is done with the row
            stage_pair_row(c, j, c.prow, tid, 256);
            __syncthreads();
        }
        float rl = __builtin_inff();  // min |rho| over the tests counted so far
        const rank_t<L> cnt = (rank_t<L>)c.count;
        for (rank_t<L> iter = 0;; ++iter) {
            const rank_t<L> rank = iter * 256 + tid;
            const bool act = rank < cnt;
            bool pass = false, near = false;
            float ar = __builtin_inff();
            int sp[L];
            if (act) {
                subset_of<L>(c, rank, jpos, sp);
                ar = rho_test<L, MODE>(c, jpos, j, sp, near, c.prow);
                pass = ar < c.rthr;
            }
            const unsigned long long pm = __ballot(pass), nm = __ballot(near);
            if (lane == 0) {
                s_pm[wave] = pm;
                s_nm[wave] = nm;
            }
            __syncthreads();
            const bool any_near = (s_nm[0] | s_nm[1] | s_nm[2] | s_nm[3]) != 0;
            int P = -1;
#pragma unroll
            for (int w2 = 0; w2 < 4; ++w2)
                if (P < 0 && s_pm[w2]) P = w2 * 64 + __ffsll(s_pm[w2]) - 1;
            bool finish;
            if (any_near) {
                finish = true;
            } else {
                if (act) ++n_exec;
                if (P >= 0) {
                    if (act && tid <= P) {
                        ++n_useful;
                        rl = fminf(rl, ar);
                    }
                    if (tid == P) write_sepset<L, MODE>(c, j, sp);
                    finish = true;
                } else {
                    if (act) {
                        ++n_useful;
                        rl = fminf(rl, ar);
                    }
                    finish = (iter + 1) * 256 >= cnt;
                }
            }
            if (finish) {
                const float m = team_min(rl, 64);
                if (lane == 0) s_r[wave] = m;
                __syncthreads();
                if (tid == 0) {
                    const float zb = z_of_rho(fminf(fminf(s_r[0], s_r[1]), fminf(s_r[2], s_r[3])));
                    if (any_near) {
                        enqueue_slow(c, j, SlowItem{iter * 256, jpos, c.job, zb, c.i}, queue, qcount);
                    } else {
                        update_zmin(*c.J, c.p, c.i, j, zb * c.zscale);
                    }
                }
            }
            __syncthreads();  // s_pm / s_nm / s_r are reused
            if (finish) break;
        }
    }
}

// teams of `team` lanes inside a wave pull pairs from s_next
template <int L, int MODE>
__device__ void team_pairs(const ItemCtx &c, int team, int *s_next, SlowItem *__restrict__ queue,
                           unsigned int *__restrict__ qcount, unsigned int &n_exec,
                           unsigned int &n_useful) {
    const int lane = threadIdx.x & 63;
    const int gbase = lane & ~(team - 1);
    const int gl = lane & (team - 1);
    const uint64_t gmask = team == 64 ? ~0ull : (((1ull << team) - 1ull) << gbase);
    const uint8_t *__restrict__ Gm = c.J->G;
    while (true) {
        int q = 0;
        if (gl == 0) q = atomicAdd(s_next, 1);
        q = __shfl(q, gbase);
        if (q >= c.q1) break;
        const int jpos = c.jbase + q;
        const int j = nbr_at<L, MODE>(c, jpos);
        if (c.side == 1 && ldg(Gm + (size_t)c.i * c.p + j) != 1) continue;  // removed, or pending in the SLOW kernel
        float *myrow = c.prow + (size_t)(threadIdx.x / team - c.prow_slot0) * c.d;
        if constexpr (MODE == 1 && !lazy_pair_row<L>()) {
            // the team's lanes stage C[j, adj(i)]; LDS accesses of one wave
            // are ordered, the fences keep the compiler from moving them
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            stage_pair_row(c, j, myrow, gl, team);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        float rl = __builtin_inff();
        const rank_t<L> cnt = (rank_t<L>)c.count;
        for (rank_t<L> iter = 0;; ++iter) {
            const rank_t<L> rank = iter * (rank_t<L>)team + (rank_t<L>)gl;
            const bool act = rank < cnt;
            bool pass = false, near = false;
            float ar = __builtin_inff();
            int sp[L];
            if (act) {
                subset_of<L>(c, rank, jpos, sp);
                ar = rho_test<L, MODE>(c, jpos, j, sp, near, myrow);
                pass = ar < c.rthr;
            }
            if (__ballot(near) & gmask) {
                rl = team_min(rl, team);
                if (gl == 0) enqueue_slow(c, j, SlowItem{iter * team, jpos, c.job, z_of_rho(rl), c.i}, queue, qcount);
                break;
            }
            if (act) ++n_exec;
            const uint64_t pm = __ballot(pass) & gmask;
            bool finish;
            if (pm) {
                const int P = __ffsll((unsigned long long)pm) - 1 - gbase;
                if (act && gl <= P) {
                    ++n_useful;
                    rl = fminf(rl, ar);
                }
                if (gl == P) write_sepset<L, MODE>(c, j, sp);
                finish = true;
            } else {
                if (act) {
                    ++n_useful;
                    rl = fminf(rl, ar);
                }
                finish = (iter + 1) * (rank_t<L>)team >= cnt;
            }
            if (finish) {
                rl = team_min(rl, team);
                if (gl == 0) update_zmin(*c.J, c.p, c.i, j, z_of_rho(rl) * c.zscale);
                break;
            }
        }
    }
}

template <int L>
__global__ __launch_bounds__(256) void gauss_level_fast_kernel(
    const Job *__restrict__ jobs, int nj, const int32_t *__restrict__ itembase, const WorkItem *__restrict__ items,
    int side, Team tm, int WORDS, float margin_k, SlowItem *__restrict__ queue, unsigned int *__restrict__ qcount,
    unsigned long long *__restrict__ bstats) {
    extern __shared__ int lds_words[];
    __shared__ unsigned long long s_pm[4], s_nm[4];
    __shared__ float s_r[4];
    __shared__ int s_next;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nitems = itembase[nj];
    unsigned int n_exec = 0, n_useful = 0;  // per lane; widened for the block totals

    for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
        const WorkItem it = items[item];
        const Job &J = jobs[it.job];
        ItemCtx c;
        c.J = &J;
        c.p = J.p;
        c.i = it.row;
        c.job = it.job;
        c.side = side;
        JobWs jw = carve_job(J.ws, c.p);
        c.rowoff = side == 0 ? jw.rowoffA : jw.rowoffB;
        c.C32 = J.C32;
        c.rthr = ldg(jw.rthr + L);
        c.margin_k = margin_k;
        c.zscale = level_zscale(J, L);
        c.d = ldg(jw.deg + c.i);
        c.cnt = ldg(c.rowoff + c.i + 1) - ldg(c.rowoff + c.i);
        c.Nn = c.d - 1;
        c.count = binom_k<L>(c.Nn);
        const int team = team_lanes(c.count, tm);
        c.q0 = it.q0;
        GPUCSL_DCHECK(c.i >= 0 && c.i < c.p && c.d < c.p && c.cnt <= c.d && c.q0 < c.cnt);
        c.q1 = min(c.cnt, c.q0 + items_pairs(c.count, tm));
        c.jbase = side == 0 ? c.d - c.cnt : 0;
        c.D = c.d + 1;
        const int d = c.d, D = c.D, i = c.i, p = c.p;
        const int teams = 256 / team;
        // ROW mode: neighbour ids, C[i, adj], diag, and one pair row per team
        const int mode =
            (L >= 2 && D * D + d <= WORDS) ? 0 : ((lazy_pair_row<L>() ? 3 : 3 + teams) * d <= WORDS ? 1 : 2);
        c.gnb = jw.nbr + (size_t)i * p;
        int *snb = lds_words;
        float *sf = reinterpret_cast<float *>(lds_words + d);
        c.snb = snb;
        c.sf = sf;
        c.prow = sf + 2 * d;
        c.prow_slot0 = 0;
        // binomial table after the staged data, when it fits
        const int tab_off = ((d + (mode == 0 ? D * D : (2 + teams) * d)) + 1) & ~1;
        const bool use_tab = L >= 2 && mode != 2 && c.Nn >= L && tab_off + 2 * L * (c.Nn + 1) <= WORDS;
        uint64_t *tab = use_tab ? reinterpret_cast<uint64_t *>(lds_words + tab_off) : nullptr;
        c.tab = tab;
        const float *__restrict__ C32 = c.C32;

        __syncthreads();  // the previous item is done with LDS
        if (tid == 0) s_next = c.q0;
        if (mode != 2)
            for (int k = tid; k < d; k += 256) snb[k] = ldg(c.gnb + k);
        if (use_tab)
            for (int m = tid; m <= c.Nn; m += 256) {
                uint64_t b = 1;  // C(m, k) = C(m, k-1) (m - k + 1) / k, exact at every step
#pragma unroll
                for (int k = 1; k <= L; ++k) {
                    b = m >= k ? b * (uint64_t)(m - k + 1) / (uint64_t)k : 0;
                    tab[(k - 1) * (c.Nn + 1) + m] = b;
                }
            }
        __syncthreads();
        if (mode == 0) {
            // sub[a * D + b] = C[A[a], A[b]]: wave w stages rows a = w, w + 4, ...,
            // four rows (eight loads per lane) per batch; D <= 90 < 128
            for (int a0 = wave; a0 < D; a0 += 16) {
                float v[4][2];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int a = a0 + 4 * u;
                    const int va = a < d ? snb[a] : i;
#pragma unroll
                    for (int cc = 0; cc < 2; ++cc) {
                        const int b = cc * 64 + lane;
                        const int vb = b < d ? snb[b] : i;
                        v[u][cc] = (a < D && b < D) ? ldg(C32 + (size_t)va * p + vb) : 0.0f;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int a = a0 + 4 * u;
#pragma unroll
                    for (int cc = 0; cc < 2; ++cc) {
                        const int b = cc * 64 + lane;
                        if (a < D && b < D) sf[a * D + b] = v[u][cc];
                    }
                }
            }
        } else if (mode == 1) {
            for (int k0 = 0; k0 < d; k0 += 256 * 2) {
                float v[2], w[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int k = k0 + u * 256 + tid;
                    const int vv = k < d ? snb[k] : 0;
                    v[u] = k < d ? ldg(C32 + (size_t)i * p + vv) : 0.0f;
                    w[u] = k < d ? ldg(C32 + (size_t)vv * p + vv) : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int k = k0 + u * 256 + tid;
                    if (k < d) {
                        sf[k] = v[u];
                        sf[d + k] = w[u];
                    }
                }
            }
        }
        __syncthreads();

        if (team == 256) {
            if (mode == 0)
                block_pairs<L, 0>(c, queue, qcount, n_exec, n_useful, s_pm, s_nm, s_r);
            else if (mode == 1)
                block_pairs<L, 1>(c, queue, qcount, n_exec, n_useful, s_pm, s_nm, s_r);
            else
                block_pairs<L, 2>(c, queue, qcount, n_exec, n_useful, s_pm, s_nm, s_r);
        } else {
            if (mode == 0)
                team_pairs<L, 0>(c, team, &s_next, queue, qcount, n_exec, n_useful);
            else if (mode == 1)
                team_pairs<L, 1>(c, team, &s_next, queue, qcount, n_exec, n_useful);
            else
                team_pairs<L, 2>(c, team, &s_next, queue, qcount, n_exec, n_useful);
        }
    }
    // block totals -> one statistics stripe
    unsigned long long t_exec = n_exec, t_useful = n_useful;
    for (int off = 32; off >= 1; off >>= 1) {
        t_exec += __shfl_xor(t_exec, off);
        t_useful += __shfl_xor(t_useful, off);
    }
    __syncthreads();
    if (lane == 0) {
        s_pm[wave] = t_exec;
        s_nm[wave] = t_useful;
    }
    __syncthreads();
    if (tid == 0) {
        const unsigned long long e = s_pm[0] + s_pm[1] + s_pm[2] + s_pm[3];
        const unsigned long long u = s_nm[0] + s_nm[1] + s_nm[2] + s_nm[3];
        if (e) {
            unsigned long long *dst = bstats + (size_t)(blockIdx.x % BSTAT_STRIPES) * BSTAT_WORDS;
            atomicAdd(dst + BS_EXEC, e);
            atomicAdd(dst + BS_USEFUL, u);
        }
    }
}

// Level l >= 1, FAST phase, one wave per work item: every wave of a block
// takes whole items on its own and its teams (gs_min lanes, or 64) pull the
// item's pairs; no block barriers, so a wave never waits for another.  A wave
// stages the row in its own LDS region of WAVE_WORDS words when it fits (ROW
// mode: neighbour ids, C[i, adj(i)], the diagonal, and one pair row per team),
// and otherwise reads C from global memory (L2).
template <int L>
__global__ __launch_bounds__(256) void gauss_level_wave_kernel(
    const Job *__restrict__ jobs, int nj, const int32_t *__restrict__ itembase, const WorkItem *__restrict__ items,
    int side, Team tm, int WAVE_WORDS, float margin_k, SlowItem *__restrict__ queue,
    unsigned int *__restrict__ qcount, unsigned long long *__restrict__ bstats) {
    extern __shared__ int lds_words[];
    __shared__ int s_next[4];
    __shared__ unsigned long long s_e[4], s_u[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nitems = itembase[nj];
    const int nwaves = gridDim.x * 4;
    unsigned int n_exec = 0, n_useful = 0;  // per lane; widened for the block totals
    for (int item = blockIdx.x * 4 + wave; item < nitems; item += nwaves) {
        const WorkItem it = items[item];
        const Job &J = jobs[it.job];
        ItemCtx c;
        c.J = &J;
        c.p = J.p;
        c.i = it.row;
        c.job = it.job;
        c.side = side;
        JobWs jw = carve_job(J.ws, c.p);
        c.rowoff = side == 0 ? jw.rowoffA : jw.rowoffB;
        c.C32 = J.C32;
        c.rthr = ldg(jw.rthr + L);
        c.margin_k = margin_k;
        c.zscale = level_zscale(J, L);
        c.d = ldg(jw.deg + c.i);
        c.cnt = ldg(c.rowoff + c.i + 1) - ldg(c.rowoff + c.i);
        c.Nn = c.d - 1;
        c.count = binom_k<L>(c.Nn);
        c.q0 = it.q0;
        c.q1 = min(c.cnt, c.q0 + items_pairs(c.count, tm));
        c.jbase = side == 0 ? c.d - c.cnt : 0;
        c.D = c.d + 1;
        c.gnb = jw.nbr + (size_t)c.i * c.p;
        c.snb = nullptr;
        c.sf = nullptr;
        c.tab = nullptr;
        c.prow = nullptr;
        c.prow_slot0 = 0;
        const int team = team_lanes(c.count, tm);
        const int teams_w = 64 / team;
        const int d = c.d;
        // the previous item's LDS reads and s_next are done before they change
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) s_next[wave] = c.q0;
        if ((lazy_pair_row<L>() ? 3 : 3 + teams_w) * d <= WAVE_WORDS) {
            int *snb = lds_words + (size_t)wave * WAVE_WORDS;
            float *sf = reinterpret_cast<float *>(snb + d);
            for (int k = lane; k < d; k += 64) snb[k] = ldg(c.gnb + k);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int k = lane; k < d; k += 64) {
                const int v = snb[k];
                sf[k] = ldg(c.C32 + (size_t)c.i * c.p + v);
                sf[d + k] = ldg(c.C32 + (size_t)v * c.p + v);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            c.snb = snb;
            c.sf = sf;
            c.prow = sf + 2 * d;
            c.prow_slot0 = wave * teams_w;
            team_pairs<L, 1>(c, team, &s_next[wave], queue, qcount, n_exec, n_useful);
        } else {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            team_pairs<L, 2>(c, team, &s_next[wave], queue, qcount, n_exec, n_useful);
        }
        __builtin_amdgcn_wave_barrier();  // every team is done with s_next before lane 0 resets it
    }
    unsigned long long t_exec = n_exec, t_useful = n_useful;
    for (int off = 32; off >= 1; off >>= 1) {
        t_exec += __shfl_xor(t_exec, off);
        t_useful += __shfl_xor(t_useful, off);
    }
    if (lane == 0) {
        s_e[wave] = t_exec;
        s_u[wave] = t_useful;
    }
    __syncthreads();
    if (tid == 0) {
        const unsigned long long e = s_e[0] + s_e[1] + s_e[2] + s_e[3];
        const unsigned long long u = s_u[0] + s_u[1] + s_u[2] + s_u[3];
        if (e) {
            unsigned long long *dst = bstats + (size_t)(blockIdx.x % BSTAT_STRIPES) * BSTAT_WORDS;
            atomicAdd(dst + BS_EXEC, e);
            atomicAdd(dst + BS_USEFUL, u);
        }
    }
}

// ----------------------------------------------------------------------------
// level l >= 1, SLOW phase, after both FAST sides: finishes the pairs the FAST
// kernels queued, from the queued subset rank on, rechecking near-threshold
// FP32 tests in FP64 with the reference's operation order.  A side-A pair
// (pending, G = 2) that no side-A test removes goes on with its side B from
// rank 0 in the same lane group and gets G = 1 back.  Items: queue A, then
// queue B.  Work: a static share per lane group, then 8 striped counters
// chosen by block index (no single hot atomic).
// ----------------------------------------------------------------------------
template <int L, int GS>
__global__ __launch_bounds__(256) void gauss_slow_kernel(
    const Job *__restrict__ jobs, float margin_k, unsigned int *__restrict__ stripe_counters,
    const SlowItem *__restrict__ queueA, const unsigned int *__restrict__ qcountA,
    const SlowItem *__restrict__ queueB, const unsigned int *__restrict__ qcountB,
    const unsigned long long *__restrict__ bstats, unsigned long long *__restrict__ st) {
    const int lane = __lane_id();
    const int gbase = lane & ~(GS - 1);
    const int gl = lane & (GS - 1);
    const uint64_t gmask = (GS == 64) ? ~0ull : (((1ull << GS) - 1ull) << gbase);
    const int totalA = (int)*qcountA;
    const int total = totalA + (int)*qcountB;

    constexpr int GPB = 256 / GS;
    const int groups = gridDim.x * GPB;
    const int gid = blockIdx.x * GPB + (threadIdx.x / GS);
    const int S_static = total <= groups ? 1 : total / groups;
    const int dyn_base = total <= groups ? total : S_static * groups;
    const int nstripes = gridDim.x < STRIPES ? gridDim.x : STRIPES;
    const int stripe = blockIdx.x % nstripes;
    const int rem = total - dyn_base;
    const int stripe_len = rem > stripe ? (rem - stripe + nstripes - 1) / nstripes : 0;
    unsigned int *counter = stripe_counters + stripe;
    int k_static = 0;

    const Job *J = nullptr;
    int p = 0, i = 0, j = 0, jpos = 0, Nn = 0, side = 0;
    const int32_t *gnb = nullptr;
    double thr = 0.0;
    float rthr = 0.0f;
    uint64_t count = 0, next = 0;
    float zl = __builtin_inff();
    bool have = false, done = false, restore = false;
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
                    const unsigned int cur = __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
                side = t < totalA ? 0 : 1;
                const SlowItem qi = side == 0 ? queueA[t] : queueB[t - totalA];
                J = &jobs[qi.job];
                p = J->p;
                JobWs jw = carve_job(J->ws, p);
                i = qi.row;
                jpos = qi.jpos;
                const int d = jw.deg[i];
                gnb = jw.nbr + (size_t)i * p;
                GPUCSL_DCHECK(i >= 0 && i < p && jpos >= 0 && jpos < d);
                j = gnb[jpos];
                Nn = d - 1;
                count = binom_k<L>(Nn);
                next = qi.start;
                zl = gl == 0 ? qi.zl : __builtin_inff();
                thr = J->thr[L];
                rthr = ldg(jw.rthr + L);
                have = count > 0;
            }
        }
        if (__ballot(!done) == 0) break;
        if (!have) continue;

        const uint64_t rank = next + gl;
        const bool act = rank < count;
        bool pass = false;
        float zf = __builtin_inff();
        int S[L];
        if (act) {
            int pos[L];
            unrank_lex<L>(Nn, rank, pos);
#pragma unroll
            for (int c = 0; c < L; ++c) {
                int idx = pos[c] - 1;
                S[c] = gnb[idx + (idx >= jpos)];
            }
            float dmin, rho;
            z32_test<L>(J->C32, p, i, j, S, dmin, rho);
            const float ar = fabsf(rho);
            const float err = rho_error<L>(dmin);
            const bool bad = rho_bad(dmin, ar);
            if (bad || fabsf(ar - rthr) <= margin_k * err) {
                const double z64 = ref64::gauss_z<L>(J->C64, p, i, j, S);
                pass = z64 < thr;
                zf = (float)z64;
                ++n_re;
                if (!bad) {
                    n_flip += (pass != (ar < rthr));
                    mr = fmaxf(mr, (float)(fabs((double)ar - tanh(z64)) / (double)err));
                }
            } else {
                pass = ar < rthr;
                zf = z_of_rho(ar);
            }
            ++n_exec;
        }
        const uint64_t pm = __ballot(pass) & gmask;
        bool finish = false;
        if (pm) {
            const int P = __ffsll((unsigned long long)pm) - 1 - gbase;
            if (act && gl <= P) {
                ++n_useful;
                if (zf < zl) zl = zf;
            }
            if (gl == P) {
                J->G[(size_t)i * p + j] = 0;
                J->G[(size_t)j * p + i] = 0;
                const int maxL = J->max_level;
#pragma unroll
                for (int c = 0; c < L; ++c) {
                    J->sepsets[((size_t)i * p + j) * maxL + c] = S[c];
                    if (!sepset_one_way(*J)) J->sepsets[((size_t)j * p + i) * maxL + c] = S[c];
                }
            }
            finish = true;
        } else {
            if (act) {
                ++n_useful;
                if (zf < zl) zl = zf;
            }
            next += GS;
            finish = next >= count;
        }
        if (finish) {
            zl = team_min(zl, GS);
            if (gl == 0) {
                const int a = i < j ? i : j, b = i < j ? j : i;
                zmin_update(J->zmin + (size_t)a * p + b, zl * level_zscale(*J, L));
            }
            have = false;
            if (side == 0 && !pm) {
                // side A kept the pending pair: its side B, row j, partner i
                // (one of the row's neighbours below j, found by bisection)
                const int row = j, partner = i;
                JobWs jw = carve_job(J->ws, p);
                const int d = jw.deg[row];
                gnb = jw.nbr + (size_t)row * p;
                int lo = 0, hi = d;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (gnb[mid] <= partner) lo = mid; else hi = mid;
                }
                i = row;
                j = partner;
                jpos = lo;
                side = 1;
                Nn = d - 1;
                count = binom_k<L>(Nn);
                next = 0;
                zl = __builtin_inff();
                have = count > 0;
                if (!have && gl == 0) {
                    J->G[(size_t)i * p + j] = 1;
                    J->G[(size_t)j * p + i] = 1;
                }
                restore = have;
            } else if (restore) {
                // a continued pair: side B done; G = 0 was written if it passed
                if (!pm && gl == 0) {
                    J->G[(size_t)i * p + j] = 1;
                    J->G[(size_t)j * p + i] = 1;
                }
                restore = false;
            }
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
    if (!st || threadIdx.x >= 5) return;
    // the block's totals and the FAST kernels' stripes blockIdx.x,
    // blockIdx.x + gridDim.x, ... (complete: both FAST launches precede this
    // one on the stream) go straight to the level's statistics, so no block
    // waits for another
    const int k = threadIdx.x;
    const bool is_max = k == BS_MAXRATIO;
    unsigned long long v = 0;
    for (int w2 = 0; w2 < 4; ++w2) {
        const unsigned long long x = wstat[w2][k];
        v = is_max ? (x > v ? x : v) : v + x;
    }
    for (int sidx = blockIdx.x; sidx < BSTAT_STRIPES; sidx += gridDim.x) {
        const unsigned long long x = bstats[(size_t)sidx * BSTAT_WORDS + k];
        v = is_max ? (x > v ? x : v) : v + x;
    }
    if (v) {
        const int slot = k == BS_EXEC ? ST_EXEC : k == BS_USEFUL ? ST_USEFUL : k == BS_RECHECK ? ST_RECHECK : ST_FLIP;
        if (is_max)
            atomicMax(st + ST_MAXRATIO, v);
        else
            atomicAdd(st + slot, v);
    }
}


// ----------------------------------------------------------------------------
// levels above the templated range (9..GPUCSL_MAX_LEVEL): one wave per
// ordered pair, the 64 lanes evaluating one CI test together in FP64 in the
// reference's operation order (gaussian_helpers.cu:16-262, the same steps as
// ref64::gauss_z), so every test takes the FP64 decision directly and no
// recheck queue is needed.  Each output element of a matrix product is one
// lane's sequential dot product in the reference's k order; the Cholesky
// factor is formed column by column and the Gauss-Jordan inverse pivot by
// pivot, every element with the reference's formula and operands, so the
// bits equal the sequential routine's.  A wave walks the pair's subsets in
// rank order (the lexicographic successor of the previous one, one lane per
// position) and stops at the first passing test.  Side A, then side B, as in
// the templated levels.
// ----------------------------------------------------------------------------
constexpr int GEN_THREADS = 64;

__host__ __device__ inline size_t gen_lds_bytes(int L) {
    // doubles: M2, T, Lo, R, M2i (L*L each), M1 and W (2L each), H (4); then S (L ints)
    return sizeof(double) * ((size_t)5 * L * L + 4 * (size_t)L + 4) + sizeof(int) * (size_t)L;
}

struct GenMats {
    double *M2, *T, *Lo, *R, *M2i, *M1, *W, *H;
    int *S;
};

__device__ __forceinline__ GenMats gen_carve(void *lds, int L) {
    GenMats m;
    double *d = (double *)lds;
    const int LL = L * L;
    m.M2 = d;
    m.T = d + LL;
    m.Lo = d + 2 * LL;
    m.R = d + 3 * LL;
    m.M2i = d + 4 * LL;
    m.M1 = d + 5 * LL;
    m.W = m.M1 + 2 * L;
    m.H = m.W + 2 * L;
    m.S = (int *)(m.H + 4);
    return m;
}

// R (D1 x D3) = op(A) op(B) over D2, one output element per lane step.
// MODE 0: A (D1 x D2) B (D2 x D3); 1: A^T with A (D2 x D1); 2: B^T with B (D3 x D2)
template <int MODE>
__device__ __forceinline__ void gen_mm(const double *A, const double *B, double *R, int D1, int D2, int D3) {
#pragma clang fp contract(off)
    for (int o = threadIdx.x; o < D1 * D3; o += GEN_THREADS) {
        const int r = o / D3, c = o - (o / D3) * D3;
        double acc = 0.0;
        for (int k = 0; k < D2; ++k) {
            const double a = MODE == 1 ? A[k * D1 + r] : A[r * D2 + k];
            const double b = MODE == 2 ? B[c * D2 + k] : B[k * D3 + c];
            acc += a * b;
        }
        R[o] = acc;
    }
    __syncthreads();
}

// Lo = cholesky(A) as ref64::cholesky: Lo[r][c] = sqrt(A[r][r] - acc) on the
// diagonal, (1 / Lo[c][c]) (A[r][c] - acc) below, acc = sum_{k<c} Lo[r][k] Lo[c][k]
__device__ void gen_cholesky(const double *A, double *Lo, int L) {
#pragma clang fp contract(off)
    for (int o = threadIdx.x; o < L * L; o += GEN_THREADS) Lo[o] = 0.0;
    __syncthreads();
    for (int c = 0; c < L; ++c) {
        if (threadIdx.x == 0) {
            double acc = 0.0;
            for (int k = 0; k < c; ++k) acc += Lo[c * L + k] * Lo[c * L + k];
            Lo[c * L + c] = sqrt(A[c * L + c] - acc);
        }
        __syncthreads();
        for (int r = c + 1 + threadIdx.x; r < L; r += GEN_THREADS) {
            double acc = 0.0;
            for (int k = 0; k < c; ++k) acc += Lo[r * L + k] * Lo[c * L + k];
            Lo[r * L + c] = (1.0 / Lo[c * L + c] * (A[r * L + c] - acc));
        }
        __syncthreads();
    }
}

// Inv = M^-1 by Gauss-Jordan without pivoting, as ref64::gj_inverse (M is
// overwritten).  Each pivot: the pivot row divided by f, then every other row
// r minus the pivot row times its old M[r][piv].
__device__ void gen_gj_inverse(double *M, double *Inv, int L) {
#pragma clang fp contract(off)
    constexpr int PER = (GPUCSL_MAX_LEVEL * GPUCSL_MAX_LEVEL + GEN_THREADS - 1) / GEN_THREADS;
    for (int o = threadIdx.x; o < L * L; o += GEN_THREADS) Inv[o] = (o / L == o % L) ? 1.0 : 0.0;
    __syncthreads();
    for (int piv = 0; piv < L; ++piv) {
        const double f = M[piv * L + piv];
        __syncthreads();
        for (int c = threadIdx.x; c < L; c += GEN_THREADS) {
            Inv[piv * L + c] = Inv[piv * L + c] / f;
            M[piv * L + c] = M[piv * L + c] / f;
        }
        __syncthreads();
        double g[PER];
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            const int o = threadIdx.x + t * GEN_THREADS;
            g[t] = o < L * L ? M[(o / L) * L + piv] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            const int o = threadIdx.x + t * GEN_THREADS;
            if (o < L * L) {
                const int r = o / L, c = o - r * L;
                if (r != piv) {
                    Inv[o] -= Inv[piv * L + c] * g[t];
                    M[o] -= M[piv * L + c] * g[t];
                }
            }
        }
        __syncthreads();
    }
}

// z of (vi, vj | S) from C64, S in m.S (L ids); every lane returns it.
__device__ double gen_gauss_z(const double *__restrict__ C, int p, int vi, int vj, const GenMats &m, int L) {
#pragma clang fp contract(off)
    for (int o = threadIdx.x; o < L * L; o += GEN_THREADS) {
        const int r = o / L, c = o - r * L;
        m.M2[o] = C[(size_t)m.S[r] * p + m.S[c]];
    }
    for (int k = threadIdx.x; k < L; k += GEN_THREADS) {
        m.M1[k] = C[(size_t)vi * p + m.S[k]];
        m.M1[L + k] = C[(size_t)vj * p + m.S[k]];
    }
    __syncthreads();
    // compute_m_2_inv (gaussian_helpers.cu:201-218)
    gen_mm<1>(m.M2, m.M2, m.T, L, L, L);
    gen_cholesky(m.T, m.Lo, L);
    gen_mm<1>(m.Lo, m.Lo, m.T, L, L, L);
    gen_gj_inverse(m.T, m.R, L);
    gen_mm<0>(m.Lo, m.R, m.T, L, L, L);
    gen_mm<0>(m.T, m.R, m.M2i, L, L, L);
    gen_mm<2>(m.M2i, m.Lo, m.T, L, L, L);
    gen_mm<2>(m.T, m.M2, m.M2i, L, L, L);
    // compute_h (gaussian_helpers.cu:78-89)
    gen_mm<2>(m.M2i, m.M1, m.W, L, L, 2);
    gen_mm<0>(m.M1, m.W, m.H, 2, L, 2);
    const double m0[4] = {1.0, C[(size_t)vi * p + vj], C[(size_t)vj * p + vi], 1.0};
    double H[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) H[t] = m0[t] - m.H[t];
    __syncthreads();  // m.H and the matrices are rewritten by the next test
    // compute_fishers_z_transform (gaussian_helpers.cu:220-227)
    const double rho = H[1] / (sqrt(fabs(H[0])) * sqrt(fabs(H[3])));
    return fabs(0.5 * (log(fabs(1 + rho)) - log(fabs(1 - rho))));
}

__global__ __launch_bounds__(GEN_THREADS) void gauss_level_generic_kernel(const Job *__restrict__ jobs, int level,
                                                                        int side, unsigned long long *__restrict__ st) {
    extern __shared__ double gen_lds[];
    const Job &J = jobs[blockIdx.y];
    if (level > J.max_level) return;
    const int p = J.p, L = level, lane = threadIdx.x;
    JobWs jw = carve_job(J.ws, p);
    const int32_t *rowoff = side == 0 ? jw.rowoffA : jw.rowoffB;
    const int total = rowoff[p];
    const double thr = J.thr[L];
    const float zscale = level_zscale(J, L);
    const bool both = !sepset_one_way(J);
    GenMats m = gen_carve(gen_lds, L);
    unsigned long long n_exec = 0;
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        int lo = 0, hi = p;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (rowoff[mid] <= w) lo = mid; else hi = mid;
        }
        const int i = lo;
        const int d = jw.deg[i];
        const int cnt = rowoff[i + 1] - rowoff[i];
        const int jpos = (side == 0 ? d - cnt : 0) + (w - rowoff[i]);
        const int32_t *nb = jw.nbr + (size_t)i * p;
        GPUCSL_DCHECK(i >= 0 && i < p && jpos >= 0 && jpos < d);
        const int j = nb[jpos];
        if (side == 1 && J.G[(size_t)i * p + j] != 1) continue;  // side A removed the edge
        const int N = d - 1;
        const uint64_t count = binom(N, L);
        // lane k < L holds the k-th 1-based position of the current subset
        int pos = lane + 1;
        float zl = __builtin_inff();
        for (uint64_t r = 0; r < count; ++r) {
            if (lane < L) {
                const int idx = pos - 1;
                GPUCSL_DCHECK(idx >= 0 && idx < N);
                m.S[lane] = nb[idx + (idx >= jpos)];
            }
            __syncthreads();
            const double z = gen_gauss_z(J.C64, p, i, j, m, L);
            ++n_exec;
            if (!(z != z)) zl = fminf(zl, (float)z);
            if (z < thr) {
                stg(J.G + (size_t)i * p + j, (uint8_t)0);
                stg(J.G + (size_t)j * p + i, (uint8_t)0);
                if (lane < L) {
                    const int idx = pos - 1;
                    const int v = nb[idx + (idx >= jpos)];
                    J.sepsets[((size_t)i * p + j) * J.max_level + lane] = v;
                    if (both) J.sepsets[((size_t)j * p + i) * J.max_level + lane] = v;
                }
                break;
            }
            // lexicographic successor: the last position below its maximum
            // N - L + 1 + k goes up by one, the positions after it follow
            const uint64_t up = __ballot(lane < L && pos < N - L + 1 + lane);
            if (up) {
                const int K = 63 - __clzll((long long)up);
                const int base = __shfl(pos, K) + 1;
                if (lane >= K && lane < L) pos = base + (lane - K);
            }
        }
        if (lane == 0 && zl < __builtin_inff()) update_zmin(J, p, i, j, zl * zscale);
    }
    if (lane == 0 && n_exec) {
        atomicAdd(&st[ST_EXEC], n_exec);
        atomicAdd(&st[ST_USEFUL], n_exec);
    }
}

// ----------------------------------------------------------------------------
// pmax and output conversion
// ----------------------------------------------------------------------------
// pmax = 2 (1 - Phi(sqrt(n-3) |zmin|)) = erfc(sqrt(n-3) |zmin| / sqrt 2); -1 on
// and below the diagonal.  gpucsl/pc/helpers.py:62-82.
__device__ __forceinline__ float pmax_of(float zmin, double sqrt_nm3) {
    const double x = fabs(sqrt_nm3 * (double)zmin);
    return (float)erfc(x * 0.70710678118654752440);
}

__global__ void pmax_kernel(const float *__restrict__ zmin, int p, double sqrt_nm3, float *__restrict__ pmax) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (b >= p) return;
    const size_t t = (size_t)a * p + b;
    pmax[t] = b <= a ? -1.0f : pmax_of(zmin[t], sqrt_nm3);
}

// pmax and (optionally) the uint16 skeleton of every job of a launch
// Job R.job[k]: one workgroup per row and 256 columns, p * ceil(p / 256).
__global__ void job_finish_kernel(const Job *__restrict__ jobs, BlockRanges R) {
    int local;
    const Job &J = jobs[range_job(R, blockIdx.x, local)];
    const int p = J.p;
    const int cbk = (p + 255) / 256;
    const int b = (local % cbk) * 256 + threadIdx.x;
    const int a = local / cbk;
    if (b >= p || a >= p) return;
    const size_t t = (size_t)a * p + b;
    // pcalg mode: zmin already carries sqrt(n - |S| - 3)
    const double s = (J.flags & GPUCSL_JOB_PCALG_COMPAT) ? 1.0 : sqrt((double)(J.n - 3));
    if (J.pmax) J.pmax[t] = b <= a ? -1.0f : pmax_of(J.zmin[t], s);
    if (J.skeleton) J.skeleton[t] = J.G[t];
}

// pmax_kernel and the uint16 skeleton of one dataset in one pass
__global__ void outputs_kernel(const float *__restrict__ zmin, const uint8_t *__restrict__ g, int p, double sqrt_nm3,
                               float *__restrict__ pmax, uint16_t *__restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (b >= p) return;
    const size_t t = (size_t)a * p + b;
    pmax[t] = b <= a ? -1.0f : pmax_of(zmin[t], sqrt_nm3);
    out[t] = g[t];
}
__global__ void to_u16_kernel(const uint8_t *__restrict__ g, size_t count, uint16_t *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) out[i] = g[i];
}

// ----------------------------------------------------------------------------
// host side: launches over a job table
// ----------------------------------------------------------------------------
static int slow_grid() { return device_cus() / 2; }
static int fast_grid(int bucket, int level) {
    return device_cus() * tuning_of(bucket, level).fast_blocks_per_cu;
}

size_t gauss_batch_ws_bytes(const Job *jobs_host, int nj) {
    return batch_ws_bytes(nj, queue_capacity(jobs_host, nj), item_capacity(jobs_host, nj));
}

static int max_p(const Job *jh, int nj) {
    int m = 1;
    for (int b = 0; b < nj; ++b) m = jh[b].p > m ? jh[b].p : m;
    return m;
}

int gauss_batch_init(const Job *jh, const Job *jd, int nj, hipStream_t s) {
    for (int b = 0; b < nj; ++b) {
        const Job &J = jh[b];
        if (J.p < 1 || J.max_level < 0 || !J.G || !J.zmin || (J.max_level > 0 && !J.sepsets) || !J.ws)
            return set_error(ERR_ARG, "gauss_init: job %d has bad arguments", b);
    }
    return for_block_ranges(
        nj, [&](int b) { return (long long)init_blocks(jh[b].p, jh[b].max_level); },
        [&](const BlockRanges &R) -> int {
            GPUCSL_LAUNCH("gauss_init", job_init_kernel, dim3(R.first[R.count]), dim3(256), 0, s, jd, R);
            return check_hip(hipGetLastError(), "gauss_init launch");
        });
}

int gauss_batch_level0(const Job *jh, const Job *jd, int nj, float margin_k, unsigned long long *st,
                       hipStream_t s) {
    for (int b = 0; b < nj; ++b)
        if (!jh[b].C32 || !jh[b].C64 || !jh[b].G || !jh[b].zmin)
            return set_error(ERR_ARG, "gauss_level0: job %d has a null matrix", b);
    return for_block_ranges(
        nj, [&](int b) { return (long long)jh[b].p * ((jh[b].p + 255) / 256); },
        [&](const BlockRanges &R) -> int {
            GPUCSL_LAUNCH("gauss_level0", gauss_level0_kernel, dim3(R.first[R.count]), dim3(256), 0, s, jd, margin_k,
                          st, R);
            return check_hip(hipGetLastError(), "gauss_level0 launch");
        });
}

static int compact_and_plan(const Job *jh, const Job *jd, int nj, int level, int gs_min, unsigned long long *st,
                            const PlanArgs &plan, hipStream_t s) {
    const int mp = max_p(jh, nj);
    const Team team = team_for(tuning_bucket(nj, mp),
                               level < 1 ? 1 : (level > GPUCSL_MAX_LEVEL_TEMPLATE ? GPUCSL_MAX_LEVEL_TEMPLATE : level),
                               gs_min);
    const int rc = for_block_ranges(
        nj, [&](int b) { return (long long)(jh[b].p + 3) / 4; },
        [&](const BlockRanges &R) -> int {
            GPUCSL_LAUNCH("compact", compact_kernel, dim3(R.first[R.count]), dim3(256), 0, s, jd, level, team, st, plan,
                          R);
            return check_hip(hipGetLastError(), "compact launch");
        });
    if (rc) return rc;
    // the scan only when a plan or the level statistics are wanted (the
    // neighbour lists and degrees alone come from the compaction)
    if (plan.done || st)
        GPUCSL_LAUNCH("compact_scan", scan_kernel, dim3(nj), dim3(1024), 0, s, jd, level, team, st, plan);
    if (plan.done && nj > 1)
        GPUCSL_LAUNCH("compact_items", items_kernel, dim3(nj), dim3(256), 0, s, jd, level, team, plan);
    return check_hip(hipGetLastError(), "compact launch");
}

int gauss_batch_compact(const Job *jh, const Job *jd, int nj, int level, int gs_min, unsigned long long *st,
                        hipStream_t s) {
    return compact_and_plan(jh, jd, nj, level, gs_min, st, PlanArgs{}, s);
}

static const char *const kPhaseFastName[] = {
    "", "gauss_level1_fast", "gauss_level2_fast", "gauss_level3_fast", "gauss_level4_fast",
    "gauss_level5_fast", "gauss_level6_fast", "gauss_level7_fast", "gauss_level8_fast"};
static const char *const kPhaseSlowName[] = {
    "", "gauss_level1_slow", "gauss_level2_slow", "gauss_level3_slow", "gauss_level4_slow",
    "gauss_level5_slow", "gauss_level6_slow", "gauss_level7_slow", "gauss_level8_slow"};

// Side 0: the FAST kernel.  Side 1: the FAST kernel, then the SLOW kernel for
// the pairs both sides queued (and the level's statistics).
template <int L>
static void launch_phase(const Job *jd, int nj, int bucket, const BatchWs &w, int side, int gs_min, float margin_k,
                         unsigned long long *st, hipStream_t s) {
    const int32_t *base = side == 0 ? w.itembaseA : w.itembaseB;
    const WorkItem *items = side == 0 ? w.itemsA : w.itemsB;
    SlowItem *queue = side == 0 ? w.queueA : w.queueB;
    unsigned int *qcount = w.counters + CNT_QUEUE + side;
    const gpucsl_level_tuning tl = tuning_of(bucket, L);
    const int lds = tl.fast_lds_bytes;
    if (tl.wave_items)
        GPUCSL_LAUNCH(kPhaseFastName[L], gauss_level_wave_kernel<L>, dim3(fast_grid(bucket, L)), dim3(256), lds, s,
                      jd, nj, base, items, side, team_for(bucket, L, gs_min), lds / 16, margin_k, queue, qcount,
                      w.bstats);
    else
        GPUCSL_LAUNCH(kPhaseFastName[L], gauss_level_fast_kernel<L>, dim3(fast_grid(bucket, L)), dim3(256), lds, s,
                      jd, nj, base, items, side, team_for(bucket, L, gs_min), lds / 4, margin_k, queue, qcount,
                      w.bstats);
    if (side == 0) return;
    GPUCSL_LAUNCH(kPhaseSlowName[L], (gauss_slow_kernel<L, 64>), dim3(slow_grid()), dim3(256), 0, s, jd, margin_k,
                  w.counters + STRIPES, w.queueA, w.counters + CNT_QUEUE, w.queueB, w.counters + CNT_QUEUE + 1,
                  w.bstats, st);
}

int gauss_batch_level(int level, const Job *jh, const Job *jd, int nj, float margin_k, int group_size,
                      unsigned long long *st, void *bws, size_t bws_bytes, hipStream_t s) {
    if (level < 1) return set_error(ERR_ARG, "gauss_level: level %d < 1", level);
    if (level > GPUCSL_MAX_LEVEL)
        return set_error(ERR_UNSUPPORTED, "gauss_level: levels above %d are not built", GPUCSL_MAX_LEVEL);
    if (group_size != 8 && group_size != 16 && group_size != 32 && group_size != 64)
        return set_error(ERR_ARG, "group size must be 8, 16, 32 or 64 (got %d)", group_size);
    if (!st || !bws) return set_error(ERR_ARG, "gauss_level: null stats or workspace");
    const size_t qcap = queue_capacity(jh, nj), icap = item_capacity(jh, nj);
    if (bws_bytes < batch_ws_bytes(nj, qcap, icap)) return set_error(ERR_WORKSPACE, "gauss_level: workspace too small");
    for (int b = 0; b < nj; ++b)
        if (!jh[b].C32 || !jh[b].C64 || !jh[b].G || !jh[b].zmin || !jh[b].ws || (jh[b].max_level > 0 && !jh[b].sepsets))
            return set_error(ERR_ARG, "gauss_level: job %d has a null buffer", b);
    BatchWs w = carve_batch(bws, nj, qcap, icap);
    const int bucket = tuning_bucket(nj, max_p(jh, nj));
    const PlanArgs plan{w.counters + CNT_PLAN_DONE, nj, w.itembaseA, w.itembaseB, w.counters, w.bstats, w.itemsA,
                        w.itemsB};
    int rc = compact_and_plan(jh, jd, nj, level, group_size, st, plan, s);
    if (rc) return rc;
    if (level > GPUCSL_MAX_LEVEL_TEMPLATE) {
        // a few waves per compute unit: each holds its test's matrices in LDS
        const dim3 grid((unsigned)(device_cus() * 8), (unsigned)nj);
        for (int side = 0; side < 2; ++side)
            GPUCSL_LAUNCH(side == 0 ? "gauss_level_generic_a" : "gauss_level_generic_b", gauss_level_generic_kernel,
                          grid, dim3(GEN_THREADS), gen_lds_bytes(level), s, jd, level, side, st);
        return check_hip(hipGetLastError(), "gauss_level launch");
    }
    for (int side = 0; side < 2; ++side) {
        switch (level) {
#define GPUCSL_CASE(LV)                                                     \
    case LV:                                                                \
        launch_phase<LV>(jd, nj, bucket, w, side, group_size, margin_k, st, s); \
        break;
            GPUCSL_CASE(1) GPUCSL_CASE(2) GPUCSL_CASE(3) GPUCSL_CASE(4)
            GPUCSL_CASE(5) GPUCSL_CASE(6) GPUCSL_CASE(7) GPUCSL_CASE(8)
#undef GPUCSL_CASE
        }
    }
    return check_hip(hipGetLastError(), "gauss_level launch");
}

int gauss_batch_finish(const Job *jh, const Job *jd, int nj, hipStream_t s) {
    for (int b = 0; b < nj; ++b)
        if (jh[b].n < 4) return set_error(ERR_ARG, "finish: job %d has fewer than 4 observations", b);
    return for_block_ranges(
        nj, [&](int b) { return (long long)jh[b].p * ((jh[b].p + 255) / 256); },
        [&](const BlockRanges &R) -> int {
            GPUCSL_LAUNCH("finish", job_finish_kernel, dim3(R.first[R.count]), dim3(256), 0, s, jd, R);
            return check_hip(hipGetLastError(), "finish launch");
        });
}

// ---- single dataset: [job descriptor][launch workspace][job workspace] -------
size_t gauss_workspace_bytes(int p) {
    Job j{};
    j.p = p;
    return JOB_DESC_BYTES + align_up(batch_ws_bytes(1, queue_capacity(&j, 1), item_capacity(&j, 1)), 256) +
           job_ws_bytes(p);
}

struct SingleWs {
    Job *desc;
    void *bws;
    size_t bws_bytes;
    void *jws;
};

static SingleWs carve_single(void *ws, int p) {
    Job j{};
    j.p = p;
    SingleWs w;
    w.desc = (Job *)ws;
    w.bws = (char *)ws + JOB_DESC_BYTES;
    w.bws_bytes = align_up(batch_ws_bytes(1, queue_capacity(&j, 1), item_capacity(&j, 1)), 256);
    w.jws = (char *)w.bws + w.bws_bytes;
    return w;
}

// Builds the one-job table of a single-dataset call in the workspace head.
Job single_job(int p, int max_level, const float *C32, const double *C64, uint8_t *Gm, int32_t *sep, float *zmin,
               const double *thr_host, int thr_count, void *ws) {
    Job j{};
    j.p = p;
    j.max_level = max_level;
    j.n = 4;
    j.C32 = const_cast<float *>(C32);
    j.C64 = const_cast<double *>(C64);
    j.G = Gm;
    j.sepsets = sep;
    j.zmin = zmin;
    j.ws = carve_single(ws, p).jws;
    for (int l = 0; l < thr_count && l < GPUCSL_MAX_LEVELS; ++l) j.thr[l] = thr_host[l];
    return j;
}

int single_write_desc(const Job &j, void *ws, hipStream_t s) {
    hipLaunchKernelGGL(write_job_kernel, dim3(1), dim3(1), 0, s, carve_single(ws, j.p).desc, j);
    return check_hip(hipGetLastError(), "job descriptor");
}

// G = 1, zmin = +inf, sepsets = -1 and the statistics zeroed, in one launch;
// with `dst`, also the job descriptor of a single-dataset workspace.
__global__ void single_init_kernel(uint8_t *__restrict__ Gm, int32_t *__restrict__ sep, float *__restrict__ zmin,
                                   size_t pp, size_t ns, unsigned long long *__restrict__ st, int nst, Job *dst,
                                   Job src) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (dst && t0 == 0) *dst = src;
    for (size_t t = t0; t < pp; t += stride) {
        Gm[t] = 1;
        zmin[t] = __builtin_inff();
    }
    if (sep)
        for (size_t t = t0; t < ns; t += stride) sep[t] = -1;
    if (st && t0 < (size_t)nst) st[t0] = 0;
}

static int init_launch(int p, int max_level, uint8_t *Gm, int32_t *sep, float *zmin, unsigned long long *st, int nst,
                       Job *dst, const Job &src, hipStream_t s) {
    if (p < 1 || max_level < 0 || !Gm || !zmin) return set_error(ERR_ARG, "gauss_init: bad arguments");
    const size_t pp = (size_t)p * p, ns = max_level > 0 && sep ? pp * max_level : 0;
    size_t blocks = ((ns > pp ? ns : pp) + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks * 256 < (size_t)nst) blocks = ((size_t)nst + 255) / 256;
    GPUCSL_LAUNCH("gauss_init", single_init_kernel, dim3((unsigned)blocks), dim3(256), 0, s, Gm,
                  max_level > 0 ? sep : nullptr, zmin, pp, ns, st, nst, dst, src);
    return check_hip(hipGetLastError(), "gauss_init launch");
}

int gauss_init_launch(int p, int max_level, uint8_t *Gm, int32_t *sep, float *zmin, unsigned long long *st, int nst,
                      hipStream_t s) {
    return init_launch(p, max_level, Gm, sep, zmin, st, nst, nullptr, Job{}, s);
}

int gauss_level0_launch(const float *C32, const double *C64, int p, double thr, float margin_k, uint8_t *Gm,
                        float *zmin, unsigned long long *stats, void *ws, size_t ws_bytes, hipStream_t s) {
    if (!C32 || !C64 || !Gm || !zmin || !stats || !ws || p < 1) return set_error(ERR_ARG, "gauss_level0: bad arguments");
    if (ws_bytes < gauss_workspace_bytes(p)) return set_error(ERR_WORKSPACE, "gauss_level0: workspace too small");
    Job j = single_job(p, 0, C32, C64, Gm, nullptr, zmin, &thr, 1, ws);
    int rc = single_write_desc(j, ws, s);
    if (rc) return rc;
    return gauss_batch_level0(&j, carve_single(ws, p).desc, 1, margin_k, stats, s);
}

int gauss_level_launch(int level, const float *C32, const double *C64, int p, int max_level, double thr,
                       float margin_k, int group_size, uint8_t *Gm, int32_t *sep, float *zmin,
                       unsigned long long *stats, void *ws, size_t ws_bytes, hipStream_t s) {
    if (level < 1 || level > max_level) return set_error(ERR_ARG, "gauss_level: level %d outside [1, %d]", level, max_level);
    if (!ws) return set_error(ERR_ARG, "gauss_level: null workspace");
    if (ws_bytes < gauss_workspace_bytes(p)) return set_error(ERR_WORKSPACE, "gauss_level: workspace too small");
    if (level > GPUCSL_MAX_LEVEL)
        return set_error(ERR_UNSUPPORTED, "gauss_level: levels above %d are not built", GPUCSL_MAX_LEVEL);
    double thr_all[GPUCSL_MAX_LEVELS] = {0};
    thr_all[level] = thr;
    Job j = single_job(p, max_level, C32, C64, Gm, sep, zmin, thr_all, GPUCSL_MAX_LEVELS, ws);
    int rc = single_write_desc(j, ws, s);
    if (rc) return rc;
    SingleWs sw = carve_single(ws, p);
    return gauss_batch_level(level, &j, sw.desc, 1, margin_k, group_size, stats, sw.bws, sw.bws_bytes, s);
}

// init (with the job descriptor), level 0 and levels 1..max_level of one
// dataset; stats holds nst words
int gauss_skeleton_launch(const float *C32, const double *C64, int p, int max_level, const double *thr,
                          float margin_k, int group_size, uint8_t *Gm, int32_t *sep, float *zmin,
                          unsigned long long *stats, int nst, void *ws, size_t ws_bytes, hipStream_t s) {
    if (!ws || ws_bytes < gauss_workspace_bytes(p)) return set_error(ERR_WORKSPACE, "gauss_skeleton: workspace too small");
    Job j = single_job(p, max_level, C32, C64, Gm, sep, zmin, thr, max_level + 1, ws);
    SingleWs sw = carve_single(ws, p);
    int rc = init_launch(p, max_level, Gm, sep, zmin, stats, nst, sw.desc, j, s);
    if (rc) return rc;
    rc = gauss_batch_level0(&j, sw.desc, 1, margin_k, stats, s);
    if (rc) return rc;
    const int built = max_level < GPUCSL_MAX_LEVEL ? max_level : GPUCSL_MAX_LEVEL;
    for (int l = 1; l <= built; ++l) {
        rc = gauss_batch_level(l, &j, sw.desc, 1, margin_k, group_size, stats + (size_t)l * ST_SLOTS, sw.bws,
                               sw.bws_bytes, s);
        if (rc) return rc;
    }
    return OK;
}

// ---- single dataset without a caller workspace ----------------------------
// For callers that cannot allocate device memory (code under graph capture,
// serving loops, harnesses that forbid hipMalloc): the workspace of
// gauss_skeleton_launch lives in this file-scope array, one copy per device,
// held until the process ends.  Host code never passes the array's name to a
// kernel: its device address, per device, comes from hipGetSymbolAddress on
// the first call with that device current.  Every call on a device shares its
// array, so calls that overlap on one device race; the kernels set every word
// they read (descriptor, counters, statistics, lists), so nothing needs
// clearing between calls.
static __device__ __align__(256) unsigned char scratch[GPUCSL_NOALLOC_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[SCRATCH_MAX_DEVICES];

static int scratch_of_current_device(void **out) {
    int dev = 0;
    int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
        return set_error(ERR_UNSUPPORTED, "gauss_skeleton_noalloc: device %d beyond %d", dev, SCRATCH_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        rc = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (rc) return rc;
        if (!ptr) return set_error(ERR_HIP, "hipGetSymbolAddress(scratch) gave a null address");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return OK;
}

int gauss_skeleton_noalloc_launch(const float *C32, const double *C64, int p, int max_level, const double *thr,
                                  float margin_k, int group_size, uint8_t *Gm, int32_t *sep, float *zmin,
                                  unsigned long long *stats, int nst, hipStream_t s) {
    if (p < 1 || max_level < 0) return set_error(ERR_ARG, "gauss_skeleton_noalloc: bad p or max_level");
    if (max_level > GPUCSL_MAX_LEVEL)
        return set_error(ERR_UNSUPPORTED, "gauss_skeleton_noalloc: max_level %d above %d", max_level,
                         GPUCSL_MAX_LEVEL);
    const size_t need = gauss_workspace_bytes(p);
    if (need > (size_t)GPUCSL_NOALLOC_SCRATCH_BYTES)
        return set_error(ERR_WORKSPACE,
                         "gauss_skeleton_noalloc: p = %d needs %zu bytes of scratch, more than its %zu-byte array; "
                         "use gpucsl_gauss_skeleton with a workspace",
                         p, need, (size_t)GPUCSL_NOALLOC_SCRATCH_BYTES);
    void *ws = nullptr;
    int rc = scratch_of_current_device(&ws);
    if (rc) return rc;
    prof::Suppress no_events;  // event timing would create events
    return gauss_skeleton_launch(C32, C64, p, max_level, thr, margin_k, group_size, Gm, sep, zmin, stats, nst, ws,
                                 need, s);
}

int compact_launch(const uint8_t *Gm, int p, int level, void *ws, unsigned long long *stats, hipStream_t s) {
    Job j = single_job(p, level, nullptr, nullptr, const_cast<uint8_t *>(Gm), nullptr, nullptr, nullptr, 0, ws);
    if (!stats) {
        GPUCSL_LAUNCH("compact", compact_one_kernel, dim3((p + 3) / 4, 1), dim3(256), 0, s, j, level,
                      team_for(tuning_bucket(1, p),
                               level < 1 ? 1 : (level > GPUCSL_MAX_LEVEL_TEMPLATE ? GPUCSL_MAX_LEVEL_TEMPLATE : level),
                               16));
        return check_hip(hipGetLastError(), "compact launch");
    }
    int rc = single_write_desc(j, ws, s);
    if (rc) return rc;
    return gauss_batch_compact(&j, carve_single(ws, p).desc, 1, level, 16, stats, s);
}

// Views of the level-start neighbour lists inside a single-dataset workspace.
void level_list_views(void *ws, int p, int32_t **nbr, int32_t **deg) {
    JobWs w = carve_job(carve_single(ws, p).jws, p);
    *nbr = w.nbr;
    *deg = w.deg;
}

int pmax_launch(const float *zmin, int p, int64_t n, float *pmax, hipStream_t s) {
    if (!zmin || !pmax || p < 1 || n < 4) return set_error(ERR_ARG, "pmax: bad arguments");
    GPUCSL_LAUNCH("pmax", pmax_kernel, dim3((p + 255) / 256, p), dim3(256), 0, s, zmin, p, sqrt((double)(n - 3)),
                  pmax);
    return check_hip(hipGetLastError(), "pmax launch");
}

int outputs_launch(const float *zmin, const uint8_t *Gm, int p, int64_t n, float *pmax, uint16_t *out, hipStream_t s) {
    if (!zmin || !Gm || !pmax || !out || p < 1 || n < 4) return set_error(ERR_ARG, "gauss_outputs: bad arguments");
    GPUCSL_LAUNCH("outputs", outputs_kernel, dim3((p + 255) / 256, p), dim3(256), 0, s, zmin, Gm, p,
                  sqrt((double)(n - 3)), pmax, out);
    return check_hip(hipGetLastError(), "gauss_outputs launch");
}

int to_u16_launch(const uint8_t *Gm, int p, uint16_t *out, hipStream_t s) {
    const size_t cnt = (size_t)p * p;
    size_t blocks = (cnt + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    GPUCSL_LAUNCH("to_u16", to_u16_kernel, dim3((unsigned)blocks), dim3(256), 0, s, Gm, cnt, out);
    return check_hip(hipGetLastError(), "to_u16 launch");
}

}  // namespace gpucsl
