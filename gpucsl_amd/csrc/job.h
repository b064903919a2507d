// Job tables and workspace layouts shared by the Gaussian kernels.
#pragma once
#include "../../include/gpucsl_hip.h"
#include "common.h"

namespace gpucsl {

using Job = gpucsl_gauss_job;

constexpr int CORR_ROW_SPLITS_MAX = 64;
constexpr int JOB_DESC_BYTES = 512;  // room for one job descriptor at the head of a workspace
static_assert(sizeof(Job) <= JOB_DESC_BYTES, "job descriptor too large");

// Launches over a table of jobs that give each job only the workgroups it
// needs: job base + rel[k] owns workgroups first[k] .. first[k + 1] - 1 of a
// 1-D grid.  Passed by value (a captured launch keeps it; about 3 KB of
// kernel arguments); a batch with more jobs than RANGE_JOBS takes several
// launches.
constexpr int RANGE_JOBS = 512;
struct BlockRanges {
    int count, base;
    uint16_t rel[RANGE_JOBS];
    int first[RANGE_JOBS + 1];
};

// The job of workgroup id and the workgroup's index within that job.
__device__ __forceinline__ int range_job(const BlockRanges &R, int id, int &local) {
    int lo = 0, hi = R.count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (R.first[mid] <= id)
            lo = mid;
        else
            hi = mid - 1;
    }
    local = id - R.first[lo];
    return R.base + R.rel[lo];
}

// Calls launch(ranges) for the jobs in chunks of at most RANGE_JOBS, job b
// taking blocks_of(b) workgroups (jobs with none are left out); returns the
// first non-zero status.
template <class BlocksOf, class Launch>
inline int for_block_ranges(int nj, BlocksOf blocks_of, Launch launch) {
    BlockRanges R{};
    for (int b = 0; b < nj; ++b) {
        const long long nb = blocks_of(b);
        if (nb > 0) {
            if (R.count == 0) R.base = b;
            R.rel[R.count] = (uint16_t)(b - R.base);
            R.first[R.count + 1] = R.first[R.count] + (int)nb;
            ++R.count;
        }
        const bool full = R.count == RANGE_JOBS || (R.count > 0 && b + 1 - R.base > 65535);
        if (R.count > 0 && (full || b == nj - 1)) {
            const int rc = launch(R);
            if (rc) return rc;
            R = BlockRanges{};
        }
    }
    return 0;
}

// FP64-recheck queue entry
struct SlowItem {
    uint64_t start;     // first subset rank still to evaluate
    int jpos;           // position of the partner j in the row's neighbour list
    int job;            // index into the launch's job table
    float zl;           // zmin over the tests already counted
    int row;            // the row i whose side the pair was queued from
};

// level work item: `pairs` consecutive pairs of one row, from pair q0
struct WorkItem {
    int job, row, q0, pad;
};

// per-job level workspace
struct JobWs {
    int32_t *nbr, *deg, *cntA, *cntB, *rowoffA, *rowoffB, *itemoffA, *itemoffB;
    float *rthr;  // per level: (float) tanh(thr[level]), written by the level's scan
};

__host__ __device__ inline JobWs carve_job(void *ws, int p) {
    JobWs w;
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
    w.rthr = (float *)(b + off);
    return w;
}

inline size_t job_ws_bytes(int p) {
    size_t off = 0;
    off = align_up(off + sizeof(int32_t) * (size_t)p * p, 256);
    off = align_up(off + sizeof(int32_t) * (size_t)p * 3, 256);
    off = align_up(off + sizeof(int32_t) * (size_t)(p + 1) * 4, 256);
    off = align_up(off + sizeof(float) * GPUCSL_MAX_LEVELS, 256);
    return off;
}

// Digit planes of the integer Gram path (float32 data): the centred, scaled
// value of column c at row k is sum_a D_a[c][k] 128^(5-a), every digit in
// [-64, 64].  Rows and columns are padded to multiples of 64 (padding is zero)
// and each digit plane is stored as 64 x 64 tiles of 4 KB, tile (column block
// cb, row step st) at ((a * CB + cb) * steps + st) * 4096 with CB = columns / 64
// and steps = rows / 64, and inside a tile column c at byte 64 c, row r at
// byte r: a column panel of one step is one contiguous 4 KB read.
constexpr int I8_PLANES = 6;
constexpr int64_t I8_MAX_STEPS = 1024;  // 64-row steps per split-K block (int32 range)
constexpr int I8_TILE = 4096;
__host__ __device__ inline int64_t i8_rows(int64_t n) { return (n + 63) / 64 * 64; }
__host__ __device__ inline int i8_cols(int p) { return (p + 63) / 64 * 64; }
__host__ __device__ inline size_t i8_tile(int a, int cb, int64_t st, int CB, int64_t steps) {
    return (((size_t)a * CB + cb) * (size_t)steps + (size_t)st) * I8_TILE;
}

// per-job correlation workspace: per row split s and column a the sum, minimum
// and maximum of X[., a] over the split; the digit planes; the ks split-K
// partial Gram matrices (p x p each, upper tiles only)
struct CorrWs {
    double *partial, *cmin, *cmax;
    int8_t *planes;
    double *gram;
};

__host__ __device__ inline CorrWs carve_corr(void *ws, int64_t n, int p) {
    CorrWs c;
    char *b = (char *)ws;
    size_t off = 0;
    const size_t split_bytes = align_up(sizeof(double) * (size_t)CORR_ROW_SPLITS_MAX * p, 256);
    c.partial = (double *)(b + off);
    off += split_bytes;
    c.cmin = (double *)(b + off);
    off += split_bytes;
    c.cmax = (double *)(b + off);
    off += split_bytes;
    c.planes = (int8_t *)(b + off);
    off = align_up(off + (size_t)I8_PLANES * i8_cols(p) * (size_t)i8_rows(n), 256);
    c.gram = (double *)(b + off);
    return c;
}

__host__ __device__ inline CorrWs carve_corr(const Job &J) { return carve_corr(J.corr_ws, J.n, J.p); }

inline size_t corr_job_ws_bytes(int64_t n, int p, int k_splits) {
    size_t off = 3 * align_up(sizeof(double) * (size_t)CORR_ROW_SPLITS_MAX * p, 256);
    off = align_up(off + (size_t)I8_PLANES * i8_cols(p) * (size_t)i8_rows(n), 256);
    off = align_up(off + sizeof(double) * (size_t)k_splits * p * p, 256);
    return off;
}

// per-launch (batch) workspace
constexpr int NUM_COUNTERS = 64;
constexpr int CNT_QUEUE = 32;       // 32, 33: queue lengths of sides A and B
constexpr int CNT_PLAN_DONE = 40;   // scan blocks finished (level planning)
constexpr int STRIPES = 8;
constexpr int BSTAT_WORDS = 8;
constexpr int BSTAT_STRIPES = 64;

struct BatchWs {
    unsigned int *counters;        // NUM_COUNTERS
    unsigned long long *bstats;    // BSTAT_STRIPES * BSTAT_WORDS
    int32_t *itembaseA, *itembaseB;  // nj + 1 each
    SlowItem *queueA, *queueB;     // qcap each
    WorkItem *itemsA, *itemsB;     // qcap + sum of p each
};

// A phase has at most one queue entry per ordered pair of its side, and at
// most one work item per pair plus one per row.
inline size_t batch_ws_bytes(int nj, size_t qcap, size_t icap) {
    size_t off = 0;
    off = align_up(off + sizeof(unsigned int) * NUM_COUNTERS, 256);
    off = align_up(off + sizeof(unsigned long long) * BSTAT_STRIPES * BSTAT_WORDS, 256);
    off = align_up(off + sizeof(int32_t) * (size_t)(nj + 1) * 2, 256);
    off = align_up(off + sizeof(SlowItem) * qcap * 2, 256);
    off = align_up(off + sizeof(WorkItem) * icap * 2, 256);
    return off;
}

inline BatchWs carve_batch(void *ws, int nj, size_t qcap, size_t icap) {
    BatchWs w;
    char *b = (char *)ws;
    size_t off = 0;
    w.counters = (unsigned int *)(b + off);
    off = align_up(off + sizeof(unsigned int) * NUM_COUNTERS, 256);
    w.bstats = (unsigned long long *)(b + off);
    off = align_up(off + sizeof(unsigned long long) * BSTAT_STRIPES * BSTAT_WORDS, 256);
    w.itembaseA = (int32_t *)(b + off);
    w.itembaseB = w.itembaseA + (nj + 1);
    off = align_up(off + sizeof(int32_t) * (size_t)(nj + 1) * 2, 256);
    w.queueA = (SlowItem *)(b + off);
    w.queueB = w.queueA + qcap;
    off = align_up(off + sizeof(SlowItem) * qcap * 2, 256);
    w.itemsA = (WorkItem *)(b + off);
    w.itemsB = w.itemsA + icap;
    return w;
}

inline size_t queue_capacity(const Job *jobs_host, int nj) {
    size_t q = 1;
    for (int b = 0; b < nj; ++b) q += (size_t)jobs_host[b].p * (jobs_host[b].p - 1) / 2;
    return q;
}

inline size_t item_capacity(const Job *jobs_host, int nj) {
    size_t q = queue_capacity(jobs_host, nj);
    for (int b = 0; b < nj; ++b) q += (size_t)jobs_host[b].p;
    return q;
}

// copies one job descriptor into device memory from a kernel argument, so a
// single-dataset call stays stream-ordered and graph-capturable
__global__ void write_job_kernel(Job *dst, Job src);

}  // namespace gpucsl
