// extern "C" entry points of libgpucsl_hip.so (declared in include/gpucsl_hip.h)
// and the library's error reporting.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpucsl_hip.h"
#include "job.h"

namespace gpucsl {

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int check_hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return OK;
    return set_error(ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// ----------------------------------------------------------------------------
// per-launch event timing
// ----------------------------------------------------------------------------
namespace prof {
struct Rec {
    const char *name;
    hipEvent_t a, b;
};
static std::mutex g_mu;
static bool g_on = false;
static std::vector<Rec> g_recs;
static std::vector<hipEvent_t> g_pool;
static thread_local hipEvent_t t_start = nullptr;

static hipEvent_t take() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

static thread_local int t_suppress = 0;
Suppress::Suppress() { ++t_suppress; }
Suppress::~Suppress() { --t_suppress; }

bool enabled() { return g_on && t_suppress == 0; }

void record_start(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    t_start = take();
    (void)hipEventRecord(t_start, s);
}

void record_stop(const char *name, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    hipEvent_t e = take();
    (void)hipEventRecord(e, s);
    g_recs.push_back(Rec{name, t_start, e});
    t_start = nullptr;
}
}  // namespace prof

// defined in corr.hip / gauss.hip
size_t corr_workspace_bytes(int64_t n, int p);
int corr_k_splits(int64_t n, int p, int dtype);
int corr_launch(const void *X, int dtype, int64_t n, int p, double *C64, float *C32, void *ws,
                size_t ws_bytes, hipStream_t s);
int c64_to_c32_launch(const double *C64, int p, float *C32, hipStream_t s);
int corr_batch_launch(const Job *jobs_host, const Job *jobs_dev, int nj, hipStream_t s);
int c64_to_c32_batch_launch(const Job *jobs_host, const Job *jobs_dev, int nj, hipStream_t s);
size_t gauss_workspace_bytes(int p);
int gauss_init_launch(int p, int max_level, uint8_t *Gm, int32_t *sep, float *zmin, unsigned long long *st, int nst,
                      hipStream_t s);
int gauss_level0_launch(const float *C32, const double *C64, int p, double thr, float margin_k, uint8_t *Gm,
                        float *zmin, unsigned long long *stats, void *ws, size_t ws_bytes, hipStream_t s);
int gauss_level_launch(int level, const float *C32, const double *C64, int p, int max_level, double thr,
                       float margin_k, int group_size, uint8_t *Gm, int32_t *sep, float *zmin,
                       unsigned long long *stats, void *ws, size_t ws_bytes, hipStream_t s);
int gauss_skeleton_launch(const float *C32, const double *C64, int p, int max_level, const double *thr,
                          float margin_k, int group_size, uint8_t *Gm, int32_t *sep, float *zmin,
                          unsigned long long *stats, int nst, void *ws, size_t ws_bytes, hipStream_t s);
int gauss_skeleton_noalloc_launch(const float *C32, const double *C64, int p, int max_level, const double *thr,
                                  float margin_k, int group_size, uint8_t *Gm, int32_t *sep, float *zmin,
                                  unsigned long long *stats, int nst, hipStream_t s);
int compact_launch(const uint8_t *Gm, int p, int level, void *ws, unsigned long long *stats, hipStream_t s);
void level_list_views(void *ws, int p, int32_t **nbr, int32_t **deg);
int pmax_launch(const float *zmin, int p, int64_t n, float *pmax, hipStream_t s);
int outputs_launch(const float *zmin, const uint8_t *Gm, int p, int64_t n, float *pmax, uint16_t *out, hipStream_t s);
size_t gauss_batch_ws_bytes(const Job *jobs_host, int nj);
int set_level_tuning(int bucket, int level, const gpucsl_level_tuning *t);
int get_level_tuning(int bucket, int level, gpucsl_level_tuning *t);
int tuning_bucket(int nj, int max_p);
int gauss_batch_init(const Job *jh, const Job *jd, int nj, hipStream_t s);
int gauss_batch_level0(const Job *jh, const Job *jd, int nj, float margin_k, unsigned long long *st, hipStream_t s);
int gauss_batch_compact(const Job *jh, const Job *jd, int nj, int level, int gs_min, unsigned long long *st,
                        hipStream_t s);
int gauss_batch_level(int level, const Job *jh, const Job *jd, int nj, float margin_k, int group_size,
                      unsigned long long *st, void *bws, size_t bws_bytes, hipStream_t s);
int gauss_batch_finish(const Job *jh, const Job *jd, int nj, hipStream_t s);
size_t disc_workspace_bytes(int p, int max_level, int max_dim);
size_t disc_min_workspace_bytes(int p, int max_level, int max_dim);
int disc_encode_launch(const void *values, int dtype, int64_t n, int p, int64_t ld, int row_major, uint8_t *codes,
                       int64_t stride, uint8_t *dims, int32_t *status, hipStream_t s);
int disc_init_launch(int p, int max_level, int max_dim, uint8_t *Gm, int32_t *sep, float *pmax, void *ws,
                     unsigned long long *st, int nst, hipStream_t s);
int disc_level0_launch(const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int p,
                       int max_level, int max_dim, double alpha, uint8_t *Gm, float *pmax,
                       unsigned long long *st, void *ws, size_t ws_bytes, hipStream_t s);
int disc_level_launch(int level, const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int p,
                      int max_level, int max_dim, double alpha, uint8_t *Gm, int32_t *sep, float *pmax,
                      unsigned long long *st, void *ws, size_t ws_bytes, hipStream_t s);
int disc_single_test_launch(const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int vi, int vj,
                            const int32_t *S, int l, int sequential, uint32_t *table_out, double *out,
                            hipStream_t s);
int lower_minus_one_launch(float *pmax, int p, hipStream_t s);
int disc_batch_init_launch(const gpucsl_disc_job *jobs, int nj, hipStream_t s);
int disc_batch_level0_launch(const gpucsl_disc_job *jobs, int nj, hipStream_t s);
int disc_batch_level_launch(int level, const gpucsl_disc_job *jobs, int nj, hipStream_t s);
int disc_batch_finish_launch(const gpucsl_disc_job *jobs, int nj, hipStream_t s);
int to_u16_launch(const uint8_t *Gm, int p, uint16_t *out, hipStream_t s);

// test hook: lexicographic unranking
template <int L>
__global__ void unrank_kernel(int N, const uint64_t *__restrict__ ranks, int count,
                              int32_t *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int pos[L];
    unrank_lex<L>(N, ranks[t], pos);
#pragma unroll
    for (int k = 0; k < L; ++k) out[(size_t)t * L + k] = pos[k];
}

__global__ void unrank_rt_kernel(int N, int l, const uint64_t *__restrict__ ranks, int count,
                                 int32_t *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int pos[GPUCSL_MAX_LEVEL];
    unrank_lex_rt(N, l, ranks[t], pos);
    for (int k = 0; k < l; ++k) out[(size_t)t * l + k] = pos[k];
}

// nbr/deg (library layout) -> reference Gc layout [deg, nbrs..., 0...]
__global__ void gc_layout_kernel(const int32_t *__restrict__ nbr, const int32_t *__restrict__ deg, int p,
                                 int32_t *__restrict__ Gc) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= p) return;
    const int d = deg[r];
    int32_t v;
    if (c == 0)
        v = d;
    else
        v = (c - 1 < d) ? nbr[(size_t)r * p + c - 1] : 0;
    Gc[(size_t)r * p + c] = v;
}

}  // namespace gpucsl

using namespace gpucsl;

static inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Levels above the built range only matter if some row still has enough
// neighbours for them: compact once more, read the max degree (one host sync).
template <typename CompactFn>
static int check_unbuilt_levels(int max_level, uint64_t *stats, CompactFn compact, hipStream_t s) {
    if (max_level <= GPUCSL_MAX_LEVEL) return OK;
    const int next = GPUCSL_MAX_LEVEL + 1;
    unsigned long long *st = reinterpret_cast<unsigned long long *>(stats) + (size_t)next * GPUCSL_STAT_SLOTS;
    int rc = compact(next, st);
    if (rc) return rc;
    unsigned long long maxdeg = 0;
    rc = check_hip(hipMemcpyAsync(&maxdeg, st + 5, sizeof(maxdeg), hipMemcpyDeviceToHost, s), "read max degree");
    if (rc) return rc;
    rc = check_hip(hipStreamSynchronize(s), "sync");
    if (rc) return rc;
    if ((long long)maxdeg - 1 >= next)
        return set_error(ERR_UNSUPPORTED, "level %d has work (max degree %llu) but only levels <= %d are built", next,
                         maxdeg, GPUCSL_MAX_LEVEL);
    return OK;
}

extern "C" {

const char *gpucsl_last_error(void) { return g_err; }

int gpucsl_abi_version(void) { return GPUCSL_ABI_VERSION; }

int gpucsl_max_built_level(void) { return GPUCSL_MAX_LEVEL; }

int gpucsl_device_info(int *compute_units, size_t *free_bytes, size_t *total_bytes) {
    int dev = 0;
    int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (compute_units) {
        rc = check_hip(hipDeviceGetAttribute(compute_units, hipDeviceAttributeMultiprocessorCount, dev),
                       "hipDeviceGetAttribute");
        if (rc) return rc;
    }
    size_t fr = 0, to = 0;
    rc = check_hip(hipMemGetInfo(&fr, &to), "hipMemGetInfo");
    if (rc) return rc;
    if (free_bytes) *free_bytes = fr;
    if (total_bytes) *total_bytes = to;
    return OK;
}

size_t gpucsl_corr_workspace_bytes(int64_t n, int p) { return corr_workspace_bytes(n, p); }

int gpucsl_corr(const void *X, int dtype, int64_t n, int p, double *C64, float *C32, void *ws,
                size_t ws_bytes, void *stream) {
    return corr_launch(X, dtype, n, p, C64, C32, ws, ws_bytes, S(stream));
}

int gpucsl_corr_c64_to_c32(const double *C64, int p, float *C32, void *stream) {
    return c64_to_c32_launch(C64, p, C32, S(stream));
}

size_t gpucsl_gauss_workspace_bytes(int p, int max_level) {
    (void)max_level;
    return gauss_workspace_bytes(p);
}

size_t gpucsl_gauss_job_workspace_bytes(int p) { return p < 1 ? 0 : job_ws_bytes(p); }

int gpucsl_corr_k_splits(int64_t n, int p, int dtype) {
    return n < 2 || p < 1 || (dtype != 0 && dtype != 1) ? 0 : corr_k_splits(n, p, dtype);
}

size_t gpucsl_corr_job_workspace_bytes(int64_t n, int p, int k_splits) {
    return n < 2 || p < 1 || k_splits < 1 ? 0 : corr_job_ws_bytes(n, p, k_splits);
}

int gpucsl_gauss_init(int p, int max_level, uint8_t *G, int32_t *sepsets, float *zmin,
                      uint64_t *stats, void *stream) {
    return gauss_init_launch(p, max_level, G, sepsets, zmin, reinterpret_cast<unsigned long long *>(stats),
                             stats ? (max_level + 1) * GPUCSL_STAT_SLOTS : 0, S(stream));
}

int gpucsl_gauss_level0(const float *C32, const double *C64, int p, double thr0, float margin,
                        uint8_t *G, float *zmin, uint64_t *stats_level0, void *ws, size_t ws_bytes, void *stream) {
    return gauss_level0_launch(C32, C64, p, thr0, margin, G, zmin,
                               reinterpret_cast<unsigned long long *>(stats_level0), ws, ws_bytes, S(stream));
}

int gpucsl_gauss_level(int level, const float *C32, const double *C64, int p, int max_level,
                       double thr, float margin, int group_size, uint8_t *G, int32_t *sepsets,
                       float *zmin, uint64_t *stats_level, void *ws, size_t ws_bytes, void *stream) {
    return gauss_level_launch(level, C32, C64, p, max_level, thr, margin, group_size, G, sepsets, zmin,
                              reinterpret_cast<unsigned long long *>(stats_level), ws, ws_bytes,
                              S(stream));
}

int gpucsl_gauss_skeleton(const float *C32, const double *C64, int p, int max_level,
                          const double *thr, float margin, int group_size, uint8_t *G,
                          int32_t *sepsets, float *zmin, uint64_t *stats, void *ws, size_t ws_bytes,
                          void *stream) {
    if (!thr || !stats) return set_error(ERR_ARG, "gauss_skeleton: null thr or stats");
    if (!C32 || !C64) return set_error(ERR_ARG, "gauss_skeleton: null correlation matrix");
    int rc = gauss_skeleton_launch(C32, C64, p, max_level, thr, margin, group_size, G, sepsets, zmin,
                                   reinterpret_cast<unsigned long long *>(stats), (max_level + 1) * GPUCSL_STAT_SLOTS,
                                   ws, ws_bytes, S(stream));
    if (rc) return rc;
    return check_unbuilt_levels(
        max_level, stats,
        [&](int lv, unsigned long long *st) { return compact_launch(G, p, lv, ws, st, S(stream)); }, S(stream));
}

int gpucsl_gauss_skeleton_noalloc(const float *C32, const double *C64, int p, int max_level, const double *thr,
                                  float margin, int group_size, uint8_t *G, int32_t *sepsets, float *zmin,
                                  uint64_t *stats, void *stream) {
    if (!thr || !stats) return set_error(ERR_ARG, "gauss_skeleton_noalloc: null thr or stats");
    if (!C32 || !C64 || !G || !zmin || (max_level > 0 && !sepsets))
        return set_error(ERR_ARG, "gauss_skeleton_noalloc: null matrix");
    return gauss_skeleton_noalloc_launch(C32, C64, p, max_level, thr, margin, group_size, G, sepsets, zmin,
                                         reinterpret_cast<unsigned long long *>(stats),
                                         (max_level + 1) * GPUCSL_STAT_SLOTS, S(stream));
}

// ---- batches of Gaussian jobs ----------------------------------------------
static int check_batch(const gpucsl_gauss_job *jobs, const gpucsl_gauss_job *jobs_dev, int nj) {
    if (nj < 1 || !jobs || !jobs_dev) return set_error(ERR_ARG, "batch: empty job table");
    for (int b = 0; b < nj; ++b)
        if (jobs[b].p < 1 || jobs[b].max_level < 0 || jobs[b].n < 4)
            return set_error(ERR_ARG, "batch: job %d has p=%d max_level=%d n=%lld", b, jobs[b].p, jobs[b].max_level,
                             (long long)jobs[b].n);
    return OK;
}

size_t gpucsl_gauss_batch_workspace_bytes(const gpucsl_gauss_job *jobs, int nj) {
    return nj < 1 || !jobs ? 0 : gauss_batch_ws_bytes(jobs, nj);
}

int gpucsl_gauss_batch_corr(const gpucsl_gauss_job *jobs, const gpucsl_gauss_job *jobs_dev, int nj, void *stream) {
    int rc = check_batch(jobs, jobs_dev, nj);
    if (rc) return rc;
    int with_x = 0;
    for (int b = 0; b < nj; ++b) with_x += jobs[b].X != nullptr;
    if (with_x == nj) return corr_batch_launch(jobs, jobs_dev, nj, S(stream));
    if (with_x == 0) return c64_to_c32_batch_launch(jobs, jobs_dev, nj, S(stream));
    return set_error(ERR_ARG, "batch_corr: either every job or no job has X");
}

int gpucsl_gauss_batch_init(const gpucsl_gauss_job *jobs, const gpucsl_gauss_job *jobs_dev, int nj, void *stream) {
    int rc = check_batch(jobs, jobs_dev, nj);
    if (rc) return rc;
    return gauss_batch_init(jobs, jobs_dev, nj, S(stream));
}

int gpucsl_gauss_batch_level0(const gpucsl_gauss_job *jobs, const gpucsl_gauss_job *jobs_dev, int nj, float margin,
                              uint64_t *stats_level0, void *stream) {
    int rc = check_batch(jobs, jobs_dev, nj);
    if (rc) return rc;
    if (!stats_level0) return set_error(ERR_ARG, "batch_level0: null stats");
    return gauss_batch_level0(jobs, jobs_dev, nj, margin, reinterpret_cast<unsigned long long *>(stats_level0),
                              S(stream));
}

int gpucsl_gauss_batch_level(int level, const gpucsl_gauss_job *jobs, const gpucsl_gauss_job *jobs_dev, int nj,
                             float margin, int group_size, uint64_t *stats_level, void *ws, size_t ws_bytes,
                             void *stream) {
    int rc = check_batch(jobs, jobs_dev, nj);
    if (rc) return rc;
    return gauss_batch_level(level, jobs, jobs_dev, nj, margin, group_size,
                             reinterpret_cast<unsigned long long *>(stats_level), ws, ws_bytes, S(stream));
}

int gpucsl_gauss_batch_finish(const gpucsl_gauss_job *jobs, const gpucsl_gauss_job *jobs_dev, int nj, void *stream) {
    int rc = check_batch(jobs, jobs_dev, nj);
    if (rc) return rc;
    return gauss_batch_finish(jobs, jobs_dev, nj, S(stream));
}

int gpucsl_gauss_batch_skeleton(const gpucsl_gauss_job *jobs, const gpucsl_gauss_job *jobs_dev, int nj, float margin,
                                int group_size, uint64_t *stats, void *ws, size_t ws_bytes, void *stream) {
    int rc = check_batch(jobs, jobs_dev, nj);
    if (rc) return rc;
    if (!stats || !ws) return set_error(ERR_ARG, "batch_skeleton: null stats or workspace");
    int max_level = 0;
    for (int b = 0; b < nj; ++b) max_level = jobs[b].max_level > max_level ? jobs[b].max_level : max_level;
    launch_fill<uint64_t>(stats, 0ull, (size_t)(max_level + 1) * GPUCSL_STAT_SLOTS, S(stream));
    if ((rc = gpucsl_gauss_batch_corr(jobs, jobs_dev, nj, stream))) return rc;
    if ((rc = gauss_batch_init(jobs, jobs_dev, nj, S(stream)))) return rc;
    if ((rc = gauss_batch_level0(jobs, jobs_dev, nj, margin, reinterpret_cast<unsigned long long *>(stats),
                                 S(stream))))
        return rc;
    const int built = max_level < GPUCSL_MAX_LEVEL ? max_level : GPUCSL_MAX_LEVEL;
    for (int l = 1; l <= built; ++l) {
        rc = gauss_batch_level(l, jobs, jobs_dev, nj, margin, group_size,
                               reinterpret_cast<unsigned long long *>(stats) + (size_t)l * GPUCSL_STAT_SLOTS, ws,
                               ws_bytes, S(stream));
        if (rc) return rc;
    }
    rc = check_unbuilt_levels(
        max_level, stats,
        [&](int lv, unsigned long long *st) { return gauss_batch_compact(jobs, jobs_dev, nj, lv, group_size, st, S(stream)); },
        S(stream));
    if (rc) return rc;
    return gauss_batch_finish(jobs, jobs_dev, nj, S(stream));
}

int gpucsl_set_level_tuning(int level, const gpucsl_level_tuning *tuning) { return set_level_tuning(-1, level, tuning); }

int gpucsl_get_level_tuning(int level, gpucsl_level_tuning *tuning) {
    return get_level_tuning(GPUCSL_TUNING_LARGE, level, tuning);
}

int gpucsl_set_bucket_tuning(int bucket, int level, const gpucsl_level_tuning *tuning) {
    return set_level_tuning(bucket, level, tuning);
}

int gpucsl_get_bucket_tuning(int bucket, int level, gpucsl_level_tuning *tuning) {
    return get_level_tuning(bucket, level, tuning);
}

int gpucsl_tuning_bucket(int jobs, int max_p) {
    if (jobs < 1 || max_p < 1) return set_error(ERR_ARG, "tuning_bucket: jobs and max_p must be >= 1"), -1;
    return tuning_bucket(jobs, max_p);
}

int gpucsl_pmax_from_zmin(const float *zmin, int p, int64_t n, float *pmax, void *stream) {
    return pmax_launch(zmin, p, n, pmax, S(stream));
}

int gpucsl_compact(const uint8_t *G, int p, int32_t *Gc, void *ws, size_t ws_bytes, void *stream) {
    if (!G || !Gc || !ws || p < 1) return set_error(ERR_ARG, "compact: bad arguments");
    if (ws_bytes < gauss_workspace_bytes(p)) return set_error(ERR_WORKSPACE, "compact: workspace too small");
    int rc = compact_launch(G, p, 1, ws, nullptr, S(stream));
    if (rc) return rc;
    int32_t *nbr = nullptr, *deg = nullptr;
    level_list_views(ws, p, &nbr, &deg);
    hipLaunchKernelGGL(gc_layout_kernel, dim3((p + 255) / 256, p), dim3(256), 0, S(stream), nbr, deg, p, Gc);
    return check_hip(hipGetLastError(), "gc layout");
}

size_t gpucsl_disc_workspace_bytes(int p, int max_level, int max_dim) {
    return disc_workspace_bytes(p, max_level, max_dim);
}

size_t gpucsl_disc_min_workspace_bytes(int p, int max_level, int max_dim) {
    return disc_min_workspace_bytes(p, max_level, max_dim);
}

int gpucsl_disc_init(int p, int max_level, int max_dim, uint8_t *G, int32_t *sepsets, float *pmax,
                     uint64_t *stats, void *ws, size_t ws_bytes, void *stream) {
    if (ws_bytes < disc_min_workspace_bytes(p, max_level, max_dim))
        return set_error(ERR_WORKSPACE, "disc_init: workspace too small");
    return disc_init_launch(p, max_level, max_dim, G, sepsets, pmax, ws,
                            reinterpret_cast<unsigned long long *>(stats),
                            stats ? (max_level + 1) * GPUCSL_STAT_SLOTS : 0, S(stream));
}

int gpucsl_disc_level0(const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int p, int max_level,
                       int max_dim, double alpha, uint8_t *G, float *pmax, uint64_t *stats_level0, void *ws,
                       size_t ws_bytes, void *stream) {
    return disc_level0_launch(data, n, stride, dims, p, max_level, max_dim, alpha, G, pmax,
                              reinterpret_cast<unsigned long long *>(stats_level0), ws, ws_bytes, S(stream));
}

int gpucsl_disc_level(int level, const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int p,
                      int max_level, int max_dim, double alpha, uint8_t *G, int32_t *sepsets, float *pmax,
                      uint64_t *stats_level, void *ws, size_t ws_bytes, void *stream) {
    return disc_level_launch(level, data, n, stride, dims, p, max_level, max_dim, alpha, G, sepsets, pmax,
                             reinterpret_cast<unsigned long long *>(stats_level), ws, ws_bytes, S(stream));
}

int gpucsl_disc_skeleton(const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int p,
                         int max_level, int max_dim, double alpha, uint8_t *G, int32_t *sepsets, float *pmax,
                         uint64_t *stats, void *ws, size_t ws_bytes, void *stream) {
    if (!stats) return set_error(ERR_ARG, "disc_skeleton: null stats");
    int rc = gpucsl_disc_init(p, max_level, max_dim, G, sepsets, pmax, stats, ws, ws_bytes, stream);
    if (rc) return rc;
    rc = gpucsl_disc_level0(data, n, stride, dims, p, max_level, max_dim, alpha, G, pmax, stats, ws, ws_bytes,
                            stream);
    if (rc) return rc;
    const int built = max_level < GPUCSL_MAX_LEVEL ? max_level : GPUCSL_MAX_LEVEL;
    for (int l = 1; l <= built; ++l) {
        rc = gpucsl_disc_level(l, data, n, stride, dims, p, max_level, max_dim, alpha, G, sepsets, pmax,
                               stats + (size_t)l * GPUCSL_STAT_SLOTS, ws, ws_bytes, stream);
        if (rc) return rc;
    }
    rc = check_unbuilt_levels(
        max_level, stats,
        [&](int lv, unsigned long long *st) { return compact_launch(G, p, lv, ws, st, S(stream)); }, S(stream));
    if (rc) return rc;
    return lower_minus_one_launch(pmax, p, S(stream));
}

int gpucsl_disc_batch_init(const gpucsl_disc_job *jobs, int nj, void *stream) {
    return disc_batch_init_launch(jobs, nj, S(stream));
}

int gpucsl_disc_batch_level0(const gpucsl_disc_job *jobs, int nj, void *stream) {
    return disc_batch_level0_launch(jobs, nj, S(stream));
}

int gpucsl_disc_batch_level(int level, const gpucsl_disc_job *jobs, int nj, void *stream) {
    return disc_batch_level_launch(level, jobs, nj, S(stream));
}

int gpucsl_disc_batch_finish(const gpucsl_disc_job *jobs, int nj, void *stream) {
    return disc_batch_finish_launch(jobs, nj, S(stream));
}

int gpucsl_disc_batch_skeleton(const gpucsl_disc_job *jobs, int nj, void *stream) {
    int rc = gpucsl_disc_batch_init(jobs, nj, stream);
    if (rc) return rc;
    rc = gpucsl_disc_batch_level0(jobs, nj, stream);
    if (rc) return rc;
    int top = 0;
    for (int b = 0; b < nj; ++b) top = jobs[b].max_level > top ? jobs[b].max_level : top;
    const int built = top < GPUCSL_MAX_LEVEL ? top : GPUCSL_MAX_LEVEL;
    for (int l = 1; l <= built; ++l) {
        rc = gpucsl_disc_batch_level(l, jobs, nj, stream);
        if (rc) return rc;
    }
    for (int b = 0; b < nj; ++b) {
        const gpucsl_disc_job &J = jobs[b];
        rc = check_unbuilt_levels(
            J.max_level, J.stats,
            [&](int lv, unsigned long long *st) { return compact_launch(J.G, J.p, lv, J.ws, st, S(stream)); },
            S(stream));
        if (rc) return rc;
    }
    return gpucsl_disc_batch_finish(jobs, nj, stream);
}

int gpucsl_disc_encode(const void *values, int dtype, int64_t n, int p, int64_t ld, int row_major, uint8_t *codes,
                       int64_t stride, uint8_t *dims, int32_t *status, void *stream) {
    return disc_encode_launch(values, dtype, n, p, ld, row_major, codes, stride, dims, status, S(stream));
}

int gpucsl_disc_finish(const uint8_t *G, int p, float *pmax, uint16_t *skeleton, void *stream) {
    if (!G || !pmax || p < 1) return set_error(ERR_ARG, "disc_finish: bad arguments");
    int rc = lower_minus_one_launch(pmax, p, S(stream));
    if (rc) return rc;
    return skeleton ? to_u16_launch(G, p, skeleton, S(stream)) : OK;
}

int gpucsl_disc_ci_test(const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int vi, int vj,
                        const int32_t *sepset, int l, int sequential, uint32_t *table_out, double *out,
                        void *stream) {
    if (!data || !dims || !table_out || !out || (l > 0 && !sepset)) return set_error(ERR_ARG, "disc_ci_test: null pointer");
    return disc_single_test_launch(data, n, stride, dims, vi, vj, sepset, l, sequential, table_out, out, S(stream));
}

int gpucsl_profile_begin(void) {
    std::lock_guard<std::mutex> lk(prof::g_mu);
    for (auto &r : prof::g_recs) {
        prof::g_pool.push_back(r.a);
        prof::g_pool.push_back(r.b);
    }
    prof::g_recs.clear();
    prof::g_on = true;
    return OK;
}

int gpucsl_profile_end(char *out, size_t cap) {
    std::vector<prof::Rec> recs;
    {
        std::lock_guard<std::mutex> lk(prof::g_mu);
        prof::g_on = false;
        recs.swap(prof::g_recs);
    }
    struct Agg {
        long long count = 0;
        double total = 0.0, mn = 1e300, mx = 0.0;
    };
    std::map<std::string, Agg> agg;
    for (auto &r : recs) {
        int rc = check_hip(hipEventSynchronize(r.b), "profile sync");
        if (rc) return rc;
        float ms = 0.0f;
        rc = check_hip(hipEventElapsedTime(&ms, r.a, r.b), "profile elapsed");
        if (rc) return rc;
        Agg &a = agg[r.name];
        a.count += 1;
        a.total += ms;
        a.mn = ms < a.mn ? ms : a.mn;
        a.mx = ms > a.mx ? ms : a.mx;
    }
    {
        std::lock_guard<std::mutex> lk(prof::g_mu);
        for (auto &r : recs) {
            prof::g_pool.push_back(r.a);
            prof::g_pool.push_back(r.b);
        }
    }
    std::string txt;
    char line[256];
    for (auto &kv : agg) {
        snprintf(line, sizeof(line), "%s\t%lld\t%.6f\t%.6f\t%.6f\n", kv.first.c_str(), kv.second.count,
                 kv.second.total, kv.second.mn, kv.second.mx);
        txt += line;
    }
    if (out && cap) {
        size_t n = txt.size() < cap - 1 ? txt.size() : cap - 1;
        memcpy(out, txt.data(), n);
        out[n] = 0;
        if (txt.size() >= cap) return set_error(ERR_ARG, "profile_end: buffer too small (%zu needed)", txt.size() + 1);
    }
    return OK;
}

int gpucsl_gauss_outputs(const float *zmin, const uint8_t *G, int p, int64_t n, float *pmax, uint16_t *skeleton,
                         void *stream) {
    return outputs_launch(zmin, G, p, n, pmax, skeleton, S(stream));
}

int gpucsl_skeleton_to_u16(const uint8_t *G, int p, uint16_t *out, void *stream) {
    if (!G || !out || p < 1) return set_error(ERR_ARG, "skeleton_to_u16: bad arguments");
    return to_u16_launch(G, p, out, S(stream));
}

int gpucsl_unrank(int N, int l, const uint64_t *ranks, int count, int32_t *out, void *stream) {
    if (!ranks || !out || count < 0 || N < 0) return set_error(ERR_ARG, "unrank: bad arguments");
    if (count == 0) return OK;
    dim3 grid((count + 255) / 256);
    switch (l) {
#define GPUCSL_U(LV)                                                                                 \
    case LV:                                                                                         \
        hipLaunchKernelGGL(unrank_kernel<LV>, grid, dim3(256), 0, S(stream), N, ranks, count, out);  \
        break;
        GPUCSL_U(1) GPUCSL_U(2) GPUCSL_U(3) GPUCSL_U(4) GPUCSL_U(5) GPUCSL_U(6) GPUCSL_U(7) GPUCSL_U(8)
#undef GPUCSL_U
    default:
        if (l < 1 || l > GPUCSL_MAX_LEVEL) return set_error(ERR_ARG, "unrank: l must be in 1..%d", GPUCSL_MAX_LEVEL);
        hipLaunchKernelGGL(unrank_rt_kernel, grid, dim3(256), 0, S(stream), N, l, ranks, count, out);
        break;
    }
    return check_hip(hipGetLastError(), "unrank launch");
}

}  // extern "C"
