/*
 * gpucsl_hip.h -- C ABI of libgpucsl_hip.so, the gfx950 PC-stable skeleton
 * library.
 *
 * Conventions
 *   - Every call returns 0 on success or a non-zero status (1 bad argument,
 *     2 HIP error, 3 unsupported, 4 workspace too small); gpucsl_last_error()
 *     returns a thread-local message for the last failure.
 *   - Every device pointer is caller-owned device memory.  The library never
 *     allocates or frees caller memory; workspaces are sized by the
 *     *_workspace_bytes functions and passed in.
 *   - Every call is stream-ordered on the caller's hipStream_t (passed as
 *     void*, NULL = default stream) and does not synchronise the host, except
 *     gpucsl_gauss_skeleton / gpucsl_disc_skeleton when max_level exceeds the
 *     built level range (they then read the per-level max degree once).
 *   - Matrices are row-major.  G is uint8 p*p (1 edge, 0 none); separation sets
 *     are int32 p*p*max_level, -1 padded, entry [(i*p + j)*max_level + k]; zmin
 *     and pmax are float p*p with the upper triangle meaningful.
 *   - stats is a device array of uint64 [(max_level+1) * 8]; slot s of level l
 *     is stats[l*8 + s]: 0 executed CI tests, 1 full enumeration size T_l,
 *     2 FP64 rechecks, 3 rechecks whose FP32 decision differed, 4 max
 *     |z32 - z64| / error-estimate over rechecks (float bits), 5 max degree,
 *     6 tests with key <= the removing test's key (the count a sequential
 *     pcalg-order loop runs; executed >= this because lane groups evaluate a
 *     whole iteration before they see the first pass).
 *
 * Reference interfaces (paths in the gpucsl repository) are cited per call.
 */
#ifndef GPUCSL_HIP_H
#define GPUCSL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPUCSL_ABI_VERSION 2
#define GPUCSL_STAT_SLOTS 8

/* Thread-local description of the last failed call. */
const char *gpucsl_last_error(void);

/* ABI version (GPUCSL_ABI_VERSION) and the highest level the kernels run
 * (GPUCSL_MAX_LEVEL).  Levels 1..8 have kernels specialised per level; levels
 * 9..GPUCSL_MAX_LEVEL run a kernel that takes the level at run time (the
 * reference instantiates every level at run time,
 * gpucsl/pc/kernel_management.py:211-212). */
int gpucsl_abi_version(void);
int gpucsl_max_built_level(void);

/* Device properties used for planning: compute units, free and total bytes.
 * Replaces the CuPy queries of gpucsl/pc/kernel_management.py:394-400 and
 * gpucsl/pc/pc_validations.py:30-42. */
int gpucsl_device_info(int *compute_units, size_t *free_bytes, size_t *total_bytes);

/* ---------------------------------------------------------------------------
 * correlation: replaces correlation_matrix_of / np.corrcoef
 * (gpucsl/pc/helpers.py:58-59) and the host-to-device copy of C
 * (gpucsl/pc/gaussian_device_manager.py:114-116).
 * X: n x p row-major, dtype 0 = float32, 1 = float64.  C64 may be NULL.
 * ------------------------------------------------------------------------- */
size_t gpucsl_corr_workspace_bytes(int64_t n, int p);
int gpucsl_corr(const void *X, int dtype, int64_t n, int p, double *C64, float *C32, void *ws,
                size_t ws_bytes, void *stream);
/* C32 = (float)C64 for a caller-supplied correlation matrix
 * (the correlation_matrix argument of gpucsl/pc/pc.py:88-102). */
int gpucsl_corr_c64_to_c32(const double *C64, int p, float *C32, void *stream);

/* ---------------------------------------------------------------------------
 * Gaussian skeleton
 * ------------------------------------------------------------------------- */
#define GPUCSL_MAX_LEVEL 32   /* highest level any kernel runs */
#define GPUCSL_MAX_LEVELS 33  /* thresholds for levels 0..GPUCSL_MAX_LEVEL */

/* One Gaussian skeleton job (one dataset).  Every pointer is device memory
 * owned by the caller: ws is a per-job level workspace of
 * gpucsl_gauss_job_workspace_bytes(p) bytes and corr_ws a correlation
 * workspace of gpucsl_corr_job_workspace_bytes(n, p, k_splits) bytes.  X (n x p
 * row-major, x_dtype 0 = float32, 1 = float64) is only read by the correlation
 * step and may be NULL when C64 is supplied.  The single-dataset entry points
 * below build one of these internally; the batched entry points take an array
 * of them, the same array in host memory and in device memory. */
/* pcalg's conventions instead of gpucsl's: the caller passes thresholds
 * |Phi^-1(alpha/2)| / sqrt(n - l - 3); zmin holds min over tests of
 * sqrt(n - |S| - 3) |z| (so pmax = 2 (1 - Phi(zmin)) is pcalg's pMax, the
 * largest p-value of any test), and a sepset is stored only for the ordered
 * pair whose test removed the edge. */
#define GPUCSL_JOB_PCALG_COMPAT 1

typedef struct gpucsl_gauss_job {
    int32_t p;
    int32_t max_level;
    int64_t n;
    int32_t x_dtype;
    int32_t k_splits;   /* split-K factor of the correlation GEMM for this job */
    int32_t flags;      /* GPUCSL_JOB_* bits */
    int32_t reserved;
    const void *X;
    float *C32;
    double *C64;
    uint8_t *G;
    int32_t *sepsets;
    float *zmin;
    float *pmax;
    uint16_t *skeleton; /* may be NULL; filled by the finish step */
    void *ws;
    void *corr_ws;
    double thr[GPUCSL_MAX_LEVELS];
} gpucsl_gauss_job;

/* Per-job workspaces of the batched entry points, and the split-K factor the
 * correlation GEMM uses for an n x p dataset of x_dtype dtype (1..16, chosen
 * from n, p, dtype and the current device's compute-unit count, and at least
 * n / 65536 for float32 data with p > 256, whose integer sums are exact up to
 * 65 536 rows per split; a batched job with fewer splits is refused; the same for a
 * dataset alone and in any batch on that device, which keeps its correlation
 * bits the same; 0 for bad arguments).  The correlation workspace holds the
 * column sums, the integer digit planes of float32 data (6 bytes per element,
 * rows and columns padded to multiples of 64) and k_splits partial Gram
 * matrices. */
size_t gpucsl_gauss_job_workspace_bytes(int p);
int gpucsl_corr_k_splits(int64_t n, int p, int dtype);
size_t gpucsl_corr_job_workspace_bytes(int64_t n, int p, int k_splits);
/* Workspace of the single-dataset entry points (level lists, work queues). */
size_t gpucsl_gauss_workspace_bytes(int p, int max_level);

/* G = 1, sepsets = -1, zmin = +inf, stats = 0.  Replaces the buffer set-up of
 * gpucsl/pc/gaussian_device_manager.py:106-133. */
int gpucsl_gauss_init(int p, int max_level, uint8_t *G, int32_t *sepsets, float *zmin,
                      uint64_t *stats, void *stream);

/* Level 0: Fisher-z of every pair, removal on z < thr0, zmin, diagonal
 * cleared.  Replaces gaussian_ci_level_0 (gpucsl/pc/cuda/gaussian_ci.cu:18-51).
 * margin scales the FP32 error estimate inside which the FP64 recheck runs. */
int gpucsl_gauss_level0(const float *C32, const double *C64, int p, double thr0, float margin,
                        uint8_t *G, float *zmin, uint64_t *stats_level0, void *ws, size_t ws_bytes,
                        void *stream);

/* Level l >= 1: compaction, then every CI test with |S| = l in pcalg order.
 * Replaces one iteration of gaussian_ci_worker
 * (gpucsl/pc/discover_skeleton_gaussian.py:64-102): compact
 * (gpucsl/pc/cuda/helpers/graph_helpers.cu:35-56) and gaussian_ci_level_n
 * (gpucsl/pc/cuda/gaussian_ci.cu:66-120).  group_size in {8, 16, 32, 64} lanes
 * per ordered pair.  stats_level points at this level's 8 slots. */
int gpucsl_gauss_level(int level, const float *C32, const double *C64, int p, int max_level,
                       double thr, float margin, int group_size, uint8_t *G, int32_t *sepsets,
                       float *zmin, uint64_t *stats_level, void *ws, size_t ws_bytes, void *stream);

/* Levels 0..max_level back to back with no host synchronisation
 * (gaussian_ci_worker, gpucsl/pc/discover_skeleton_gaussian.py:30-113).
 * thr: HOST array of max_level+1 thresholds (gpucsl/pc/helpers.py:42-50). */
int gpucsl_gauss_skeleton(const float *C32, const double *C64, int p, int max_level,
                          const double *thr, float margin, int group_size, uint8_t *G,
                          int32_t *sepsets, float *zmin, uint64_t *stats, void *ws, size_t ws_bytes,
                          void *stream);

/* gpucsl_gauss_skeleton without a workspace argument: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * workspace is a library array of GPUCSL_NOALLOC_SCRATCH_BYTES per device,
 * held until the process ends; p is limited to what it holds
 * (gpucsl_gauss_workspace_bytes(p, max_level) <= GPUCSL_NOALLOC_SCRATCH_BYTES,
 * p <= 1234), else status 4 (workspace) and nothing runs; max_level <=
 * GPUCSL_MAX_LEVEL, else status 3.  Calls no allocating routine (and records
 * no timing events).  All calls on one device share its array: calls that
 * overlap on one device race and must be ordered by the caller (one stream,
 * or events).  Results are the same bits as gpucsl_gauss_skeleton's. */
#define GPUCSL_NOALLOC_SCRATCH_BYTES (64u << 20)
int gpucsl_gauss_skeleton_noalloc(const float *C32, const double *C64, int p, int max_level, const double *thr,
                                  float margin, int group_size, uint8_t *G, int32_t *sepsets, float *zmin,
                                  uint64_t *stats, void *stream);

/* ---------------------------------------------------------------------------
 * Batches of Gaussian jobs.  Every call takes the job table twice: `jobs` in
 * host memory (sizes and validation) and `jobs_dev`, the same bytes in device
 * memory (read by the kernels).  All jobs of a batch advance level by level in
 * shared launches: one launch per kernel and level for the whole batch.  A job
 * whose max_level is below the level of a launch takes no part in it.  The
 * launch workspace `ws` holds the cross-job work queues
 * (gpucsl_gauss_batch_workspace_bytes).  Results of a job do not depend on the
 * other jobs of its batch.  This replaces running gaussian_ci_worker
 * (gpucsl/pc/discover_skeleton_gaussian.py:30-113) once per dataset.
 * ------------------------------------------------------------------------- */
size_t gpucsl_gauss_batch_workspace_bytes(const gpucsl_gauss_job *jobs, int nj);
/* Correlation of every job: from X when every job has X, else C32 = (float)C64. */
int gpucsl_gauss_batch_corr(const gpucsl_gauss_job *jobs, const gpucsl_gauss_job *jobs_dev, int nj, void *stream);
/* G = 1, sepsets = -1, zmin = +inf for every job. */
int gpucsl_gauss_batch_init(const gpucsl_gauss_job *jobs, const gpucsl_gauss_job *jobs_dev, int nj, void *stream);
/* Level 0 of every job with thr[0]; stats summed over the jobs. */
int gpucsl_gauss_batch_level0(const gpucsl_gauss_job *jobs, const gpucsl_gauss_job *jobs_dev, int nj, float margin,
                              uint64_t *stats_level0, void *stream);
/* Level l >= 1 of every job with max_level >= l, thresholds thr[l]. */
int gpucsl_gauss_batch_level(int level, const gpucsl_gauss_job *jobs, const gpucsl_gauss_job *jobs_dev, int nj,
                             float margin, int group_size, uint64_t *stats_level, void *ws, size_t ws_bytes,
                             void *stream);
/* pmax from zmin (when pmax is set) and the uint16 skeleton (when set). */
int gpucsl_gauss_batch_finish(const gpucsl_gauss_job *jobs, const gpucsl_gauss_job *jobs_dev, int nj, void *stream);
/* corr + init + levels 0..max(max_level) + finish; stats: (max level + 1) x 8. */
int gpucsl_gauss_batch_skeleton(const gpucsl_gauss_job *jobs, const gpucsl_gauss_job *jobs_dev, int nj, float margin,
                                int group_size, uint64_t *stats, void *ws, size_t ws_bytes, void *stream);

/* Launch shapes of the level-l CI-test kernel (no counterpart in the
 * reference, whose kernel_management.py:61-65 fixes threads per pair).  The
 * lanes that test one pair together are group_size while C(deg-1, l) <=
 * group_size * small_team_iters, 64 while it is <= 64 * wave_team_iters, else
 * 256; fast_blocks_per_cu and fast_lds_bytes (4096..65536, a multiple of 1024)
 * size the grid and its LDS staging; a work item holds about item_tests
 * (64..2^24) subset tests; wave_items = 1 lets every wave take whole items on
 * its own (64 lanes per pair, reads from L2, no LDS staging).  Results do not
 * depend on these values.  level 0 sets every level.  The table is
 * process-wide; launches read it at launch time, so set it before they run.  Defaults:
 * {8, 8, 4, 32768, 4096, 0}. */
typedef struct gpucsl_level_tuning {
    int32_t small_team_iters;
    int32_t wave_team_iters;
    int32_t fast_blocks_per_cu;
    int32_t fast_lds_bytes;
    int32_t item_tests;
    int32_t wave_items;
} gpucsl_level_tuning;
int gpucsl_set_level_tuning(int level, const gpucsl_level_tuning *tuning);
int gpucsl_get_level_tuning(int level, gpucsl_level_tuning *tuning);

/* The table is kept per launch bucket: GPUCSL_TUNING_SMALL (one dataset, p <=
 * 256), GPUCSL_TUNING_LARGE (one dataset, p > 256) and GPUCSL_TUNING_BATCH
 * (several datasets in one launch).  gpucsl_set_level_tuning sets every
 * bucket, gpucsl_get_level_tuning reads GPUCSL_TUNING_LARGE; the calls below
 * address one bucket (level 0 in a set: every level).  A row's degree is
 * handled on the device: its subset count C(deg - 1, l) picks the team size
 * against the shape's thresholds.  gpucsl_tuning_bucket gives the bucket of a
 * launch of `jobs` datasets whose largest has max_p variables (-1 on bad
 * arguments). */
#define GPUCSL_TUNING_SMALL 0
#define GPUCSL_TUNING_LARGE 1
#define GPUCSL_TUNING_BATCH 2
#define GPUCSL_TUNING_BUCKETS 3
int gpucsl_set_bucket_tuning(int bucket, int level, const gpucsl_level_tuning *tuning);
int gpucsl_get_bucket_tuning(int bucket, int level, gpucsl_level_tuning *tuning);
int gpucsl_tuning_bucket(int jobs, int max_p);

/* pmax = 2 (1 - Phi(sqrt(n-3) |zmin|)), -1 on and below the diagonal.
 * Replaces transform_to_pmax_cupy + postprocess_pmax_cupy
 * (gpucsl/pc/helpers.py:62-82). */
int gpucsl_pmax_from_zmin(const float *zmin, int p, int64_t n, float *pmax, void *stream);

/* ---------------------------------------------------------------------------
 * Edge orientation (host function; all pointers are HOST memory).
 * skeleton: p x p bytes, symmetric 0/1; sepsets: p x p x max_level int32, -1
 * padded (NULL when max_level is 0); dag: p x p bytes out, dag[i*p+j] = arc
 * i -> j (both arcs = undirected edge).  v-structures, then Meek R1-R3 until
 * no rule fires, in the visiting order of gpucsl/pc/edge_orientation/
 * (v_structures.py:8-35, orientation_rules.py:7-89).  passes (may be NULL)
 * receives the number of rule passes.  Returns 0, or 1 on bad arguments.
 * ------------------------------------------------------------------------- */
int gpucsl_orient_edges(const uint8_t *skeleton, int p, const int32_t *sepsets, int max_level, uint8_t *dag,
                        int *passes);
/* The same orientation, and the CPDAG's arcs as (i, j) int32 pairs in
 * row-major order (i ascending, then j): at most arcs_cap pairs are written,
 * *n_arcs receives the number of arcs (at most the skeleton's nonzero count).
 * Replaces the DiGraph the reference builds edge by edge
 * (gpucsl/pc/edge_orientation/edge_orientation.py:25-36); the Python layer
 * builds its networkx result from the pairs without scanning the p x p
 * bytes.  Returns 0, 1 on bad arguments, or 2 when arcs_cap is too small
 * (dag and *n_arcs are then complete). */
int gpucsl_orient_edges_arcs(const uint8_t *skeleton, int p, const int32_t *sepsets, int max_level, uint8_t *dag,
                             int32_t *arcs, int64_t arcs_cap, int64_t *n_arcs, int *passes);

/* ---------------------------------------------------------------------------
 * discrete skeleton (Pearson chi^2)
 * data: uint8 category codes, column-major with a padded column stride:
 * data[v*stride + k], stride a multiple of 16 and >= n rounded up to 16, data
 * 16-byte aligned; dims[v] = number of categories of v; max_dim = max dims.
 * ------------------------------------------------------------------------- */
/* Workspace of the discrete entry points; SIZE_MAX when the contingency tables
 * of max_level can never be held (the out-of-memory case of
 * find_execution_specification, gpucsl/pc/kernel_management.py:382-456). */
size_t gpucsl_disc_workspace_bytes(int p, int max_level, int max_dim);
/* The smallest workspace the discrete entry points accept: with less than
 * gpucsl_disc_workspace_bytes they run fewer tests at once (one global table
 * slice per running block), as find_execution_specification reduces
 * parallelism under a memory restriction. */
size_t gpucsl_disc_min_workspace_bytes(int p, int max_level, int max_dim);

/* Ordinal codes on the device, as sklearn's OrdinalEncoder: a value's code is
 * its rank among its column's sorted distinct values (replaces the encoding of
 * gpucsl/pc/discover_skeleton_discrete.py:38-39).  values: device n x p array
 * with element (k, v) at values[k*ld + v] (row_major = 1) or values[v*ld + k]
 * (row_major = 0); dtype 0 int32, 1 int64, 2 float32, 3 float64.  codes: p x
 * stride uint8 (stride >= n, a multiple of 16; padding set to 0); dims: p
 * category counts; status: p int32, 0 ok, 1 more than 255 distinct values
 * (dims 0, codes unset), 2 NaN present (codes unset). */
int gpucsl_disc_encode(const void *values, int dtype, int64_t n, int p, int64_t ld, int row_major, uint8_t *codes,
                       int64_t stride, uint8_t *dims, int32_t *status, void *stream);

/* G = 1, sepsets = -1, pmax = 0, stats = 0, removal keys cleared and the
 * level-0 table matrix zeroed, in one launch.  Replaces the buffer set-up of
 * gpucsl/pc/discover_skeleton_discrete.py:41-55. */
int gpucsl_disc_init(int p, int max_level, int max_dim, uint8_t *G, int32_t *sepsets, float *pmax,
                     uint64_t *stats, void *ws, size_t ws_bytes, void *stream);

/* Level 0: chi^2 test of every pair, removal on p >= alpha, pmax.  Replaces
 * discrete_ci_level_0 (gpucsl/pc/cuda/discrete_ci.cu:37-97).  Runs after
 * gpucsl_disc_init on the same workspace.  With max_dim <= 16 and
 * p * max_dim <= 4096 every pair's table comes from one int8-MFMA Gram matrix
 * of the one-hot codes (zeroed by gpucsl_disc_init and left zero by each run,
 * so level 0 may run again on the same workspace); otherwise tables are
 * counted pair by pair.  Results are the same either way. */
int gpucsl_disc_level0(const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int p,
                       int max_level, int max_dim, double alpha, uint8_t *G, float *pmax,
                       uint64_t *stats_level0, void *ws, size_t ws_bytes, void *stream);

/* Level l >= 1: compaction, every test of every level-start adjacent pair,
 * pmax, removal with the pcalg-order first sepset.  Replaces one iteration of
 * the level loop of gpucsl/pc/discover_skeleton_discrete.py:80-112 (compact +
 * discrete_ci_level_n, gpucsl/pc/cuda/discrete_ci.cu:185-246). */
int gpucsl_disc_level(int level, const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int p,
                      int max_level, int max_dim, double alpha, uint8_t *G, int32_t *sepsets, float *pmax,
                      uint64_t *stats_level, void *ws, size_t ws_bytes, void *stream);

/* Levels 0..max_level and the pmax post-processing (-1 on and below the
 * diagonal, gpucsl/pc/helpers.py:96-101), no host synchronisation.  Replaces
 * gpucsl/pc/discover_skeleton_discrete.py:24-128 after ordinal encoding. */
int gpucsl_disc_skeleton(const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int p,
                         int max_level, int max_dim, double alpha, uint8_t *G, int32_t *sepsets, float *pmax,
                         uint64_t *stats, void *ws, size_t ws_bytes, void *stream);
/* After level-by-level calls: pmax = -1 on and below the diagonal
 * (postprocess_pmax, gpucsl/pc/helpers.py:96-101) and, when skeleton is not
 * NULL, the uint16 skeleton (the dtype of gpucsl's SkeletonResult). */
int gpucsl_disc_finish(const uint8_t *G, int p, float *pmax, uint16_t *skeleton, void *stream);

/* One CI test, for known-answer checks: table_out gets the dims_s*di*dj counts
 * N[g][a][b], out[0..2] = chi^2, p value, degrees of freedom.  sepset is a
 * device array of l ids; sequential selects the level-0 summation order.  The
 * table and its marginals (dim_s*(di*dj + di + dj + 1) words) must fit 8192
 * words; a larger test is not counted and gets out = {NaN, NaN, -1}.
 * (gpucsl/pc/cuda/helpers/discrete_helpers.cu:25-186, math_helpers.cu:128-175) */
int gpucsl_disc_ci_test(const uint8_t *data, int64_t n, int64_t stride, const uint8_t *dims, int vi, int vj,
                        const int32_t *sepset, int l, int sequential, uint32_t *table_out, double *out,
                        void *stream);

/* ---------------------------------------------------------------------------
 * batched discrete skeletons: many datasets (or resamples) of uneven n, p and
 * categories advance level by level together, one launch per kernel of a
 * level for up to GPUCSL_DISC_BATCH_JOBS jobs (the job table travels in the
 * kernel arguments, so a captured launch keeps it).  No counterpart in gpucsl,
 * which runs gpucsl/pc/discover_skeleton_discrete.py:24-128 once per dataset;
 * each job's results are bit-identical to the single-dataset entry points on
 * the same buffers.  A job's buffers and workspace are those of the
 * single-dataset calls (gpucsl_disc_workspace_bytes or at least
 * gpucsl_disc_min_workspace_bytes); the tests of one launch run in as many
 * workgroups at once as the smallest workspace among its jobs allows.
 * ------------------------------------------------------------------------- */
#define GPUCSL_DISC_BATCH_JOBS 16

typedef struct gpucsl_disc_job {
    const uint8_t *data;   /* codes, column-major, padded stride (as above) */
    const uint8_t *dims;   /* p category counts */
    int64_t n;
    int64_t stride;
    int32_t p;
    int32_t max_level;
    int32_t max_dim;
    int32_t reserved;
    double alpha;
    uint8_t *G;            /* p x p */
    int32_t *sepsets;      /* p x p x max_level (may be NULL when max_level is 0) */
    float *pmax;           /* p x p */
    uint64_t *stats;       /* (max_level + 1) x GPUCSL_STAT_SLOTS */
    uint16_t *skeleton;    /* may be NULL; filled by gpucsl_disc_batch_finish */
    void *ws;
    size_t ws_bytes;
} gpucsl_disc_job;

/* gpucsl_disc_init, gpucsl_disc_level0, gpucsl_disc_level and
 * gpucsl_disc_finish for every job (a job whose max_level is below `level`
 * is left alone), and all of them in order. */
int gpucsl_disc_batch_init(const gpucsl_disc_job *jobs, int nj, void *stream);
int gpucsl_disc_batch_level0(const gpucsl_disc_job *jobs, int nj, void *stream);
int gpucsl_disc_batch_level(int level, const gpucsl_disc_job *jobs, int nj, void *stream);
int gpucsl_disc_batch_finish(const gpucsl_disc_job *jobs, int nj, void *stream);
int gpucsl_disc_batch_skeleton(const gpucsl_disc_job *jobs, int nj, void *stream);

/* ---------------------------------------------------------------------------
 * graph helpers
 * ------------------------------------------------------------------------- */
/* Reference layout of compact (gpucsl/pc/cuda/helpers/graph_helpers.cu:35-56):
 * row r of Gc = [deg, ascending neighbours..., 0...].  ws as for the Gaussian
 * skeleton. */
int gpucsl_compact(const uint8_t *G, int p, int32_t *Gc, void *ws, size_t ws_bytes, void *stream);

/* uint8 adjacency -> the uint16 skeleton the Python API returns
 * (gpucsl/pc/helpers.py:53-55). */
int gpucsl_skeleton_to_u16(const uint8_t *G, int p, uint16_t *out, void *stream);

/* gpucsl_pmax_from_zmin and gpucsl_skeleton_to_u16 of one dataset in one
 * launch, with the same results (gpucsl/pc/helpers.py:53-82). */
int gpucsl_gauss_outputs(const float *zmin, const uint8_t *G, int p, int64_t n, float *pmax, uint16_t *skeleton,
                         void *stream);

/* Lexicographic unranking used by the CI kernels, for tests: out[c*l + k] is
 * the k-th 1-based element of combination ranks[c] of l out of N
 * (gpucsl/pc/cuda/helpers/set_helpers.cu:25-49).  1 <= l <= GPUCSL_MAX_LEVEL;
 * levels 1..8 use the kernels' specialised unranking, higher ones the run-time one. */
int gpucsl_unrank(int N, int l, const uint64_t *ranks, int count, int32_t *out, void *stream);

/* ---------------------------------------------------------------------------
 * timing (the reference's only instrumentation is wall-clock @timed,
 * gpucsl/pc/helpers.py:30-39, and per-level log_time,
 * gpucsl/pc/discover_skeleton_gaussian.py:99-102)
 * ------------------------------------------------------------------------- */
/* While enabled, every kernel launch of this library is bracketed by a pair of
 * hipEvents recorded on the stream it is launched on.  Do not enable while a
 * stream is being captured into a graph. */
int gpucsl_profile_begin(void);
/* Disables timing, waits for the recorded events and writes one line per
 * kernel name: "name\tlaunches\ttotal_ms\tmin_ms\tmax_ms\n". */
int gpucsl_profile_end(char *out, size_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* GPUCSL_HIP_H */
