/*
 * jsvd_c.h -- C-ABI of the MI355X (gfx950) one-sided Jacobi SVD.
 *
 * Plain C types only (pointers, sizes, status codes); no HIP or torch types
 * appear in any signature.  The C++ API in include/jacobi/svd/ wraps these
 * entry points and rethrows the reference's exception types and messages.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the root of kleedaisuki/jacobi-svd-cuda-opt).
 *
 * Layout conventions (identical to the reference's JacobiSvdResult,
 * include/jacobi/svd/domain/jacobi_svd_result.cuh:26-36):
 *   A  m x n row-major (m >= n)        U  m x n row-major
 *   sigma  n values, in final column order (not sorted)
 *   V  n x n row-major
 * Packed batches store matrix k right after matrix k-1 with no padding:
 * A and U use m_k*n_k elements, sigma n_k, V n_k*n_k, sweeps one int32.
 *
 * Threading: every function is thread-safe.  Calls on the library's own
 * stream are serialised per device; plan executions run on the caller's
 * stream and may overlap with other plans.
 */
#ifndef JSVD_C_H
#define JSVD_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum jsvd_status
{
    JSVD_OK = 0,
    JSVD_EINVAL = 1,  /* invalid argument (reference: std::invalid_argument) */
    JSVD_EDEVICE = 2, /* HIP runtime failure (reference: jacobi::svd::CudaError) */
    JSVD_ENOMEM = 3   /* device allocation failed */
} jsvd_status;

/* Arithmetic of a solve. */
enum
{
    JSVD_PRECISION_FP64 = 0,  /* plain FP64 Jacobi in the reference's pair order */
    JSVD_PRECISION_MIXED = 1  /* FP32 sweeps + FP64 rotation parameters + FP64 refinement */
};

/* Why a solve stopped, per matrix (no reference counterpart: the reference
 * stops silently after max_sweeps, src/domain/jacobi_svd.cu:156-185). */
enum
{
    JSVD_STOP_CONVERGED = 0,  /* the last sweep rotated no pair: every pair passed the threshold test */
    JSVD_STOP_MAX_SWEEPS = 1, /* max_sweeps sweeps ran and the last one still rotated a pair */
    /* Matrices too large for LDS run kernels whose workgroups wait for each
     * other (software barriers over a grid sized for full residency).  A
     * wait longer than 10 s plus 1 s per GB of the launch's matrices (or
     * the environment variable JSVD_BARRIER_TIMEOUT_MS, read at each
     * execution) ends the launch instead of hanging: its matrices get this
     * code and their results are invalid.  The synchronous entry points then
     * return JSVD_EDEVICE. */
    JSVD_STOP_BARRIER_TIMEOUT = 2
};

/* Layout policy values (reference LayoutTransposeMode,
 * include/jacobi/svd/domain/layout_transpose.cuh:10-26). */
enum
{
    JSVD_LAYOUT_AUTO = 0,
    JSVD_LAYOUT_ON = 1,
    JSVD_LAYOUT_OFF = 2
};

/* Mirrors JacobiSvdConfig (include/jacobi/svd/domain/jacobi_svd_config.cuh:12-58)
 * field for field, then appends the knobs the reference lacks. */
typedef struct jsvd_config
{
    double epsilon;              /* default 1e-9: rotate pair when |pq| > eps*sqrt(pp*qq) */
    int32_t max_sweeps;          /* default 128 */
    int32_t threads_per_block;   /* default 256; rounded to a multiple of 64 in [64,1024] */
    int32_t layout_mode;         /* JSVD_LAYOUT_*; accepted and validated, no numeric effect */
    int32_t layout_min_columns;  /* default 16 */
    uint64_t layout_min_elements;/* default 4096 */
    int32_t layout_auto_tune;    /* default 0 */
    int32_t layout_bench_reps;   /* default 2 */
    int32_t layout_bench_sweeps; /* default 8 */
    /* --- no reference counterpart --- */
    int32_t precision;           /* JSVD_PRECISION_*; default FP64 */
    int32_t block_width;         /* 0 (default) = column pairs everywhere; 16 or 32 = matrices too large
                                    for LDS with n >= 2*block_width use blocked Jacobi (Gram blocks and
                                    block updates on FP64 MFMA; same convergence test and sweep count) */
    int32_t refine_sweeps;       /* mixed: FP64 sweep cap after the FP32 phase; 0 (default) = max_sweeps */
} jsvd_config;

/* Fills *cfg with the reference defaults (jacobi_svd_config.cuh:17-57). */
void jsvd_config_default(jsvd_config *cfg);

/* Message of the last failing call on this thread ("" if none).  Same text
 * the C++ API throws. */
const char *jsvd_last_error(void);

/* Number of the library's ABI revision (2: jsvd_plan_info gained
 * ring_matrices; a caller built against revision 1 passes a shorter struct). */
int32_t jsvd_abi_version(void);

/* Single solve from host memory, synchronous.
 * Replaces jacobi::svd::one_sided_jacobi_svd
 * (include/jacobi/svd/domain/jacobi_svd.cuh:25-28, src/domain/jacobi_svd.cu:402-412). */
jsvd_status jsvd_solve_f64(const double *a_rm, uint64_t m, uint64_t n, const jsvd_config *cfg, double *u_rm,
                           double *sigma, double *v_rm, int32_t *sweeps);

/* Batched solve from host memory: count independent matrices, pointer per
 * matrix.  Synchronous.  Replaces the per-testcase loop of
 * src/application/pipeline.cu:59-83 -> src/application/kernel_stage.cuh:30-58. */
jsvd_status jsvd_solve_batched_f64(int64_t count, const double *const *a_rm, const uint64_t *m, const uint64_t *n,
                                   const jsvd_config *cfg, double *const *u_rm, double *const *sigma,
                                   double *const *v_rm, int32_t *sweeps);

/* jsvd_solve_batched_f64 plus the stop reason of every matrix
 * (JSVD_STOP_*, one int32 per matrix; NULL = not wanted). */
jsvd_status jsvd_solve_batched_status_f64(int64_t count, const double *const *a_rm, const uint64_t *m,
                                          const uint64_t *n, const jsvd_config *cfg, double *const *u_rm,
                                          double *const *sigma, double *const *v_rm, int32_t *sweeps,
                                          int32_t *stop);

/* ---- device-resident batches (no reference counterpart: the reference has no
 * batched or asynchronous entry point; this is the work-list launch the
 * batched pipeline above is built on) ---- */
typedef struct jsvd_plan jsvd_plan;

typedef struct jsvd_plan_info
{
    int64_t count;             /* matrices in the plan */
    int64_t lds_matrices;      /* solved by the LDS-resident work-list kernel */
    int64_t large_matrices;    /* too large for LDS: wave-per-pair kernels (grouped barriers or one
                                  workgroup per matrix), the grid-barrier kernel, or blocked Jacobi */
    int64_t a_elements;        /* packed A / U elements */
    int64_t v_elements;        /* packed V elements */
    int64_t sigma_elements;    /* packed sigma elements */
    int32_t workgroup_size;    /* threads per workgroup of the first (largest) bucket */
    int32_t grid_size;         /* workgroups of that bucket's launch */
    int32_t lds_bytes;         /* dynamic LDS per workgroup of that bucket */
    int32_t precision;
    int32_t buckets;           /* work-list launches (shape class x LDS occupancy) */
    int32_t blocked_matrices;  /* large matrices on the blocked path */
    int32_t ring_matrices;     /* large matrices solved by the ring form of the wave kernels */
    int32_t pad;
} jsvd_plan_info;

/* Packed layout of a batch (no reference counterpart: the reference indexes
 * one matrix with 32-bit int arithmetic, src/domain/matrix_index.cuh:12-27,
 * which overflows at 2^31 elements).  Host only; needs no device.  Checks
 * every shape as jsvd_plan_create does, then writes the 64-bit element
 * offsets of matrix k in the packed A / U, V and sigma buffers (each output
 * array may be NULL) and the totals {A / U, V, sigma} to totals[0..2] (may be
 * NULL).  JSVD_EINVAL when a buffer would exceed 2^60 elements. */
jsvd_status jsvd_packed_layout(int64_t count, const int64_t *m, const int64_t *n, int64_t *a_off, int64_t *v_off,
                               int64_t *s_off, int64_t *totals);

/* Device workspace of one matrix on the blocked path (no reference
 * counterpart).  Host only.  For block_width 16 or 32 and an m x n matrix
 * with n >= 2 * block_width: rows padded to lda (a multiple of 64), columns
 * to nv (an even number of blocks), the block pairs per round, the Gram row
 * chunks, and the element counts of the column-major A and V workspaces and
 * of the partial-Gram buffer.  JSVD_EINVAL for other widths or shapes. */
typedef struct jsvd_blocked_layout
{
    int64_t lda, nv, blocks, pairs, chunk_rows, chunks;
    int64_t a_elements, v_elements, gram_elements;
} jsvd_blocked_layout;
jsvd_status jsvd_blocked_layout_f64(int64_t m, int64_t n, int32_t block_width, jsvd_blocked_layout *out);

/* Builds the device work-list for count matrices of shapes m[k] x n[k].
 * Allocates all device memory the plan needs; execute allocates nothing. */
jsvd_status jsvd_plan_create(int64_t count, const int64_t *m, const int64_t *n, const jsvd_config *cfg,
                             jsvd_plan **out);

/* Runs the plan on device buffers (packed layout above) on `stream`
 * (a hipStream_t, or NULL for the null stream).  Asynchronous.  d_sweeps may
 * be NULL.  For JSVD_PRECISION_MIXED the inputs and outputs are still FP64.
 * The convergence decisions of every path stay on the device (the blocked
 * path's later sweeps return at once after convergence), so an execution
 * never waits for the device and can be captured in a HIP graph. */
jsvd_status jsvd_plan_execute_f64(jsvd_plan *plan, const double *d_a, double *d_u, double *d_sigma, double *d_v,
                                  int32_t *d_sweeps, void *stream);

/* jsvd_plan_execute_f64 plus the stop reason of every matrix in device
 * memory (JSVD_STOP_*, one int32 per matrix, batch order; may be NULL). */
jsvd_status jsvd_plan_execute_status_f64(jsvd_plan *plan, const double *d_a, double *d_u, double *d_sigma,
                                         double *d_v, int32_t *d_sweeps, int32_t *d_stop, void *stream);

/* Error metrics of a plan's results (no reference counterpart: the reference
 * reports sweeps only, src/interfaces/cli_support.cu:427-464).  From the
 * device buffers of a jsvd_plan_execute_f64 call (d_a its input, d_u /
 * d_sigma / d_v its outputs), writes three doubles per matrix to d_out:
 *   ||U^T U - D||_F with D = diag(sigma_j > epsilon) (the U columns the U
 *   rule keeps), ||V^T V - I||_F, and ||A - U diag(sigma) V^T||_F / ||A||_F.
 * Asynchronous on `stream`; the first call on a plan allocates and uploads
 * its tile list (synchronously), later calls allocate nothing. */
jsvd_status jsvd_plan_metrics_f64(jsvd_plan *plan, const double *d_a, const double *d_u, const double *d_sigma,
                                  const double *d_v, double *d_out, void *stream);

jsvd_status jsvd_plan_get_info(const jsvd_plan *plan, jsvd_plan_info *info);
void jsvd_plan_destroy(jsvd_plan *plan);

/* No counterpart in the reference.  For callers that cannot allocate (code
 * under stream capture, serving loops, harnesses that forbid hipMalloc): the
 * solve of jsvd_plan_create + jsvd_plan_execute_f64 in one call that
 * allocates nothing, on the device or the host.  Same packed layout, same
 * per-matrix kernels and results (bitwise) as a plan of the same batch;
 * asynchronous on `stream`, d_sweeps may be NULL.
 * Limits: at most 32768 matrices, FP64 only (cfg->precision must be
 * JSVD_PRECISION_FP64), and every matrix must fit the LDS-resident work-list
 * kernel with A and V in LDS (every m <= 160 with n <= 128 does, and every
 * m <= 256 with n <= 64); anything else returns JSVD_EINVAL with a message.  The work-lists live in
 * a static 2.5 MiB device array per GPU, which keeps the last call's
 * contents: calls that overlap on one GPU (two streams, two threads) race on
 * it and must be serialised by the caller. */
jsvd_status jsvd_solve_device_noalloc_f64(int64_t count, const int64_t *m, const int64_t *n,
                                          const jsvd_config *cfg, const double *d_a, double *d_u, double *d_sigma,
                                          double *d_v, int32_t *d_sweeps, void *stream);

/* SVD adjoint (no reference counterpart: the reference has no derivative).
 * For each matrix of the plan, from the solve's U, sigma, V (any column
 * order, as returned) and upstream gradients gU, gsigma, gV (each may be
 * NULL = zero), writes
 *   gA = U [ (F o (U^T gU - gU^T U)) S + S (F o (V^T gV - gV^T V)) + diag(gsigma) ] V^T
 *        + (I - U U^T) gU S^-1 V^T,   F_ij = 1/(s_j^2 - s_i^2) (i != j),
 * packed like A.  Requires distinct nonzero singular values, m < 2^31, and
 * (n'^2 + max(n', 64)(n'+1)) doubles of LDS per matrix, n' being n rounded up
 * to 16 (n <= 96, any m), in FP64 and FP32 alike.
 * Asynchronous on `stream`; allocates nothing. */
jsvd_status jsvd_plan_backward_f64(jsvd_plan *plan, const double *d_u, const double *d_sigma, const double *d_v,
                                   const double *d_gu, const double *d_gsigma, const double *d_gv, double *d_ga,
                                   void *stream);
jsvd_status jsvd_plan_backward_f32(jsvd_plan *plan, const float *d_u, const float *d_sigma, const float *d_v,
                                   const float *d_gu, const float *d_gsigma, const float *d_gv, float *d_ga,
                                   void *stream);

/* Layout-threshold autotune.  Replaces jacobi::svd::auto_tune_layout_transpose_threshold
 * (include/jacobi/svd/domain/jacobi_svd.cuh:36-37, src/domain/layout_transpose_autotune.cu:78-126). */
typedef struct jsvd_autotune_report
{
    int32_t executed;
    int32_t recommended_min_columns;
    uint64_t recommended_min_elements;
    double estimated_best_speedup;
    uint64_t sample_count;
} jsvd_autotune_report;

jsvd_status jsvd_autotune(const jsvd_config *cfg, jsvd_autotune_report *out);

/* Shape-class table of the work-list kernel (no reference counterpart).
 * Column counts are bucketed as n <= 8, 16, 32, 48, 64, 96, 128 (buckets
 * 0..6).  jsvd_tune_class sets the class used for a bucket by plans created
 * afterwards (-1 = default choice; an index that does not fit a shape falls
 * back to the default) and returns the previous setting, or -2 for a bad
 * bucket.  jsvd_autotune fills this table from measurements. */
int32_t jsvd_tune_class(int32_t bucket, int32_t cls);
int32_t jsvd_class_count(void);

/* Shape-class autotune on a caller's own batch (no reference counterpart:
 * the reference tunes only its layout threshold).  Times whole executions of
 * a plan for the count matrices m[k] x n[k] on device buffers of the packed
 * layout (d_a is read, d_u / d_sigma / d_v are overwritten), one warm-up
 * and `reps` timed runs per candidate, on `stream`.  Buckets present in the
 * batch are visited smallest first; each keeps its current class unless the
 * default (-1) or a class that fits all of its matrices makes the whole plan
 * at least 1% faster.  The winners stay installed in the table above, so
 * plans created afterwards use them; the report gives the table and the
 * plan's time before and after.  Blocks the host until the timing is done. */
#define JSVD_TUNE_BUCKETS 7
typedef struct jsvd_batch_tune_report
{
    int32_t classes[JSVD_TUNE_BUCKETS]; /* class per bucket after tuning (-1 = default) */
    int32_t samples;                    /* plans timed */
    double ms_initial;                  /* mean ms per execution with the table as it was */
    double ms_tuned;                    /* mean ms per execution with the tuned table */
} jsvd_batch_tune_report;

jsvd_status jsvd_autotune_batch(int64_t count, const int64_t *m, const int64_t *n, const jsvd_config *cfg,
                                const double *d_a, double *d_u, double *d_sigma, double *d_v, int32_t reps,
                                void *stream, jsvd_batch_tune_report *out);

#ifdef __cplusplus
}
#endif

#endif /* JSVD_C_H */
