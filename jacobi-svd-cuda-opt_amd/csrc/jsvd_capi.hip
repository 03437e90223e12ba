// jsvd_capi.hip -- the C-ABI (include/jsvd_c.h): validation, device work-list
// plans, workspace, launches and error reporting.  The solver kernels live in
// jsvd_kernels_*.hip; this file holds only the two small work-list kernels of
// the no-allocation entry point, beside its static scratch array.
#include "jsvd_c.h"
#include "jsvd_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

namespace
{
    thread_local std::string g_last_error;

    jsvd_status fail(jsvd_status status, const std::string &message)
    {
        g_last_error = message;
        return status;
    }

    std::string hip_message(hipError_t e, const char *expr, const char *file, int line)
    {
        char buf[512];
        std::snprintf(buf, sizeof(buf), "HIP call failed: %s @ %s:%d, code=%d, message=%s", expr, file, line,
                      static_cast<int>(e), hipGetErrorString(e));
        return buf;
    }

    // The runtime keeps the last failing call's code until hipGetLastError
    // reads it, and the launch wrappers report hipGetLastError after each
    // launch: an earlier failure (the caller's, or an allocation this library
    // already reported) would be taken for a launch error.  Entry points that
    // launch therefore start by discarding it, and a reported failure is
    // consumed as it is reported.
    void discard_stale_error()
    {
        (void)hipGetLastError();
    }

#define JSVD_HIP(expr)                                                                                         \
    do                                                                                                         \
    {                                                                                                          \
        const hipError_t jsvd_e_ = (expr);                                                                     \
        if (jsvd_e_ != hipSuccess)                                                                             \
        {                                                                                                      \
            discard_stale_error();                                                                             \
            return fail(jsvd_e_ == hipErrorOutOfMemory ? JSVD_ENOMEM : JSVD_EDEVICE,                           \
                        hip_message(jsvd_e_, #expr, __FILE__, __LINE__));                                      \
        }                                                                                                      \
    } while (0)

#define JSVD_CHECK(expr)                                                                                       \
    do                                                                                                         \
    {                                                                                                          \
        const jsvd_status jsvd_s_ = (expr);                                                                    \
        if (jsvd_s_ != JSVD_OK)                                                                                \
            return jsvd_s_;                                                                                    \
    } while (0)

    // Validation with the reference's messages
    // (src/domain/jacobi_svd.cu:197-223 and :36-54 of that file).  The
    // epsilon test is the reference's `epsilon <= 0.0` (jacobi_svd.cu:211), so
    // a NaN epsilon is accepted as there: every threshold comparison with it
    // is false, no pair rotates, the solve stops after one sweep with sigma =
    // the column norms, U = 0 and V = I.
    jsvd_status validate_config(const jsvd_config &c)
    {
        if (c.epsilon <= 0.0)
            return fail(JSVD_EINVAL, "epsilon must be positive.");
        if (c.max_sweeps <= 0)
            return fail(JSVD_EINVAL, "max_sweeps must be positive.");
        if (c.layout_min_columns <= 0)
            return fail(JSVD_EINVAL, "layout_transpose_min_columns must be positive.");
        if (c.layout_min_elements == 0)
            return fail(JSVD_EINVAL, "layout_transpose_min_elements must be positive.");
        if (c.layout_bench_reps <= 0)
            return fail(JSVD_EINVAL, "layout_transpose_benchmark_repetitions must be positive.");
        if (c.layout_bench_sweeps <= 0)
            return fail(JSVD_EINVAL, "layout_transpose_benchmark_sweeps must be positive.");
        if (c.layout_mode < JSVD_LAYOUT_AUTO || c.layout_mode > JSVD_LAYOUT_OFF)
            return fail(JSVD_EINVAL, "layout_transpose_mode is not a valid policy.");
        if (c.precision != JSVD_PRECISION_FP64 && c.precision != JSVD_PRECISION_MIXED)
            return fail(JSVD_EINVAL, "precision must be JSVD_PRECISION_FP64 or JSVD_PRECISION_MIXED.");
        if (c.block_width != 0 && c.block_width != 16 && c.block_width != 32)
            return fail(JSVD_EINVAL, "block_width must be 0, 16 or 32.");
        if (c.refine_sweeps < 0)
            return fail(JSVD_EINVAL, "refine_sweeps must be non-negative.");
        return JSVD_OK;
    }

    jsvd_status validate_shape(uint64_t rows, uint64_t columns)
    {
        if (rows == 0 || columns == 0)
            return fail(JSVD_EINVAL, "Input matrix shape must be non-zero.");
        if (rows < columns)
            return fail(JSVD_EINVAL, "One-sided Jacobi SVD requires rows >= columns in this implementation.");
        if (columns > (1u << 30) || rows > (1ull << 40))
            return fail(JSVD_EINVAL, "Input matrix shape is too large.");
        return JSVD_OK;
    }

    const char *const kBarrierTimeoutMessage =
        "A device barrier timed out: the workgroups of a barrier kernel were not all resident (other kernels "
        "holding CUs?); the results of the matrices reported with stop reason JSVD_STOP_BARRIER_TIMEOUT are invalid.";

    // Largest packed buffer, in elements (2^63 bytes).
    constexpr int64_t kMaxPackedElements = int64_t(1) << 60;

    // Offsets of the packed layout (every shape checked first); any output
    // may be null.  All arithmetic is 64-bit, with m * n compared by
    // division so that no product can overflow.
    jsvd_status packed_layout(int64_t count, const int64_t *m, const int64_t *n, int64_t *a_off, int64_t *v_off,
                              int64_t *s_off, int64_t *totals)
    {
        if (count < 0 || (count > 0 && (!m || !n)))
            return fail(JSVD_EINVAL, "Input batch description is invalid.");
        int64_t ao = 0, vo = 0, so = 0;
        for (int64_t k = 0; k < count; ++k)
        {
            if (m[k] <= 0 || n[k] <= 0)
                return fail(JSVD_EINVAL, "Input matrix shape must be non-zero.");
            JSVD_CHECK(validate_shape(static_cast<uint64_t>(m[k]), static_cast<uint64_t>(n[k])));
            if (m[k] > kMaxPackedElements / n[k])
                return fail(JSVD_EINVAL, "Packed batch exceeds 2^60 elements.");
            const int64_t mn = m[k] * n[k], nn = n[k] * n[k]; // n <= 2^30: nn <= 2^60
            if (ao > kMaxPackedElements - mn || vo > kMaxPackedElements - nn || so > kMaxPackedElements - n[k])
                return fail(JSVD_EINVAL, "Packed batch exceeds 2^60 elements.");
            if (a_off)
                a_off[k] = ao;
            if (v_off)
                v_off[k] = vo;
            if (s_off)
                s_off[k] = so;
            ao += mn;
            vo += nn;
            so += n[k];
        }
        if (totals)
        {
            totals[0] = ao;
            totals[1] = vo;
            totals[2] = so;
        }
        return JSVD_OK;
    }

    // threads_per_block maps to a workgroup size: a multiple of 64 in [64,1024]
    // (the reference rounds to a multiple of 32 in [32,1024],
    // src/domain/jacobi_schedule.cuh:64-69).
    int normalize_workgroup(int raw)
    {
        const int clamped = std::clamp(raw, 64, 1024);
        return std::min(((clamped + 63) / 64) * 64, 1024);
    }

    int lds_kernel_block(int raw)
    {
        const int w = normalize_workgroup(raw);
        if (w <= 64)
            return 64;
        if (w <= 128)
            return 128;
        if (w <= 256)
            return 256;
        return 512;
    }

    // Blocked path: matrices too large for LDS with at least two column
    // blocks when the caller asked for block_width 16 or 32.
    bool uses_blocked(const jsvd_config &cfg, int64_t n)
    {
        return cfg.block_width != 0 && n >= 2 * int64_t(cfg.block_width);
    }

    // Column blocking and row chunking of one matrix on the blocked path:
    // zero columns pad n to nv = nb * block with nb even (no idle block in
    // any round), zero rows pad m to lda (a multiple of 64); Gram row chunks
    // (multiples of 64 rows) are sized for about 512 workgroups per launch.
    struct BlockedShape
    {
        int block = 0, nb = 0, nv = 0, pairs = 0, chunks = 0;
        int64_t lda = 0, chunk_rows = 0;
    };

    BlockedShape blocked_shape(int block, int64_t m, int64_t n)
    {
        BlockedShape s;
        s.block = block;
        s.nb = static_cast<int>((n + 2 * block - 1) / (2 * block)) * 2;
        s.nv = s.nb * block;
        s.pairs = s.nb / 2;
        s.lda = (m + 63) / 64 * 64;
        const int64_t panels = s.lda / 64;
        // Row chunks per block pair for the partial Grams: about 512 Gram
        // workgroups at block width 16 and 256 at 32 (more chunks keep more
        // workgroups streaming, but the Gram solve sums every chunk; measured
        // on 8192 x 2048: 11.7 / 12.5 ms per sweep at 512 / 256 for width 16,
        // 9.5 / 9.3 for width 32).
        const int64_t target = block == 32 ? 256 : 512;
        const int64_t chunks = std::clamp<int64_t>((target + s.pairs - 1) / s.pairs, 1, panels);
        s.chunk_rows = ((panels + chunks - 1) / chunks) * 64;
        s.chunks = static_cast<int>((s.lda + s.chunk_rows - 1) / s.chunk_rows);
        return s;
    }

    int persistent_kernel_block(int raw)
    {
        const int w = normalize_workgroup(raw);
        if (w <= 256)
            return 256;
        if (w <= 512)
            return 512;
        return 1024;
    }

    template <typename T>
    struct DevBuf
    {
        T *ptr = nullptr;
        size_t count = 0;
        DevBuf() = default;
        DevBuf(const DevBuf &) = delete;
        DevBuf &operator=(const DevBuf &) = delete;
        ~DevBuf()
        {
            if (ptr)
                (void)hipFree(ptr);
        }
        jsvd_status alloc(size_t n)
        {
            if (ptr)
                (void)hipFree(ptr);
            ptr = nullptr;
            count = 0;
            if (n == 0)
                return JSVD_OK;
            JSVD_HIP(hipMalloc(reinterpret_cast<void **>(&ptr), n * sizeof(T)));
            count = n;
            return JSVD_OK;
        }
        jsvd_status reserve(size_t n)
        {
            return n <= count ? JSVD_OK : alloc(n);
        }
    };
} // namespace

namespace
{
    // Process-wide shape-class choices per column-count bucket, written by
    // the autotuner (or jsvd_tune_set) and read at plan creation.  -1 keeps
    // the default (first fitting class).
    constexpr int kTuneBuckets = 7;
    constexpr int64_t kTuneUpper[kTuneBuckets] = {8, 16, 32, 48, 64, 96, 128};
    std::mutex g_tune_mu;
    int g_tuned[kTuneBuckets] = {-1, -1, -1, -1, -1, -1, -1};
    static_assert(kTuneBuckets == JSVD_TUNE_BUCKETS, "bucket table size is part of the C ABI");

    int tune_bucket(int64_t n)
    {
        for (int b = 0; b < kTuneBuckets; ++b)
            if (n <= kTuneUpper[b])
                return b;
        return -1;
    }

    // Class pinned by the autotuner for this column count, or -1.
    int tuned_class(int64_t m, int64_t n)
    {
        const int b = tune_bucket(n);
        if (b < 0)
            return -1;
        std::lock_guard<std::mutex> lock(g_tune_mu);
        const int c = g_tuned[b];
        return (c >= 0 && jsvd::class_fits(c, m, n)) ? c : -1;
    }

    // join_four_wave: the plan already has four-wave (class 4) matrices;
    // the one-wave classes' matrices (n <= 16) then join that work-list,
    // where the longest-first order runs them at its end instead of in a
    // bucket that waits for registers behind it (1% on the uneven batch).
    int preferred_class(int64_t m, int64_t n, bool join_four_wave)
    {
        const int t = tuned_class(m, n);
        if (t >= 0)
            return t;
        const int c = jsvd::select_class_f64(m, n);
        if (join_four_wave && (c == 0 || c == 1) && jsvd::class_fits(4, m, n))
            return 4;
        return c;
    }
} // namespace

// Work-list counters per bucket: two per stage (FP32 sweeps,
// orthonormalisation, B = A Q, FP64 sweeps, then the QR preconditioner's two
// kernels); plain buckets use the FP64 pair.
constexpr int kCountersPerBucket = 12;
// Mixed-precision buckets of matrices wider than kMixedSplitMinColumns are
// split into up to kMixedSplitParts parts of at least kMixedSplitMin
// matrices, each on a stream of its own.
constexpr size_t kMixedSplitMin = 64;
constexpr size_t kMixedSplitParts = 2;
constexpr int64_t kMixedSplitMinColumns = 64;

// A bucket is one persistent launch of the LDS-resident work-list kernel:
// matrices of one shape class whose LDS needs allow the same number of
// workgroups per CU.
struct JsvdBucket
{
    int cls = -1;           // shape class, -1 = generic kernel
    int block = 256;
    int lds_bytes = 0;
    int grid = 0;
    int64_t first = 0;      // first work-list entry of the bucket
    int64_t count = 0;
    double work = 0.0;      // sum of n^2 (m + n), for launch order
    int64_t max_n = 0, max_m = 0;
    int stream = 0;         // 0 = the caller's stream, k = auxiliary stream k-1
    int per_cu = 0;         // resident workgroups per CU
    // Rotation-log classes: one tangent log per resident workgroup.
    double *vlog = nullptr;
    int64_t vlog_stride = 0;
    // Mixed precision: the FP32 sweep class and the launch shapes of the
    // FP32, orthonormalisation and B = A Q stages; `cls` is then the class
    // of the FP64 refinement.
    int f32cls = -1;
    int lds32 = 0, grid32 = 0;
    int lds_ortho = 0, grid_ortho = 0, grid_gemm = 0;
    // Mixed precision on X = R2^T of two QRs, A = Q1 R and R^T = Q2 R2 (the
    // stages above and the FP64 refinement then run on n x n matrices): the
    // launch shapes of both QRs and of the back-transform.
    bool precond = false;
    int qr_variant = -1, lds_qr = 0, grid_qr = 0, lds_fin = 0, grid_fin = 0, grid_polar = 0;
    int qr2_variant = -1, lds_qr2 = 0, grid_qr2 = 0;
};

struct jsvd_plan
{
    int device = 0;
    jsvd_config cfg{};
    int64_t count = 0;
    std::vector<int64_t> m, n, a_off, v_off, s_off;
    // LDS-resident work-list part: entries grouped by bucket, largest first.
    std::vector<jsvd::WorkItem> items;
    std::vector<JsvdBucket> buckets;
    DevBuf<jsvd::WorkItem> d_items;
    DevBuf<unsigned int> d_counters; // kCountersPerBucket per bucket
    std::vector<std::unique_ptr<DevBuf<double>>> vlogs;
    DevBuf<double> q_scratch;        // mixed precision: packed column-major Q
    // QR-preconditioned mixed buckets: the square work-list entries of X
    // (m = n, A/U offsets = V offsets), in the order of `items`, and the
    // scratch of the preconditioner.
    std::vector<jsvd::WorkItem> pre_items;
    DevBuf<jsvd::WorkItem> d_pre_items;
    DevBuf<double> pc_refl, pc_tau, pc_x, pc_b, pc_ux, pc_vx, pc_refl2, pc_tau2, pc_t1, pc_t2;
    DevBuf<int32_t> pc_perm;
    // Every matrix in batch order (the adjoint kernel's work-list) and the
    // largest LDS need of the adjoint in FP64 / FP32.
    DevBuf<jsvd::WorkItem> d_all;
    int64_t bw_lds_f64 = 0, bw_lds_f32 = 0;
    int64_t max_rows = 0; // largest m of the plan
    // Buckets beyond the first run on plan-owned streams forked from the
    // caller's stream and joined back (event record / wait only).
    std::vector<hipStream_t> aux;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    // Large matrices: persistent kernel over a column-major workspace.
    std::vector<int64_t> large;
    DevBuf<double> ws_a, ws_v, ws_s;
    DevBuf<int32_t> ws_flags;
    DevBuf<unsigned int> ws_bar;
    DevBuf<unsigned int> d_fault; // set by a barrier kernel whose wait hit its spin bound
    double wave_bytes = 0.0;      // bytes of the grouped-wave matrices (spin bound)
    int big_block = 256;
    int big_resident = 0; // resident workgroups of the grid-barrier kernel (occupancy x CUs)
    hipEvent_t big_in = nullptr, big_out = nullptr;
    // Large matrices with m <= wave_sweeps_max_rows() (and no blocked
    // path): the grouped wave kernel, one column-major workspace per matrix.
    // Entries are ordered by row-register count (one launch each), longest
    // first inside a launch.
    struct WaveLaunch
    {
        int64_t m_max = 0, n_max = 0;
        int first = 0, count = 0;
        int members = 1; // workgroups per group; 0 = one workgroup per matrix
        int groups = 1;
        int waves = 8;   // waves per workgroup of the grouped form
        bool ring = false;      // ring form (wave_ring_sweeps_f64)
        int64_t mbox_group = 0; // ring form: mailbox words per group
        int xcd_spread = 0;     // ring form: groups placed one XCD each (0: group-major grid)
        bool ring_db = false;   // ring form: two LDS slot sets (one barrier per round)
    };
    std::vector<int64_t> wave;                 // batch index per entry
    std::vector<int64_t> wave_a_off, wave_v_off; // workspace offsets per entry (barrier forms)
    std::vector<char> wave_ring;                 // entry solved by the ring form (no workspace)
    std::vector<WaveLaunch> wave_launches;
    DevBuf<jsvd::WaveMatrix> d_wave;
    DevBuf<double> wv_a, wv_v;
    DevBuf<int32_t> wv_flags;
    DevBuf<unsigned int> wv_bars;
    DevBuf<unsigned long long> wv_mbox; // ring-form column mailboxes
    DevBuf<unsigned int> wv_xcc;        // ring form: XCC id of every member
    // Blocked Jacobi (cfg.block_width 16 or 32): per-round partial Grams,
    // rotations Q and flags, sized for the largest blocked matrix; the sweep
    // flag is read back through pinned host memory once per sweep.
    int64_t blocked = 0;
    DevBuf<double> bk_gram, bk_q;
    DevBuf<int32_t> bk_flags;
    int32_t *h_flag = nullptr;            // two pinned copies of the blocked solve's done word
    hipEvent_t bk_seen[2] = {nullptr, nullptr};
    int64_t a_elems = 0, v_elems = 0, s_elems = 0;
    // Error metrics (jsvd_plan_metrics_f64): tile list, per-matrix tile
    // ranges and partial sums, built on the first metrics call.
    DevBuf<jsvd::MetricTile> mt_tiles;
    DevBuf<int64_t> mt_first;
    DevBuf<double> mt_partial;
    int64_t mt_count = -1;

    ~jsvd_plan()
    {
        for (hipStream_t st : aux)
            (void)hipStreamDestroy(st);
        for (hipEvent_t e : join)
            (void)hipEventDestroy(e);
        if (fork)
            (void)hipEventDestroy(fork);
        if (big_in)
            (void)hipEventDestroy(big_in);
        if (big_out)
            (void)hipEventDestroy(big_out);
        if (h_flag)
            (void)hipHostFree(h_flag);
        for (hipEvent_t e : bk_seen)
            if (e)
                (void)hipEventDestroy(e);
    }
};

namespace
{
    // One stream per device carries every grid-barrier launch, so at most one
    // such kernel is on the GPU at a time (see launch_persistent_sweeps_f64).
    std::mutex g_barrier_mu;
    std::vector<hipStream_t> g_barrier_streams;

    // Wall-clock ticks a software barrier may wait before its launch is
    // marked faulted: 10 s plus 1 s per GB of the launch's matrices (a round
    // streams the largest of them about three times), or
    // JSVD_BARRIER_TIMEOUT_MS when that is set (0 makes every wait fail at
    // once, which is how the failure path is tested).
    uint64_t barrier_spin_limit(int device, double bytes)
    {
        static std::mutex mu;
        static std::vector<int> khz;
        int rate = 0;
        {
            std::lock_guard<std::mutex> lock(mu);
            if (static_cast<int>(khz.size()) <= device)
                khz.resize(device + 1, 0);
            if (khz[device] <= 0)
            {
                int r = 0;
                if (hipDeviceGetAttribute(&r, hipDeviceAttributeWallClockRate, device) != hipSuccess || r <= 0)
                {
                    (void)hipGetLastError();
                    r = 100000; // 100 MHz, the usual rate
                }
                khz[device] = r;
            }
            rate = khz[device];
        }
        double seconds = 10.0 + bytes / 1e9;
        if (const char *env = std::getenv("JSVD_BARRIER_TIMEOUT_MS"))
            seconds = std::max(0.0, std::atof(env)) / 1000.0;
        return static_cast<uint64_t>(seconds * rate * 1000.0);
    }

    jsvd_status barrier_stream(int device, hipStream_t *out)
    {
        std::lock_guard<std::mutex> lock(g_barrier_mu);
        if (static_cast<int>(g_barrier_streams.size()) <= device)
            g_barrier_streams.resize(device + 1, nullptr);
        if (!g_barrier_streams[device])
            JSVD_HIP(hipStreamCreateWithFlags(&g_barrier_streams[device], hipStreamNonBlocking));
        *out = g_barrier_streams[device];
        return JSVD_OK;
    }

    // Writes sigma, U (row-major m x n, columns divided by sigma) and V
    // (row-major n x n) from the column-major workspaces, whose leading
    // dimensions are lda >= m and ldv >= n.
    jsvd_status finish_large(jsvd_plan &p, int64_t k, int64_t m, int n, int64_t lda, int64_t ldv, double *d_u,
                             double *d_sigma, double *d_v, hipStream_t stream)
    {
        JSVD_HIP(jsvd::launch_column_norms_f64(p.ws_a.ptr, lda, p.ws_s.ptr, m, n, stream));
        JSVD_HIP(hipMemcpyAsync(d_sigma + p.s_off[k], p.ws_s.ptr, sizeof(double) * n, hipMemcpyDeviceToDevice,
                                stream));
        // U (m x n row-major) = transpose of the n x m row-major view of A_cm,
        // each A column divided by its sigma.
        JSVD_HIP(jsvd::launch_transpose_ld_f64(p.ws_a.ptr, lda, d_u + p.a_off[k], n, n, m, p.ws_s.ptr,
                                               p.cfg.epsilon, stream));
        JSVD_HIP(jsvd::launch_transpose_ld_f64(p.ws_v.ptr, ldv, d_v + p.v_off[k], n, n, n, nullptr, 0.0, stream));
        return JSVD_OK;
    }

    // Blocked Jacobi: sweeps of block-pair rounds until a sweep applies no
    // rotation (one 4-byte read-back per sweep, as the reference's sweep
    // loop does, src/domain/jacobi_svd.cu:156-185).
    jsvd_status plan_blocked(jsvd_plan &p, int64_t k, const double *d_a, double *d_u, double *d_sigma, double *d_v,
                             int32_t *d_sweeps, int32_t *d_stop, hipStream_t stream)
    {
        const int64_t m = p.m[k];
        const int n = static_cast<int>(p.n[k]);
        const BlockedShape sh = blocked_shape(p.cfg.block_width, m, n);
        // Zero padding rows and columns (a fill kernel, so the sequence can
        // be captured), then A (row-major) into the column-major workspace
        // with leading dimension lda.
        JSVD_HIP(jsvd::launch_fill_i32(reinterpret_cast<int32_t *>(p.ws_a.ptr), 0, 2 * sh.lda * sh.nv, stream));
        JSVD_HIP(jsvd::launch_transpose_ld_f64(d_a + p.a_off[k], n, p.ws_a.ptr, sh.lda, m, n, nullptr, 0.0, stream));
        JSVD_HIP(jsvd::launch_identity_f64(p.ws_v.ptr, sh.nv, stream));
        jsvd::BlockedArgs args{};
        args.block = sh.block;
        args.nb = sh.nb;
        args.lda = sh.lda;
        args.nv = sh.nv;
        args.a = p.ws_a.ptr;
        args.v = p.ws_v.ptr;
        args.chunk_rows = sh.chunk_rows;
        args.chunks = sh.chunks;
        args.eps = p.cfg.epsilon;
        args.gram = p.bk_gram.ptr;
        args.q = p.bk_q.ptr;
        args.pair_rotated = p.bk_flags.ptr;
        args.state = p.bk_flags.ptr + sh.pairs;
        const int rounds = sh.nb - 1;
        // The convergence decision stays on the device: every sweep starts
        // with a one-thread step kernel that closes the previous sweep, and
        // once a sweep has rotated nothing every later launch returns at
        // once.  The host never waits.  Outside stream capture it stops
        // enqueueing once it sees, without blocking, that an earlier
        // sweep's done word has reached host memory set; under capture the
        // whole max_sweeps sequence is recorded.
        hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
        JSVD_HIP(hipStreamIsCapturing(stream, &capture));
        const bool watch = capture == hipStreamCaptureStatusNone;
        JSVD_HIP(jsvd::launch_fill_i32(args.state, 0, 4, stream));
        if (watch)
            p.h_flag[0] = p.h_flag[1] = 0;
        int enqueued = 0, probes = 0;
        for (int sweep = 0; sweep < p.cfg.max_sweeps; ++sweep)
        {
            JSVD_HIP(jsvd::launch_blocked_sweep_step(args.state, sweep, stream));
            for (int r = 0; r < rounds; ++r)
                JSVD_HIP(jsvd::launch_blocked_round(args, r, stream));
            enqueued = sweep + 1;
            if (watch)
            {
                const int slot = probes & 1;
                if (probes > 0)
                {
                    const int prev = slot ^ 1;
                    const hipError_t q = hipEventQuery(p.bk_seen[prev]);
                    if (q == hipSuccess && p.h_flag[prev] != 0)
                        break;
                    if (q != hipSuccess && q != hipErrorNotReady)
                        JSVD_HIP(q);
                    if (q == hipErrorNotReady)
                        continue; // the previous probe is still in flight: keep its slot
                }
                JSVD_HIP(hipMemcpyAsync(p.h_flag + slot, args.state + 1, sizeof(int32_t), hipMemcpyDeviceToHost,
                                        stream));
                JSVD_HIP(hipEventRecord(p.bk_seen[slot], stream));
                ++probes;
            }
        }
        JSVD_HIP(jsvd::launch_blocked_sweep_close(args.state, enqueued, d_sweeps ? d_sweeps + k : nullptr,
                                                  d_stop ? d_stop + k : nullptr, stream));
        return finish_large(p, k, m, n, sh.lda, sh.nv, d_u, d_sigma, d_v, stream);
    }

    jsvd_status plan_large(jsvd_plan &p, int64_t k, const double *d_a, double *d_u, double *d_sigma, double *d_v,
                           int32_t *d_sweeps, int32_t *d_stop, hipStream_t stream)
    {
        if (uses_blocked(p.cfg, p.n[k]))
            return plan_blocked(p, k, d_a, d_u, d_sigma, d_v, d_sweeps, d_stop, stream);
        const int64_t m = p.m[k];
        const int n = static_cast<int>(p.n[k]);
        const double *a = d_a + p.a_off[k];
        // Row-major A -> column-major workspace (padded-LDS transpose).
        JSVD_HIP(jsvd::launch_transpose_f64(a, p.ws_a.ptr, m, n, nullptr, 0.0, stream));
        JSVD_HIP(jsvd::launch_identity_f64(p.ws_v.ptr, n, stream));
        JSVD_HIP(jsvd::launch_fill_i32(p.ws_flags.ptr, 0, p.cfg.max_sweeps, stream));
        const int half = (n + (n & 1)) / 2;
        int32_t *sweeps_slot = d_sweeps ? d_sweeps + k : nullptr;
        // Resident workgroups of the grid-barrier kernel, from the plan.  The
        // barrier words and the fault word start every launch at zero (a
        // launch that timed out leaves them anywhere).
        const int grid = std::max(1, std::min(p.big_resident, half));
        JSVD_HIP(jsvd::launch_fill_i32(reinterpret_cast<int32_t *>(p.ws_bar.ptr), 0, 2, stream));
        JSVD_HIP(jsvd::launch_fill_i32(reinterpret_cast<int32_t *>(p.d_fault.ptr), 0, 1, stream));
        JSVD_HIP(jsvd::launch_persistent_sweeps_f64(
            p.big_block, grid, p.ws_a.ptr, p.ws_v.ptr, m, n, p.cfg.epsilon, p.cfg.max_sweeps, p.ws_flags.ptr,
            p.ws_bar.ptr, sweeps_slot, d_stop ? d_stop + k : nullptr, p.d_fault.ptr,
            barrier_spin_limit(p.device, 8.0 * double(m) * n), stream));
        return finish_large(p, k, m, n, m, n, d_u, d_sigma, d_v, stream);
    }

    // Plain-path matrices too large for LDS take the wave kernels up to this
    // many columns: a group of at most 64 sixteen-wave or 256 four-wave
    // workgroups (one flag word each) holds one wave per pair of 2048
    // columns.  Measured per sweep against one workgroup per pair with a grid
    // barrier: 8192 x 2048 175 vs 374 ms, 2048 x 1024 26 vs 68 ms, 1100 x 520
    // 6.3 vs 11.4 ms.
    constexpr int64_t kWaveMaxColumns = 2048;

    // A launch of `count` matrices whose widest has `max_half` column pairs
    // per round runs one workgroup per matrix when it has more than
    // 4 + max_half^2 / 200 matrices of at most 512 columns: alone, a
    // sixteen-wave workgroup takes half / 16 pairs per wave and round, while
    // the grouped form's rounds slow down as more groups share the GPU.
    // Measured on MI355X (grouped form with sixteen-wave workgroups and flag
    // barriers), the two forms cross at about 30 matrices of 160 x 150 and 90
    // of 256 x 256, and tie at 96 of 512 x 512.  Wider matrices keep the
    // barrier kernels whatever their count (one workgroup would take dozens
    // of pairs per wave and round).
    // The ring form of the grouped wave solves is the default; JSVD_WAVE_RING=0
    // keeps the barrier form (for comparisons).
    bool ring_enabled()
    {
        const char *env = std::getenv("JSVD_WAVE_RING");
        return !(env && env[0] == '0');
    }

    // MI355X: 8 XCDs; the dispatcher hands workgroup b of a launch to XCD
    // b % 8.  The ring form places each group on one XCD (its kernel checks
    // the placement and falls back to device-scope mailbox stores when a
    // group is spread); JSVD_WAVE_RING_XCD=0 keeps the group-major grid.
    constexpr int kXcds = 8;
    bool xcd_ring_enabled()
    {
        const char *env = std::getenv("JSVD_WAVE_RING_XCD");
        return !(env && env[0] == '0');
    }

    // The ring form takes two LDS slot sets (one workgroup barrier per round
    // instead of two) when twice its LDS still leaves room for every group
    // of the launch where it is placed; JSVD_WAVE_RING_DB=0 keeps one set.
    bool ring_db_enabled()
    {
        const char *env = std::getenv("JSVD_WAVE_RING_DB");
        return !(env && env[0] == '0');
    }

    bool wave_block_mode(int64_t count, int64_t max_half)
    {
        return max_half <= 256 && count > 4 + max_half * max_half / 200;
    }

    // The same choice for a launch of matrices taller than the register
    // kernels (rows streamed in both forms): the streamed per-matrix
    // workgroup is slower per matrix, so it pays off only once the matrices
    // outnumber 64 + max_half^2 / 100.  Measured on MI355X, the forms cross at
    // about 90 matrices of 1024 x 64 and 70 of 2000 x 100; at 128 of 600 x 300
    // the grouped form still wins (354 vs 458 ms).
    bool tall_block_mode(int64_t count, int64_t max_half)
    {
        return max_half <= 256 && count > 64 + max_half * max_half / 100;
    }

    // Every grouped-wave matrix of the plan: row-major A into its
    // column-major workspace, V <- I, the grouped sweep launches, then sigma,
    // U = A / sigma and V back to the caller's row-major layout.
    jsvd_status plan_wave(jsvd_plan &p, const double *d_a, double *d_u, double *d_sigma, double *d_v,
                          int32_t *d_sweeps, int32_t *d_stop, hipStream_t stream)
    {
        const int64_t entries = static_cast<int64_t>(p.wave.size());
        JSVD_HIP(jsvd::launch_fill_i32(p.wv_flags.ptr, 0, entries * p.cfg.max_sweeps, stream));
        for (int64_t e = 0; e < entries; ++e)
        {
            if (p.wave_ring[e])
                continue;
            const int64_t k = p.wave[e];
            const int64_t m = p.m[k];
            const int n = static_cast<int>(p.n[k]);
            JSVD_HIP(jsvd::launch_transpose_f64(d_a + p.a_off[k], p.wv_a.ptr + p.wave_a_off[e], m, n, nullptr, 0.0,
                                                stream));
            JSVD_HIP(jsvd::launch_identity_f64(p.wv_v.ptr + p.wave_v_off[e], n, stream));
        }
        for (const auto &l : p.wave_launches)
        {
            jsvd::WaveGroupArgs args{};
            args.mats = p.d_wave.ptr + l.first;
            args.count = l.count;
            args.members = l.members;
            args.eps = p.cfg.epsilon;
            args.max_sweeps = p.cfg.max_sweeps;
            args.bars = p.wv_bars.ptr;
            args.sweeps = d_sweeps;
            args.stop = d_stop;
            args.fault = p.d_fault.ptr;
            args.spin_limit = barrier_spin_limit(p.device, p.wave_bytes);
            if (l.members == 0)
                JSVD_HIP(jsvd::launch_wave_block_sweeps_f64(l.m_max, l.count, args, stream));
            else if (l.ring)
            {
                // Flag words, the fault word and every mailbox word start the
                // launch at zero (a stale round tag must never match).
                JSVD_HIP(jsvd::launch_fill_i32(reinterpret_cast<int32_t *>(p.wv_bars.ptr), 0,
                                               int64_t(l.groups) * l.members, stream));
                JSVD_HIP(jsvd::launch_fill_i32(reinterpret_cast<int32_t *>(p.d_fault.ptr), 0, 1, stream));
                JSVD_HIP(jsvd::launch_fill_i32(reinterpret_cast<int32_t *>(p.wv_mbox.ptr), 0,
                                               2 * int64_t(l.groups) * l.mbox_group, stream));
                args.mbox = p.wv_mbox.ptr;
                args.mbox_group = l.mbox_group;
                args.in_a = d_a;
                args.out_u = d_u;
                args.out_s = d_sigma;
                args.out_v = d_v;
                args.ring_groups = l.groups;
                args.xcd_spread = l.xcd_spread;
                args.xcc = p.wv_xcc.ptr;
                args.ring_db = l.ring_db ? 1 : 0;
                JSVD_HIP(jsvd::launch_wave_ring_sweeps_f64(l.m_max, l.n_max, l.groups, args, stream));
            }
            else
            {
                // The group barriers' flag words and the fault word start
                // every launch at zero.
                JSVD_HIP(jsvd::launch_fill_i32(reinterpret_cast<int32_t *>(p.wv_bars.ptr), 0,
                                               int64_t(l.groups) * l.members, stream));
                JSVD_HIP(jsvd::launch_fill_i32(reinterpret_cast<int32_t *>(p.d_fault.ptr), 0, 1, stream));
                JSVD_HIP(jsvd::launch_wave_group_sweeps_f64(l.m_max, l.waves, l.groups, args, stream));
            }
        }
        for (int64_t e = 0; e < entries; ++e)
        {
            if (p.wave_ring[e])
                continue;
            const int64_t k = p.wave[e];
            const int64_t m = p.m[k];
            const int n = static_cast<int>(p.n[k]);
            const double *a_cm = p.wv_a.ptr + p.wave_a_off[e];
            double *sigma = d_sigma + p.s_off[k];
            JSVD_HIP(jsvd::launch_column_norms_f64(a_cm, m, sigma, m, n, stream));
            JSVD_HIP(jsvd::launch_transpose_ld_f64(a_cm, m, d_u + p.a_off[k], n, n, m, sigma, p.cfg.epsilon, stream));
            JSVD_HIP(jsvd::launch_transpose_ld_f64(p.wv_v.ptr + p.wave_v_off[e], n, d_v + p.v_off[k], n, n, n,
                                                   nullptr, 0.0, stream));
        }
        return JSVD_OK;
    }

    // Orders the grouped-wave entries (row-register count, then longest
    // first), sizes each launch and allocates the per-matrix workspaces.
    // A launch of many matrices (wave_block_mode) runs one workgroup per
    // matrix (members = 0): no cross-workgroup barrier, and the dispatcher
    // hands the longest matrices out first as CUs free up.  A smaller launch
    // runs groups of workgroups with barriers of their own: every
    // workgroup of it must be resident at once (the group barriers wait for
    // each other's members), so it has at most the occupancy-guaranteed
    // resident count of workgroups; a group gets one wave per column pair of
    // its largest matrix when the GPU has room for a group per matrix at
    // that size, else fewer members (each wave then takes several pairs per
    // round).
    jsvd_status plan_wave_setup(jsvd_plan &p, int cus)
    {
        // Row-register count; every matrix beyond the register kernels'
        // height shares one streamed launch.
        const int64_t rw_max = jsvd::wave_sweeps_max_rows() / 64;
        auto rw = [&](int64_t k) { return std::min<int64_t>((p.m[k] + 63) / 64, rw_max + 1); };
        auto cost = [&](int64_t k) { return double(p.n[k]) * p.n[k] * (p.m[k] + p.n[k]); };
        std::stable_sort(p.wave.begin(), p.wave.end(), [&](int64_t x, int64_t y) {
            if (rw(x) != rw(y))
                return rw(x) > rw(y);
            return cost(x) > cost(y);
        });
        const int64_t entries = static_cast<int64_t>(p.wave.size());
        for (int64_t e = 0; e < entries; ++e)
            p.wave_bytes += 8.0 * double(p.m[p.wave[e]]) * double(p.n[p.wave[e]]);
        int64_t max_flags = 1, max_mbox = 0, max_xcc = 0;
        for (int64_t e = 0; e < entries;)
        {
            jsvd_plan::WaveLaunch l;
            l.first = static_cast<int>(e);
            int64_t max_half = 1;
            while (e < entries && rw(p.wave[e]) == rw(p.wave[l.first]))
            {
                const int64_t k = p.wave[e];
                l.m_max = std::max(l.m_max, p.m[k]);
                l.n_max = std::max(l.n_max, p.n[k]);
                max_half = std::max<int64_t>(max_half, (p.n[k] + 1) / 2);
                ++e;
            }
            l.count = static_cast<int>(e - l.first);
            const bool streamed = l.m_max > jsvd::wave_sweeps_max_rows();
            // Ring form: a full group (one wave per pair, every column pair
            // held in registers) for at least one matrix at a time; groups
            // take the launch's matrices in turn when fewer groups than
            // matrices are resident.
            if (ring_enabled() && jsvd::ring_fits(l.m_max, l.n_max) && max_half >= 2)
            {
                const int waves = jsvd::ring_waves(l.m_max, l.n_max);
                int occupancy = 0;
                JSVD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &occupancy, jsvd::wave_ring_sweeps_f64_symbol(l.m_max, l.n_max), 64 * waves,
                    jsvd::ring_lds_bytes(l.m_max, l.n_max)));
                const int64_t resident = int64_t(occupancy) * cus;
                const int64_t members = (max_half + waves - 1) / waves;
                if (occupancy > 0 && members <= 256 && members <= resident)
                {
                    l.ring = true;
                    l.waves = waves;
                    l.members = static_cast<int>(members);
                    l.groups = static_cast<int>(std::min<int64_t>(l.count, resident / members));
                    // One XCD per group when the GPU's CUs split into kXcds
                    // equal XCDs and every XCD has room for its share of the
                    // groups: mailbox words then meet in that XCD's L2.
                    const int64_t per_xcd = cus % kXcds == 0 ? int64_t(occupancy) * (cus / kXcds) : 0;
                    if (xcd_ring_enabled() && members <= per_xcd)
                    {
                        const int64_t g8 = std::min<int64_t>(l.count, (per_xcd / members) * kXcds);
                        l.groups = static_cast<int>(g8);
                        l.xcd_spread = kXcds;
                    }
                    // Two slot sets when the doubled LDS still holds the
                    // launch's groups (on their XCDs when spread).
                    const int64_t lds2 = 2 * jsvd::ring_lds_bytes(l.m_max, l.n_max);
                    if (ring_db_enabled() && lds2 <= jsvd::kMaxLdsBytes)
                    {
                        int occ2 = 0;
                        JSVD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                            &occ2, jsvd::wave_ring_sweeps_f64_symbol(l.m_max, l.n_max), 64 * waves,
                            static_cast<int>(lds2)));
                        const int64_t need = int64_t(l.groups) * members;
                        const bool fits = l.xcd_spread > 0
                                              ? (int64_t(occ2) * (cus / kXcds) >= members &&
                                                 ((int64_t(occ2) * (cus / kXcds)) / members) * kXcds >= l.groups)
                                              : int64_t(occ2) * cus >= need;
                        l.ring_db = occ2 > 0 && fits;
                    }
                    max_xcc = std::max<int64_t>(max_xcc, int64_t(l.groups) * l.members);
                    l.mbox_group = jsvd::ring_mailbox_words(l.m_max, l.n_max, l.members);
                    max_flags = std::max<int64_t>(max_flags, int64_t(l.groups) * l.members);
                    max_mbox = std::max<int64_t>(max_mbox, int64_t(l.groups) * l.mbox_group);
                    p.wave_launches.push_back(l);
                    continue;
                }
            }
            if (streamed ? tall_block_mode(l.count, max_half) : wave_block_mode(l.count, max_half))
            {
                l.members = 0;
                l.groups = l.count;
                p.wave_launches.push_back(l);
                continue;
            }
            // One or two matrices: eight-wave workgroups, or four-wave ones (one
            // wave per SIMD, twice the workgroups) for matrices of more than
            // 512 rows and at most 512 columns, or more than 256 rows and at
            // most 256 columns, whose rounds are bound by each wave's row
            // loop, and when eight-wave groups would exceed 64
            // workgroups (so a lone wide matrix spreads over up to 256 CUs);
            // sixteen-wave groups for more matrices, or eight-wave ones for
            // matrices of more than 512 rows when a group of eight-wave
            // workgroups per matrix still fits one workgroup per CU (twice the
            // CUs at work on the long row loops).
            const bool four = (l.m_max > 512 && max_half <= 256) || (l.m_max > 256 && max_half <= 128) ||
                              (max_half + 7) / 8 > 64;
            const bool spread = l.m_max > 512 && int64_t(l.count) * ((max_half + 7) / 8) <= cus;
            l.waves = l.count > 2 ? (spread ? 8 : 16) : (four ? 4 : 8);
            int occupancy = 0;
            JSVD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occupancy, jsvd::wave_group_sweeps_f64_symbol(l.m_max, l.waves), 64 * l.waves, 0));
            const int64_t resident = int64_t(std::max(1, occupancy)) * cus;
            const int64_t need = (max_half + l.waves - 1) / l.waves;
            const int64_t members = std::max<int64_t>(1, std::min(need, resident / l.count));
            // At most 256 members: four flag words per polling lane.
            l.members = static_cast<int>(std::min<int64_t>(std::min(members, resident), 256));
            l.groups = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(l.count, resident / l.members)));
            max_flags = std::max<int64_t>(max_flags, int64_t(l.groups) * l.members);
            p.wave_launches.push_back(l);
        }
        // Column-major workspaces for the matrices of the barrier forms (the
        // ring form reads and writes the caller's buffers directly).
        p.wave_ring.assign(entries, 0);
        for (const auto &l : p.wave_launches)
            if (l.ring)
                std::fill(p.wave_ring.begin() + l.first, p.wave_ring.begin() + l.first + l.count, 1);
        p.wave_a_off.assign(entries, 0);
        p.wave_v_off.assign(entries, 0);
        int64_t a_total = 0, v_total = 0;
        for (int64_t e = 0; e < entries; ++e)
        {
            if (p.wave_ring[e])
                continue;
            const int64_t k = p.wave[e];
            p.wave_a_off[e] = a_total;
            p.wave_v_off[e] = v_total;
            a_total += p.m[k] * p.n[k];
            v_total += p.n[k] * p.n[k];
        }
        JSVD_CHECK(p.wv_a.alloc(a_total));
        JSVD_CHECK(p.wv_v.alloc(v_total));
        JSVD_CHECK(p.wv_flags.alloc(entries * p.cfg.max_sweeps));
        JSVD_CHECK(p.wv_bars.alloc(max_flags));
        JSVD_CHECK(p.wv_mbox.alloc(max_mbox));
        JSVD_CHECK(p.wv_xcc.alloc(max_xcc));
        std::vector<jsvd::WaveMatrix> mats(entries);
        for (int64_t e = 0; e < entries; ++e)
        {
            const int64_t k = p.wave[e];
            mats[e].a = p.wave_ring[e] ? nullptr : p.wv_a.ptr + p.wave_a_off[e];
            mats[e].v = p.wave_ring[e] ? nullptr : p.wv_v.ptr + p.wave_v_off[e];
            mats[e].flags = p.wv_flags.ptr + e * p.cfg.max_sweeps;
            mats[e].m = static_cast<int32_t>(p.m[k]);
            mats[e].n = static_cast<int32_t>(p.n[k]);
            mats[e].index = static_cast<int32_t>(k);
            mats[e].a_off = p.a_off[k];
            mats[e].v_off = p.v_off[k];
            mats[e].s_off = p.s_off[k];
        }
        JSVD_CHECK(p.d_wave.alloc(entries));
        JSVD_HIP(hipMemcpy(p.d_wave.ptr, mats.data(), sizeof(jsvd::WaveMatrix) * entries, hipMemcpyHostToDevice));
        return JSVD_OK;
    }

    // Plans of the host-array entry points, reused while the batch shapes,
    // the configuration and the shape-class table are the same (a repeated
    // one_sided_jacobi_svd call of one shape then allocates nothing and
    // queries nothing).  At most kPlanCache per device, least recently used
    // dropped first.  Kept until the process ends (not destroyed at exit,
    // when the runtime may already be gone).
    constexpr size_t kPlanCache = 4;

    bool same_config(const jsvd_config &a, const jsvd_config &b)
    {
        return a.epsilon == b.epsilon && a.max_sweeps == b.max_sweeps && a.threads_per_block == b.threads_per_block &&
               a.layout_mode == b.layout_mode && a.layout_min_columns == b.layout_min_columns &&
               a.layout_min_elements == b.layout_min_elements && a.layout_auto_tune == b.layout_auto_tune &&
               a.layout_bench_reps == b.layout_bench_reps && a.layout_bench_sweeps == b.layout_bench_sweeps &&
               a.precision == b.precision && a.block_width == b.block_width && a.refine_sweeps == b.refine_sweeps;
    }

    struct CachedPlan
    {
        std::vector<int64_t> m, n;
        jsvd_config cfg{};
        std::vector<int> tune; // the shape-class table the plan was built with
        int ring = 7;          // ring form of the wave solves enabled (bit 0), its XCD placement (bit 1), two slot sets (bit 2)
        jsvd_plan *plan = nullptr;
        uint64_t last_use = 0;
    };

    std::vector<int> tune_table()
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        return std::vector<int>(g_tuned, g_tuned + kTuneBuckets);
    }

    struct DeviceContext
    {
        std::mutex mu;
        hipStream_t stream = nullptr;
        DevBuf<double> a, u, s, v;
        DevBuf<int32_t> sweeps;
        std::vector<CachedPlan> plans;
        uint64_t clock = 0;
    };

    std::mutex g_ctx_mu;
    std::vector<std::unique_ptr<DeviceContext>> g_ctx;

    jsvd_status device_context(DeviceContext **out)
    {
        int dev = 0;
        JSVD_HIP(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lock(g_ctx_mu);
        if (static_cast<int>(g_ctx.size()) <= dev)
            g_ctx.resize(dev + 1);
        if (!g_ctx[dev])
        {
            auto ctx = std::make_unique<DeviceContext>();
            JSVD_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            g_ctx[dev] = std::move(ctx);
        }
        *out = g_ctx[dev].get();
        return JSVD_OK;
    }
} // namespace

extern "C" {

void jsvd_config_default(jsvd_config *cfg)
{
    if (!cfg)
        return;
    cfg->epsilon = 1.0e-9;
    cfg->max_sweeps = 128;
    cfg->threads_per_block = 256;
    cfg->layout_mode = JSVD_LAYOUT_AUTO;
    cfg->layout_min_columns = 16;
    cfg->layout_min_elements = 4096;
    cfg->layout_auto_tune = 0;
    cfg->layout_bench_reps = 2;
    cfg->layout_bench_sweeps = 8;
    cfg->precision = JSVD_PRECISION_FP64;
    cfg->block_width = 0;
    cfg->refine_sweeps = 0;
}

const char *jsvd_last_error(void)
{
    return g_last_error.c_str();
}

int32_t jsvd_abi_version(void)
{
    return 2;
}

jsvd_status jsvd_plan_create(int64_t count, const int64_t *m, const int64_t *n, const jsvd_config *cfg_in,
                             jsvd_plan **out)
{
    g_last_error.clear();
    discard_stale_error();
    if (!out)
        return fail(JSVD_EINVAL, "plan output pointer is null.");
    *out = nullptr;
    if (count < 0 || (count > 0 && (!m || !n)))
        return fail(JSVD_EINVAL, "Input batch description is invalid.");
    jsvd_config cfg;
    jsvd_config_default(&cfg);
    if (cfg_in)
        cfg = *cfg_in;
    JSVD_CHECK(validate_config(cfg));
    std::vector<int64_t> a_off(count), v_off(count), s_off(count);
    int64_t totals[3];
    JSVD_CHECK(packed_layout(count, m, n, a_off.data(), v_off.data(), s_off.data(), totals));

    auto plan = std::make_unique<jsvd_plan>();
    JSVD_HIP(hipGetDevice(&plan->device));
    plan->cfg = cfg;
    plan->count = count;
    plan->m.assign(m, m + count);
    plan->n.assign(n, n + count);
    plan->a_off = std::move(a_off);
    plan->v_off = std::move(v_off);
    plan->s_off = std::move(s_off);
    int64_t big_mn = 0, big_nn = 0, big_n = 0;
    int64_t big_gram = 0, big_q = 0, big_pairs = 0;
    hipDeviceProp_t prop;
    JSVD_HIP(hipGetDeviceProperties(&prop, plan->device));
    const int generic_block = lds_kernel_block(cfg.threads_per_block);
    // (class, workgroups per CU by LDS, FP32 class or -1) -> bucket
    // (class, workgroups per CU by LDS, FP32 class or -1, QR kernel shape + 1 or 0)
    using GroupKey = std::tuple<int, int, int, int>;
    std::vector<std::pair<GroupKey, std::vector<jsvd::WorkItem>>> groups;
    bool any_mixed = false, any_precond = false;
    static const int kSlots[] = {32, 24, 16, 12, 8, 6, 4, 3, 2, 1};
    bool any_four_wave = false;
    for (int64_t k = 0; k < count && !any_four_wave; ++k)
        any_four_wave = tuned_class(m[k], n[k]) < 0 && jsvd::select_class_f64(m[k], n[k]) == 4;
    // Workgroups per CU a kernel's registers allow (LDS aside), per class and
    // for the generic kernel: buckets are keyed by the resident count a matrix
    // actually gets, so matrices that differ only in an LDS-allowed count
    // above the register limit share one bucket (and its work-list).
    std::vector<int> reg_slots(jsvd::class_count() + 1, 0);
    auto register_slots = [&](int cls, int block) -> int {
        int &slot = reg_slots[cls >= 0 ? cls : jsvd::class_count()];
        if (slot == 0)
        {
            const void *fn = cls >= 0 ? jsvd::lds_class_f64_symbol(cls) : jsvd::lds_batch_f64_symbol(block);
            int occ = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, block, 0) != hipSuccess)
            {
                (void)hipGetLastError();
                occ = 32;
            }
            slot = occ > 0 ? occ : 32;
        }
        return slot;
    };
    for (int64_t k = 0; k < count; ++k)
    {
        int cls = preferred_class(m[k], n[k], any_four_wave);
        int64_t need = cls >= 0 ? jsvd::class_lds_bytes(cls, m[k], n[k]) : jsvd::lds_bytes_f64(m[k], n[k]);
        if (need <= jsvd::kMaxLdsBytes)
        {
            jsvd::WorkItem w{};
            w.m = static_cast<int32_t>(m[k]);
            w.n = static_cast<int32_t>(n[k]);
            w.index = static_cast<int32_t>(k);
            w.a_off = plan->a_off[k];
            w.v_off = plan->v_off[k];
            w.s_off = plan->s_off[k];
            const int block = cls >= 0 ? jsvd::class_block(cls) : generic_block;
            int slots = 1;
            const int reg = register_slots(cls, block);
            for (int sl : kSlots)
                if (need * sl <= jsvd::kMaxLdsBytes && (block / 64) * sl <= 32 && sl <= reg)
                {
                    slots = sl;
                    break;
                }
            // Mixed precision prefers the QR preconditioner (QR of A and of
            // R^T, then the FP32 sweeps, polar factor, B = X Q and the FP64
            // refinement on X = R2^T, n x n); without it the stages run on A.
            // Either needs an FP32 class, a refinement class keeping V in
            // LDS and Q (n x n FP64) in LDS; other matrices of a mixed plan
            // take the plain FP64 path.
            int f32cls = -1, pre = 0;
            if (cfg.precision == JSVD_PRECISION_MIXED && cls >= 0 &&
                jsvd::orthonormalize_lds_bytes(n[k]) <= jsvd::kMaxLdsBytes)
            {
                const int rc = preferred_class(n[k], n[k], any_four_wave);
                const int fc = jsvd::select_f32_class(n[k], n[k]);
                if (jsvd::precond_fits(m[k], n[k]) && rc >= 0 && !jsvd::class_uses_vlog(rc) && fc >= 0)
                {
                    pre = 1 + jsvd::qrcp_variant(m[k], n[k]); // the QR kernel shape joins the bucket key
                    f32cls = fc;
                    cls = rc;
                    need = jsvd::class_lds_bytes(cls, n[k], n[k]);
                    for (int sl : kSlots)
                        if (need * sl <= jsvd::kMaxLdsBytes && (jsvd::class_block(cls) / 64) * sl <= 32 &&
                            sl <= register_slots(cls, jsvd::class_block(cls)))
                        {
                            slots = sl;
                            break;
                        }
                }
                else
                    f32cls = jsvd::select_f32_class(m[k], n[k]);
            }
            any_mixed = any_mixed || f32cls >= 0;
            any_precond = any_precond || pre;
            const GroupKey key{cls, slots, f32cls, pre};
            auto it = std::find_if(groups.begin(), groups.end(), [&](auto &g) { return g.first == key; });
            if (it == groups.end())
            {
                groups.push_back({key, {}});
                it = groups.end() - 1;
            }
            it->second.push_back(w);
        }
        else if (!uses_blocked(cfg, n[k]) && m[k] <= jsvd::wave_sweeps_max_rows())
            plan->wave.push_back(k);
        else if (!uses_blocked(cfg, n[k]) && n[k] <= kWaveMaxColumns && m[k] <= INT32_MAX)
            plan->wave.push_back(k); // the wave kernels count rows in 32-bit ints
        else
        {
            plan->large.push_back(k);
            int64_t nv = n[k], rows = m[k];
            if (uses_blocked(cfg, n[k]))
            {
                const BlockedShape sh = blocked_shape(cfg.block_width, m[k], n[k]);
                const int64_t w2 = 4 * int64_t(sh.block) * sh.block;
                nv = sh.nv;
                rows = sh.lda;
                big_gram = std::max(big_gram, int64_t(sh.pairs) * sh.chunks * w2);
                big_q = std::max(big_q, int64_t(sh.pairs) * w2);
                big_pairs = std::max(big_pairs, int64_t(sh.pairs));
                ++plan->blocked;
            }
            big_mn = std::max(big_mn, rows * nv);
            big_nn = std::max(big_nn, nv * nv);
            big_n = std::max(big_n, n[k]);
        }
    }
    plan->a_elems = totals[0];
    plan->v_elems = totals[1];
    plan->s_elems = totals[2];

    // Longest-processing-time-first inside each bucket (a sweep costs about
    // n^2 (m + n)), and the buckets with the most work are launched first.
    auto cost = [](const jsvd::WorkItem &x) { return double(x.n) * x.n * (x.m + x.n); };
    // A mixed-precision bucket is a chain of dependent launches (QR, FP32
    // sweeps, orthonormal factor and B, FP64 refinement, back-transform), and
    // each ends in a tail where few workgroups still run.  A large bucket of
    // wide matrices is split into parts (items of its longest-first order
    // dealt out in turn) that the stream assignment below puts on streams of
    // their own: one part's stages then fill the other's tails (1024 x
    // 128x128 with the blocked QR: 10.3 ms whole, 9.7 in two parts, 10.2 in
    // three, 10.0 in four, 10.6 in eight).  Buckets of at most 64 columns stay
    // whole: their stages are short, and four chains of five launches on
    // each stream cost more than the tails (uneven 4096 batch: 4.12 ms in
    // four parts, 2.98 ms whole).  Each matrix's arithmetic does not depend on
    // its bucket.
    {
        const size_t initial = groups.size();
        for (size_t i = 0; i < initial; ++i)
        {
            int64_t widest = 0;
            for (const jsvd::WorkItem &w : groups[i].second)
                widest = std::max<int64_t>(widest, w.n);
            const size_t parts = std::min(kMixedSplitParts, groups[i].second.size() / kMixedSplitMin);
            if (std::get<2>(groups[i].first) < 0 || parts < 2 || widest <= kMixedSplitMinColumns)
                continue;
            std::vector<jsvd::WorkItem> items;
            items.swap(groups[i].second);
            std::stable_sort(items.begin(), items.end(),
                             [&](const jsvd::WorkItem &x, const jsvd::WorkItem &y) { return cost(x) > cost(y); });
            std::vector<std::vector<jsvd::WorkItem>> split(parts);
            for (size_t j = 0; j < items.size(); ++j)
                split[j % parts].push_back(items[j]);
            const GroupKey key = groups[i].first;
            groups[i].second.swap(split[0]);
            for (size_t q = 1; q < parts; ++q)
                groups.push_back({key, std::move(split[q])});
        }
    }
    for (auto &g : groups)
    {
        std::stable_sort(g.second.begin(), g.second.end(),
                         [&](const jsvd::WorkItem &x, const jsvd::WorkItem &y) { return cost(x) > cost(y); });
        JsvdBucket b;
        b.cls = std::get<0>(g.first);
        b.f32cls = std::get<2>(g.first);
        b.precond = std::get<3>(g.first) != 0;
        b.qr_variant = std::get<3>(g.first) - 1;
        b.block = b.cls >= 0 ? jsvd::class_block(b.cls) : generic_block;
        b.count = static_cast<int64_t>(g.second.size());
        for (const auto &w : g.second)
        {
            // A preconditioned bucket's sweeps run on X = R2^T (n x n).
            const int64_t rows = b.precond ? w.n : w.m;
            const int64_t need = b.cls >= 0 ? jsvd::class_lds_bytes(b.cls, rows, w.n) : jsvd::lds_bytes_f64(w.m, w.n);
            b.lds_bytes = std::max<int>(b.lds_bytes, static_cast<int>(need));
            b.work += cost(w);
            b.max_n = std::max<int64_t>(b.max_n, w.n);
            if (b.f32cls >= 0)
            {
                b.lds32 = std::max<int>(b.lds32, static_cast<int>(jsvd::f32_class_lds_bytes(b.f32cls, rows, w.n)));
                b.lds_ortho = std::max<int>(b.lds_ortho, static_cast<int>(jsvd::orthonormalize_lds_bytes(w.n)));
            }
            b.max_m = std::max<int64_t>(b.max_m, w.m);
        }
        if (b.precond)
        {
            // Both preconditioner kernels size their rows for the bucket's
            // tallest matrix.
            b.lds_qr = static_cast<int>(jsvd::qrcp_lds_bytes(b.qr_variant));
            // The second QR factors the n x n R^T of every matrix.
            b.qr2_variant = jsvd::qrcp_variant(b.max_n, b.max_n);
            b.lds_qr2 = static_cast<int>(jsvd::qrcp_lds_bytes(b.qr2_variant));
            b.lds_fin = static_cast<int>(jsvd::precond_finish_lds_bytes(b.qr_variant, b.max_m, b.max_n));
        }
        plan->buckets.push_back(b);
    }
    std::vector<size_t> order(groups.size());
    std::iota(order.begin(), order.end(), size_t(0));
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return plan->buckets[x].work > plan->buckets[y].work; });
    {
        std::vector<JsvdBucket> sorted;
        for (size_t i : order)
        {
            JsvdBucket b = plan->buckets[i];
            b.first = static_cast<int64_t>(plan->items.size());
            plan->items.insert(plan->items.end(), groups[i].second.begin(), groups[i].second.end());
            sorted.push_back(b);
        }
        plan->buckets.swap(sorted);
    }

    for (auto &b : plan->buckets)
    {
        const void *fn = b.cls >= 0 ? jsvd::lds_class_f64_symbol(b.cls) : jsvd::lds_batch_f64_symbol(b.block);
        // The attribute is a per-kernel upper bound shared by every plan (and
        // the no-allocation entry point): it is set to the budget, never to
        // one bucket's need, so a later plan cannot lower it under an
        // earlier one.  The occupancy query uses the bucket's own size.
        JSVD_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(jsvd::kMaxLdsBytes)));
        int occupancy = 0;
        JSVD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occupancy, fn, b.block, b.lds_bytes));
        if (occupancy <= 0)
            return fail(JSVD_EDEVICE, "work-list kernel cannot be resident with the requested LDS size.");
        const int64_t resident = int64_t(occupancy) * prop.multiProcessorCount;
        b.per_cu = occupancy;
        b.grid = static_cast<int>(std::min<int64_t>(resident, b.count));
        if (b.f32cls >= 0)
        {
            auto resident_of = [&](const void *f, int block, int lds, int *grid) -> jsvd_status {
                // Above 48 KiB of dynamic LDS the kernel needs the attribute;
                // it is set to the whole budget (the stage kernels' static LDS
                // is a few words).
                if (lds > 48 * 1024)
                    JSVD_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 int(jsvd::kMaxLdsBytes)));
                int occ = 0;
                JSVD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, f, block, lds));
                if (occ <= 0)
                    return fail(JSVD_EDEVICE, "mixed-precision stage cannot be resident with its LDS size.");
                *grid = static_cast<int>(std::min<int64_t>(int64_t(occ) * prop.multiProcessorCount, b.count));
                return JSVD_OK;
            };
            JSVD_CHECK(resident_of(jsvd::f32_class_symbol(b.f32cls), jsvd::f32_class_block(b.f32cls), b.lds32,
                                   &b.grid32));
            JSVD_CHECK(resident_of(jsvd::orthonormalize_symbol(), jsvd::orthonormalize_block(), b.lds_ortho,
                                   &b.grid_ortho));
            JSVD_CHECK(resident_of(jsvd::gemm_aq_symbol(), jsvd::gemm_aq_block(), 0, &b.grid_gemm));
            if (b.precond)
            {
                JSVD_CHECK(resident_of(jsvd::qrcp_symbol(b.qr_variant), jsvd::qrcp_block(b.qr_variant), b.lds_qr,
                                       &b.grid_qr));
                JSVD_CHECK(resident_of(jsvd::qrcp_symbol(b.qr2_variant), jsvd::qrcp_block(b.qr2_variant), b.lds_qr2,
                                       &b.grid_qr2));
                JSVD_CHECK(resident_of(jsvd::precond_finish_symbol(b.qr_variant, b.max_m), jsvd::precond_finish_block(b.qr_variant), b.lds_fin,
                                       &b.grid_fin));
                JSVD_CHECK(resident_of(jsvd::polar_bq_symbol(b.max_n), jsvd::polar_bq_block(), jsvd::polar_bq_lds_bytes(b.max_n),
                                       &b.grid_polar));
            }
        }
        if (b.cls >= 0 && jsvd::class_uses_vlog(b.cls))
        {
            // One tangent per (sweep, round, pair slot): max_sweeps * E(E-1)/2
            // doubles per workgroup.  Fewer workgroups if that exceeds 8 GiB.
            const int64_t even = b.max_n + (b.max_n & 1);
            const int64_t sweeps_logged = (b.f32cls >= 0 && cfg.refine_sweeps > 0) ? cfg.refine_sweeps : cfg.max_sweeps;
            b.vlog_stride = sweeps_logged * even * (even - 1) / 2;
            const int64_t cap = (int64_t(8) << 30) / 8;
            b.grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(b.grid, cap / b.vlog_stride)));
            DevBuf<double> *buf = new DevBuf<double>();
            plan->vlogs.emplace_back(buf);
            JSVD_CHECK(buf->alloc(size_t(b.grid) * b.vlog_stride));
            b.vlog = buf->ptr;
        }
    }
    if (!plan->items.empty())
    {
        JSVD_CHECK(plan->d_items.alloc(plan->items.size()));
        JSVD_HIP(hipMemcpy(plan->d_items.ptr, plan->items.data(), sizeof(jsvd::WorkItem) * plan->items.size(),
                           hipMemcpyHostToDevice));
        JSVD_CHECK(plan->d_counters.alloc(kCountersPerBucket * plan->buckets.size()));
        JSVD_HIP(hipMemset(plan->d_counters.ptr, 0, sizeof(unsigned int) * kCountersPerBucket * plan->buckets.size()));
        if (any_mixed)
            JSVD_CHECK(plan->q_scratch.alloc(plan->v_elems));
        if (any_precond)
        {
            // The square entries of X for the stages after the QR.
            plan->pre_items = plan->items;
            for (auto &w : plan->pre_items)
            {
                w.m = w.n;
                w.a_off = w.v_off;
            }
            JSVD_CHECK(plan->d_pre_items.alloc(plan->pre_items.size()));
            JSVD_HIP(hipMemcpy(plan->d_pre_items.ptr, plan->pre_items.data(),
                               sizeof(jsvd::WorkItem) * plan->pre_items.size(), hipMemcpyHostToDevice));
            JSVD_CHECK(plan->pc_refl.alloc(plan->a_elems));
            JSVD_CHECK(plan->pc_tau.alloc(plan->s_elems));
            JSVD_CHECK(plan->pc_tau2.alloc(plan->s_elems));
            // The blocked QR's panel factors T for Q1 and Q2.
            JSVD_CHECK(plan->pc_t1.alloc(jsvd::qrcp_t_elems(plan->s_elems, plan->count)));
            JSVD_CHECK(plan->pc_t2.alloc(jsvd::qrcp_t_elems(plan->s_elems, plan->count)));
            JSVD_CHECK(plan->pc_perm.alloc(plan->s_elems));
            for (DevBuf<double> *buf : {&plan->pc_x, &plan->pc_b, &plan->pc_ux, &plan->pc_vx, &plan->pc_refl2})
                JSVD_CHECK(buf->alloc(plan->v_elems));
        }
        const size_t streams = std::min<size_t>(plan->buckets.size(), 4);
        // Buckets to streams, longest first onto the least loaded stream;
        // a bucket's length is its work per resident workgroup (a bucket of
        // a few large matrices, one workgroup per CU, runs longest).
        {
            std::vector<size_t> by_len(plan->buckets.size());
            std::iota(by_len.begin(), by_len.end(), size_t(0));
            auto len = [&](size_t i) { return plan->buckets[i].work / std::max(1, plan->buckets[i].grid); };
            std::stable_sort(by_len.begin(), by_len.end(), [&](size_t x, size_t y) { return len(x) > len(y); });
            std::vector<double> load(streams, 0.0);
            for (size_t i : by_len)
            {
                const size_t s = static_cast<size_t>(std::min_element(load.begin(), load.end()) - load.begin());
                plan->buckets[i].stream = static_cast<int>(s);
                load[s] += len(i);
            }
        }
        if (streams > 1)
        {
            JSVD_HIP(hipEventCreateWithFlags(&plan->fork, hipEventDisableTiming));
            for (size_t i = 1; i < streams; ++i)
            {
                hipStream_t st;
                JSVD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
                plan->aux.push_back(st);
                hipEvent_t e;
                JSVD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                plan->join.push_back(e);
            }
        }
    }
    if (count > 0)
    {
        std::vector<jsvd::WorkItem> all(count);
        for (int64_t k = 0; k < count; ++k)
        {
            all[k].m = static_cast<int32_t>(m[k]);
            all[k].n = static_cast<int32_t>(n[k]);
            all[k].index = static_cast<int32_t>(k);
            all[k].a_off = plan->a_off[k];
            all[k].v_off = plan->v_off[k];
            all[k].s_off = plan->s_off[k];
            plan->max_rows = std::max(plan->max_rows, m[k]);
            plan->bw_lds_f64 = std::max(plan->bw_lds_f64, jsvd::backward_lds_bytes(m[k], n[k], 8));
            plan->bw_lds_f32 = std::max(plan->bw_lds_f32, jsvd::backward_lds_bytes(m[k], n[k], 4));
        }
        JSVD_CHECK(plan->d_all.alloc(count));
        JSVD_HIP(hipMemcpy(plan->d_all.ptr, all.data(), sizeof(jsvd::WorkItem) * count, hipMemcpyHostToDevice));
        for (int elem : {8, 4})
        {
            const int64_t need = elem == 8 ? plan->bw_lds_f64 : plan->bw_lds_f32;
            if (need <= jsvd::kMaxLdsBytes)
                JSVD_HIP(hipFuncSetAttribute(jsvd::backward_symbol(elem), hipFuncAttributeMaxDynamicSharedMemorySize,
                                             static_cast<int>(need)));
        }
    }
    if (!plan->wave.empty())
        JSVD_CHECK(plan_wave_setup(*plan, prop.multiProcessorCount));
    if (!plan->large.empty() || !plan->wave.empty())
    {
        JSVD_HIP(hipEventCreateWithFlags(&plan->big_in, hipEventDisableTiming));
        JSVD_HIP(hipEventCreateWithFlags(&plan->big_out, hipEventDisableTiming));
        JSVD_CHECK(plan->d_fault.alloc(1));
        JSVD_HIP(hipMemset(plan->d_fault.ptr, 0, sizeof(unsigned int)));
    }
    if (!plan->large.empty())
    {
        plan->big_block = persistent_kernel_block(cfg.threads_per_block);
        {
            int occupancy = 0;
            JSVD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occupancy, jsvd::persistent_sweeps_f64_symbol(plan->big_block), plan->big_block, 0));
            plan->big_resident = std::max(1, occupancy) * prop.multiProcessorCount;
        }
        JSVD_CHECK(plan->ws_a.alloc(big_mn));
        JSVD_CHECK(plan->ws_v.alloc(big_nn));
        JSVD_CHECK(plan->ws_s.alloc(big_n));
        JSVD_CHECK(plan->ws_flags.alloc(cfg.max_sweeps));
        JSVD_CHECK(plan->ws_bar.alloc(2));
        JSVD_HIP(hipMemset(plan->ws_bar.ptr, 0, sizeof(unsigned int) * 2));
        if (plan->blocked > 0)
        {
            JSVD_CHECK(plan->bk_gram.alloc(big_gram));
            JSVD_CHECK(plan->bk_q.alloc(big_q));
            JSVD_CHECK(plan->bk_flags.alloc(big_pairs + 4));
            JSVD_HIP(hipHostMalloc(reinterpret_cast<void **>(&plan->h_flag), 2 * sizeof(int32_t), hipHostMallocDefault));
            for (hipEvent_t &e : plan->bk_seen)
                JSVD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
    }
    JSVD_HIP(hipDeviceSynchronize());
    *out = plan.release();
    return JSVD_OK;
}

jsvd_status jsvd_plan_execute_f64(jsvd_plan *p, const double *d_a, double *d_u, double *d_sigma, double *d_v,
                                  int32_t *d_sweeps, void *stream_ptr)
{
    return jsvd_plan_execute_status_f64(p, d_a, d_u, d_sigma, d_v, d_sweeps, nullptr, stream_ptr);
}

jsvd_status jsvd_plan_execute_status_f64(jsvd_plan *p, const double *d_a, double *d_u, double *d_sigma,
                                         double *d_v, int32_t *d_sweeps, int32_t *d_stop, void *stream_ptr)
{
    g_last_error.clear();
    discard_stale_error();
    if (!p)
        return fail(JSVD_EINVAL, "plan is null.");
    if (p->count == 0)
        return JSVD_OK;
    if (!d_a || !d_u || !d_sigma || !d_v)
        return fail(JSVD_EINVAL, "device buffer pointer is null.");
    int dev = 0;
    JSVD_HIP(hipGetDevice(&dev));
    if (dev != p->device)
        return fail(JSVD_EINVAL, "plan was created on another device.");
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    if (!p->aux.empty())
    {
        JSVD_HIP(hipEventRecord(p->fork, stream));
        for (hipStream_t st : p->aux)
            JSVD_HIP(hipStreamWaitEvent(st, p->fork, 0));
    }
    for (size_t b = 0; b < p->buckets.size(); ++b)
    {
        const JsvdBucket &bk = p->buckets[b];
        const size_t lane = static_cast<size_t>(bk.stream);
        hipStream_t st = lane == 0 ? stream : p->aux[lane - 1];
        const jsvd::WorkItem *items = p->d_items.ptr + bk.first;
        unsigned int *ctr = p->d_counters.ptr + kCountersPerBucket * b;
        const bool mixed = bk.f32cls >= 0;
        // A preconditioned bucket: QR of each A (A = Q1 R), QR of R^T in
        // place (R^T = Q2 R2), then every mixed stage on X = R2^T (square
        // entries, preconditioner scratch), then U = Q1 Ux, V = P Q2 Vx.
        const bool pre = bk.precond;
        const jsvd::WorkItem *sq_items = pre ? p->d_pre_items.ptr + bk.first : items;
        jsvd::PrecondArgs pargs{};
        if (pre)
        {
            pargs.items = items;
            pargs.count = static_cast<int>(bk.count);
            pargs.a = d_a;
            pargs.refl = p->pc_refl.ptr;
            pargs.tau = p->pc_tau.ptr;
            pargs.perm = p->pc_perm.ptr;
            pargs.x = p->pc_x.ptr;
            pargs.refl2 = p->pc_refl2.ptr;
            pargs.tau2 = p->pc_tau2.ptr;
            pargs.tq = p->pc_t1.ptr;
            pargs.tq2 = p->pc_t2.ptr;
            pargs.ux = p->pc_ux.ptr;
            pargs.vx = p->pc_vx.ptr;
            pargs.sigma = d_sigma;
            pargs.u = d_u;
            pargs.v = d_v;
            pargs.eps = p->cfg.epsilon;
            pargs.counters = ctr + 8;
            JSVD_HIP(jsvd::launch_qrcp_f64(bk.qr_variant, bk.grid_qr, bk.lds_qr, pargs, st));
            // The second QR reads R^T from X and overwrites it with R2^T.
            jsvd::PrecondArgs qargs = pargs;
            qargs.items = sq_items;
            qargs.a = p->pc_x.ptr;
            qargs.refl = p->pc_refl2.ptr;
            qargs.tau = p->pc_tau2.ptr;
            qargs.tq = p->pc_t2.ptr;
            qargs.perm = nullptr;
            qargs.counters = ctr + 4; // the B = A Q pair, unused by preconditioned buckets
            JSVD_HIP(jsvd::launch_qrcp_f64(bk.qr2_variant, bk.grid_qr2, bk.lds_qr2, qargs, st));
        }
        if (mixed)
        {
            // FP32 sweeps -> Q = orth(V32) -> B = A Q (into the U buffer),
            // then the FP64 refinement below starts from B and Q.
            jsvd::MixedArgs margs{};
            margs.items = sq_items;
            margs.count = static_cast<int>(bk.count);
            margs.a = pre ? p->pc_x.ptr : d_a;
            margs.q = p->q_scratch.ptr;
            margs.u = pre ? p->pc_b.ptr : d_u;
            margs.sweeps = d_sweeps;
            margs.eps = p->cfg.epsilon;
            margs.max_sweeps = p->cfg.max_sweeps;
            margs.counters = ctr;
            JSVD_HIP(jsvd::launch_mixed_sweeps_f32(bk.f32cls, bk.grid32, bk.lds32, margs, st));
            if (pre)
            {
                // Q = polar factor of V32 and B = X Q, on MFMA.
                JSVD_HIP(jsvd::launch_polar_bq_f64(bk.max_n, bk.grid_polar, pargs, margs, st));
            }
            else
            {
                JSVD_HIP(jsvd::launch_orthonormalize_f64(bk.grid_ortho, bk.lds_ortho, margs, st));
                JSVD_HIP(jsvd::launch_gemm_aq_f64(bk.grid_gemm, margs, st));
            }
        }
        if (bk.cls >= 0)
        {
            jsvd::ClassArgs args{};
            args.items = sq_items;
            args.count = static_cast<int>(bk.count);
            args.a = pre ? p->pc_x.ptr : d_a;
            args.u = pre ? p->pc_ux.ptr : d_u;
            args.sigma = d_sigma;
            args.v = pre ? p->pc_vx.ptr : d_v;
            args.sweeps = d_sweeps;
            args.stop = d_stop;
            args.eps = p->cfg.epsilon;
            args.max_sweeps = p->cfg.max_sweeps;
            args.vlog = bk.vlog;
            args.vlog_stride = bk.vlog_stride;
            args.counters = ctr + 6;
            // One workgroup per CU: each such matrix is a long single solve
            // that other buckets' workgroups would otherwise slow down.
            args.priority = bk.per_cu == 1 ? 1 : 0;
            if (mixed)
            {
                args.b_cm = pre ? p->pc_b.ptr : d_u;
                args.q_init = p->q_scratch.ptr;
                args.add_sweeps = 1;
                args.max_sweeps = p->cfg.refine_sweeps > 0 ? p->cfg.refine_sweeps : p->cfg.max_sweeps;
                // X's left vectors give A's U: every nonzero column is
                // normalised here, the U rule applies in the back-transform.
                args.u_all_nonzero = pre ? 1 : 0;
            }
            JSVD_HIP(jsvd::launch_lds_class_f64(bk.cls, bk.grid, bk.lds_bytes, args, st));
            if (pre)
                JSVD_HIP(jsvd::launch_precond_finish_f64(bk.qr_variant, bk.max_m, bk.grid_fin, bk.lds_fin, pargs, st));
        }
        else
            JSVD_HIP(jsvd::launch_lds_batch_f64(bk.block, bk.grid, bk.lds_bytes, items, static_cast<int>(bk.count),
                                                d_a, d_u, d_sigma, d_v, d_sweeps, d_stop, p->cfg.epsilon,
                                                p->cfg.max_sweeps, ctr, st));
    }
    for (size_t i = 0; i < p->aux.size(); ++i)
    {
        JSVD_HIP(hipEventRecord(p->join[i], p->aux[i]));
        JSVD_HIP(hipStreamWaitEvent(stream, p->join[i], 0));
    }
    if (!p->large.empty() || !p->wave.empty())
    {
        // Large matrices run on the device's barrier stream, ordered after
        // the caller's prior work and before its later work.
        hipStream_t bs = nullptr;
        JSVD_CHECK(barrier_stream(p->device, &bs));
        JSVD_HIP(hipEventRecord(p->big_in, stream));
        JSVD_HIP(hipStreamWaitEvent(bs, p->big_in, 0));
        if (!p->wave.empty())
            JSVD_CHECK(plan_wave(*p, d_a, d_u, d_sigma, d_v, d_sweeps, d_stop, bs));
        for (int64_t k : p->large)
            JSVD_CHECK(plan_large(*p, k, d_a, d_u, d_sigma, d_v, d_sweeps, d_stop, bs));
        JSVD_HIP(hipEventRecord(p->big_out, bs));
        JSVD_HIP(hipStreamWaitEvent(stream, p->big_out, 0));
    }
    return JSVD_OK;
}

jsvd_status jsvd_plan_metrics_f64(jsvd_plan *p, const double *d_a, const double *d_u, const double *d_sigma,
                                  const double *d_v, double *d_out, void *stream_ptr)
{
    g_last_error.clear();
    discard_stale_error();
    if (!p)
        return fail(JSVD_EINVAL, "plan is null.");
    if (p->count == 0)
        return JSVD_OK;
    if (!d_a || !d_u || !d_sigma || !d_v || !d_out)
        return fail(JSVD_EINVAL, "device buffer pointer is null.");
    int dev = 0;
    JSVD_HIP(hipGetDevice(&dev));
    if (dev != p->device)
        return fail(JSVD_EINVAL, "plan was created on another device.");
    if (p->mt_count < 0)
    {
        // 32 x 32 output tiles per matrix: U^T U and V^T V (n x n each), then
        // the m x n reconstruction.  The tile kernel reads the row count as
        // a 32-bit int and the tile list is indexed by 32-bit ints.
        int64_t tile_total = 0;
        for (int64_t k = 0; k < p->count; ++k)
        {
            if (p->m[k] > INT32_MAX)
                return fail(JSVD_EINVAL, "metrics: matrices of more than 2^31 - 1 rows are not supported.");
            const int64_t tn = (p->n[k] + 31) / 32, tm = (p->m[k] + 31) / 32;
            tile_total += tn * (2 * tn + tm);
            if (tile_total > int64_t(0x7fffffff))
                return fail(JSVD_EINVAL, "metrics: the batch needs more than 2^31 - 1 output tiles.");
        }
        std::vector<jsvd::MetricTile> tiles;
        std::vector<int64_t> first(p->count + 1, 0);
        for (int64_t k = 0; k < p->count; ++k)
        {
            first[k] = static_cast<int64_t>(tiles.size());
            const int tn = static_cast<int>((p->n[k] + 31) / 32), tm = static_cast<int>((p->m[k] + 31) / 32);
            for (int kind = 0; kind < 3; ++kind)
                for (int ti = 0; ti < (kind == 2 ? tm : tn); ++ti)
                    for (int tj = 0; tj < tn; ++tj)
                        tiles.push_back({static_cast<int32_t>(k), kind, ti, tj});
        }
        first[p->count] = static_cast<int64_t>(tiles.size());
        JSVD_CHECK(p->mt_tiles.alloc(tiles.size()));
        JSVD_CHECK(p->mt_first.alloc(first.size()));
        JSVD_CHECK(p->mt_partial.alloc(2 * std::max<size_t>(tiles.size(), 1)));
        if (!tiles.empty())
            JSVD_HIP(hipMemcpy(p->mt_tiles.ptr, tiles.data(), sizeof(jsvd::MetricTile) * tiles.size(),
                               hipMemcpyHostToDevice));
        JSVD_HIP(hipMemcpy(p->mt_first.ptr, first.data(), sizeof(int64_t) * first.size(), hipMemcpyHostToDevice));
        p->mt_count = static_cast<int64_t>(tiles.size());
    }
    JSVD_HIP(jsvd::launch_metrics_f64(p->mt_tiles.ptr, p->mt_count, p->mt_first.ptr, p->d_all.ptr,
                                      static_cast<int>(p->count), d_a, d_u, d_sigma, d_v, p->cfg.epsilon,
                                      p->mt_partial.ptr, d_out, static_cast<hipStream_t>(stream_ptr)));
    return JSVD_OK;
}

jsvd_status jsvd_packed_layout(int64_t count, const int64_t *m, const int64_t *n, int64_t *a_off, int64_t *v_off,
                               int64_t *s_off, int64_t *totals)
{
    g_last_error.clear();
    return packed_layout(count, m, n, a_off, v_off, s_off, totals);
}

jsvd_status jsvd_blocked_layout_f64(int64_t m, int64_t n, int32_t block_width, jsvd_blocked_layout *out)
{
    g_last_error.clear();
    if (!out)
        return fail(JSVD_EINVAL, "blocked layout output pointer is null.");
    if (block_width != 16 && block_width != 32)
        return fail(JSVD_EINVAL, "block_width must be 0, 16 or 32.");
    if (m <= 0 || n <= 0)
        return fail(JSVD_EINVAL, "Input matrix shape must be non-zero.");
    JSVD_CHECK(validate_shape(static_cast<uint64_t>(m), static_cast<uint64_t>(n)));
    if (n < 2 * int64_t(block_width))
        return fail(JSVD_EINVAL, "The blocked path needs n >= 2 * block_width.");
    const BlockedShape sh = blocked_shape(block_width, m, n);
    if (sh.lda > kMaxPackedElements / sh.nv)
        return fail(JSVD_EINVAL, "Packed batch exceeds 2^60 elements.");
    out->lda = sh.lda;
    out->nv = sh.nv;
    out->blocks = sh.nb;
    out->pairs = sh.pairs;
    out->chunk_rows = sh.chunk_rows;
    out->chunks = sh.chunks;
    out->a_elements = sh.lda * sh.nv;
    out->v_elements = int64_t(sh.nv) * sh.nv;
    out->gram_elements = int64_t(sh.pairs) * sh.chunks * 4 * int64_t(block_width) * block_width;
    return JSVD_OK;
}

jsvd_status jsvd_plan_get_info(const jsvd_plan *p, jsvd_plan_info *info)
{
    if (!p || !info)
        return fail(JSVD_EINVAL, "plan or info pointer is null.");
    info->count = p->count;
    info->lds_matrices = static_cast<int64_t>(p->items.size());
    info->large_matrices = static_cast<int64_t>(p->large.size() + p->wave.size());
    info->a_elements = p->a_elems;
    info->v_elements = p->v_elems;
    info->sigma_elements = p->s_elems;
    info->workgroup_size = p->buckets.empty() ? 0 : p->buckets[0].block;
    info->grid_size = p->buckets.empty() ? 0 : p->buckets[0].grid;
    info->lds_bytes = p->buckets.empty() ? 0 : p->buckets[0].lds_bytes;
    info->buckets = static_cast<int32_t>(p->buckets.size());
    info->precision = p->cfg.precision;
    info->blocked_matrices = static_cast<int32_t>(p->blocked);
    int64_t ring = 0;
    for (const auto &l : p->wave_launches)
        ring += l.ring ? l.count : 0;
    info->ring_matrices = static_cast<int32_t>(ring);
    info->pad = 0;
    return JSVD_OK;
}

void jsvd_plan_destroy(jsvd_plan *p)
{
    delete p;
}

} // extern "C"

namespace
{
    template <typename T>
    jsvd_status plan_backward(jsvd_plan *p, const T *u, const T *s, const T *v, const T *gu, const T *gs, const T *gv,
                              T *ga, void *stream_ptr)
    {
        g_last_error.clear();
        discard_stale_error();
        if (!p)
            return fail(JSVD_EINVAL, "plan is null.");
        if (p->count == 0)
            return JSVD_OK;
        if (!u || !s || !v || !ga)
            return fail(JSVD_EINVAL, "device buffer pointer is null.");
        const int64_t need = sizeof(T) == 8 ? p->bw_lds_f64 : p->bw_lds_f32;
        if (need > jsvd::kMaxLdsBytes)
            return fail(JSVD_EINVAL, "SVD backward needs (n'^2 + max(n', 64)(n'+1)) doubles of LDS per matrix (n' = n "
                                     "rounded up to 16); a matrix of this plan exceeds the 159 KiB budget.");
        if (p->max_rows > INT32_MAX)
            return fail(JSVD_EINVAL, "SVD backward supports matrices of up to 2^31 - 1 rows.");
        int dev = 0;
        JSVD_HIP(hipGetDevice(&dev));
        if (dev != p->device)
            return fail(JSVD_EINVAL, "plan was created on another device.");
        hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
        if constexpr (sizeof(T) == 8)
            JSVD_HIP(jsvd::launch_backward_f64(p->d_all.ptr, static_cast<int>(p->count), static_cast<int>(need), u, s,
                                               v, gu, gs, gv, ga, stream));
        else
            JSVD_HIP(jsvd::launch_backward_f32(p->d_all.ptr, static_cast<int>(p->count), static_cast<int>(need), u, s,
                                               v, gu, gs, gv, ga, stream));
        return JSVD_OK;
    }
} // namespace

extern "C" {

jsvd_status jsvd_plan_backward_f64(jsvd_plan *plan, const double *d_u, const double *d_sigma, const double *d_v,
                                   const double *d_gu, const double *d_gsigma, const double *d_gv, double *d_ga,
                                   void *stream)
{
    return plan_backward<double>(plan, d_u, d_sigma, d_v, d_gu, d_gsigma, d_gv, d_ga, stream);
}

jsvd_status jsvd_plan_backward_f32(jsvd_plan *plan, const float *d_u, const float *d_sigma, const float *d_v,
                                   const float *d_gu, const float *d_gsigma, const float *d_gv, float *d_ga,
                                   void *stream)
{
    return plan_backward<float>(plan, d_u, d_sigma, d_v, d_gu, d_gsigma, d_gv, d_ga, stream);
}

jsvd_status jsvd_solve_batched_f64(int64_t count, const double *const *a_rm, const uint64_t *m, const uint64_t *n,
                                   const jsvd_config *cfg, double *const *u_rm, double *const *sigma,
                                   double *const *v_rm, int32_t *sweeps)
{
    return jsvd_solve_batched_status_f64(count, a_rm, m, n, cfg, u_rm, sigma, v_rm, sweeps, nullptr);
}

jsvd_status jsvd_solve_batched_status_f64(int64_t count, const double *const *a_rm, const uint64_t *m,
                                          const uint64_t *n, const jsvd_config *cfg, double *const *u_rm,
                                          double *const *sigma, double *const *v_rm, int32_t *sweeps,
                                          int32_t *stop)
{
    g_last_error.clear();
    if (count < 0 || (count > 0 && (!a_rm || !m || !n || !u_rm || !sigma || !v_rm)))
        return fail(JSVD_EINVAL, "Input batch description is invalid.");
    if (count == 0)
        return JSVD_OK;
    std::vector<int64_t> mm(count), nn(count);
    for (int64_t k = 0; k < count; ++k)
    {
        JSVD_CHECK(validate_shape(m[k], n[k]));
        if (!a_rm[k] || !u_rm[k] || !sigma[k] || !v_rm[k])
            return fail(JSVD_EINVAL, "Input buffer size mismatch.");
        mm[k] = static_cast<int64_t>(m[k]);
        nn[k] = static_cast<int64_t>(n[k]);
    }
    {
        // All argument checks happen before the first device call.
        jsvd_config c;
        jsvd_config_default(&c);
        if (cfg)
            c = *cfg;
        JSVD_CHECK(validate_config(c));
    }
    DeviceContext *ctx = nullptr;
    JSVD_CHECK(device_context(&ctx));
    std::lock_guard<std::mutex> lock(ctx->mu);
    jsvd_config c;
    jsvd_config_default(&c);
    if (cfg)
        c = *cfg;
    const std::vector<int> tune = tune_table();
    const int ring = (ring_enabled() ? 1 : 0) | (xcd_ring_enabled() ? 2 : 0) | (ring_db_enabled() ? 4 : 0);
    jsvd_plan *plan = nullptr;
    for (CachedPlan &e : ctx->plans)
        if (e.m == mm && e.n == nn && same_config(e.cfg, c) && e.tune == tune && e.ring == ring)
        {
            plan = e.plan;
            e.last_use = ++ctx->clock;
            break;
        }
    if (!plan)
    {
        JSVD_CHECK(jsvd_plan_create(count, mm.data(), nn.data(), &c, &plan));
        if (ctx->plans.size() >= kPlanCache)
        {
            auto lru = std::min_element(ctx->plans.begin(), ctx->plans.end(),
                                        [](const CachedPlan &x, const CachedPlan &y) { return x.last_use < y.last_use; });
            // The context's stream is idle here (every call ends with a
            // stream synchronisation), so the plan can go.
            jsvd_plan_destroy(lru->plan);
            ctx->plans.erase(lru);
        }
        ctx->plans.push_back({mm, nn, c, tune, ring, plan, ++ctx->clock});
    }

    JSVD_CHECK(ctx->a.reserve(plan->a_elems));
    JSVD_CHECK(ctx->u.reserve(plan->a_elems));
    JSVD_CHECK(ctx->s.reserve(plan->s_elems));
    JSVD_CHECK(ctx->v.reserve(plan->v_elems));
    JSVD_CHECK(ctx->sweeps.reserve(2 * count));
    hipStream_t st = ctx->stream;
    for (int64_t k = 0; k < count; ++k)
        JSVD_HIP(hipMemcpyAsync(ctx->a.ptr + plan->a_off[k], a_rm[k], sizeof(double) * mm[k] * nn[k],
                                hipMemcpyHostToDevice, st));
    JSVD_CHECK(jsvd_plan_execute_status_f64(plan, ctx->a.ptr, ctx->u.ptr, ctx->s.ptr, ctx->v.ptr, ctx->sweeps.ptr,
                                            ctx->sweeps.ptr + count, st));
    for (int64_t k = 0; k < count; ++k)
    {
        JSVD_HIP(hipMemcpyAsync(u_rm[k], ctx->u.ptr + plan->a_off[k], sizeof(double) * mm[k] * nn[k],
                                hipMemcpyDeviceToHost, st));
        JSVD_HIP(hipMemcpyAsync(sigma[k], ctx->s.ptr + plan->s_off[k], sizeof(double) * nn[k],
                                hipMemcpyDeviceToHost, st));
        JSVD_HIP(hipMemcpyAsync(v_rm[k], ctx->v.ptr + plan->v_off[k], sizeof(double) * nn[k] * nn[k],
                                hipMemcpyDeviceToHost, st));
    }
    std::vector<int32_t> sw(2 * count, 0);
    JSVD_HIP(hipMemcpyAsync(sw.data(), ctx->sweeps.ptr, sizeof(int32_t) * 2 * count, hipMemcpyDeviceToHost, st));
    JSVD_HIP(hipStreamSynchronize(st));
    if (sweeps)
        std::copy(sw.begin(), sw.begin() + count, sweeps);
    if (stop)
        std::copy(sw.begin() + count, sw.end(), stop);
    if (std::find(sw.begin() + count, sw.end(), JSVD_STOP_BARRIER_TIMEOUT) != sw.end())
        return fail(JSVD_EDEVICE, kBarrierTimeoutMessage);
    return JSVD_OK;
}

jsvd_status jsvd_solve_f64(const double *a_rm, uint64_t m, uint64_t n, const jsvd_config *cfg, double *u_rm,
                           double *sigma, double *v_rm, int32_t *sweeps)
{
    g_last_error.clear();
    JSVD_CHECK(validate_shape(m, n));
    if (!a_rm || !u_rm || !sigma || !v_rm)
        return fail(JSVD_EINVAL, "Input buffer size mismatch.");
    return jsvd_solve_batched_f64(1, &a_rm, &m, &n, cfg, &u_rm, &sigma, &v_rm, sweeps);
}

jsvd_status jsvd_autotune(const jsvd_config *cfg, jsvd_autotune_report *out)
{
    g_last_error.clear();
    if (!out)
        return fail(JSVD_EINVAL, "report pointer is null.");
    jsvd_config c;
    jsvd_config_default(&c);
    if (cfg)
        c = *cfg;
    JSVD_CHECK(validate_config(c));
    // The benchmark mirrors the reference's (src/domain/layout_transpose_autotune.cu:78-126):
    // deterministic sin/cos matrices, a capped sweep count, one warm-up and
    // `layout_bench_reps` timed repetitions per candidate.  What it tunes here
    // is the shape class (workgroup size and lanes per column pair) of each
    // column-count bucket; the layout thresholds are reported unchanged
    // because the solver always works on a column-major copy.
    jsvd_config bench = c;
    bench.max_sweeps = std::min(c.max_sweeps, std::max(1, c.layout_bench_sweeps));
    bench.precision = JSVD_PRECISION_FP64;
    out->executed = 1;
    out->recommended_min_columns = c.layout_min_columns;
    out->recommended_min_elements = c.layout_min_elements;
    out->estimated_best_speedup = 1.0;
    out->sample_count = 0;

    DeviceContext *ctx = nullptr;
    JSVD_CHECK(device_context(&ctx));
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t st = ctx->stream;
    // Events are released and a bucket's class choice is restored on every
    // return path (a failed candidate must not stay installed).
    struct Scope
    {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        int bucket = -1, saved = -1;
        ~Scope()
        {
            if (e0)
                (void)hipEventDestroy(e0);
            if (e1)
                (void)hipEventDestroy(e1);
            if (bucket >= 0)
            {
                std::lock_guard<std::mutex> tl(g_tune_mu);
                g_tuned[bucket] = saved;
            }
        }
    } scope;
    JSVD_HIP(hipEventCreate(&scope.e0));
    JSVD_HIP(hipEventCreate(&scope.e1));
    hipEvent_t e0 = scope.e0, e1 = scope.e1;
    int64_t lo = 0;
    for (int b = 0; b < kTuneBuckets; ++b)
    {
        const int64_t hi = kTuneUpper[b];
        constexpr int64_t kCount = 512;
        std::vector<int64_t> ms(kCount), ns(kCount);
        for (int64_t k = 0; k < kCount; ++k)
        {
            ns[k] = lo + 1 + (k * 7919) % (hi - lo);
            ms[k] = ns[k] + (k * 104729) % (ns[k] + 1);
        }
        const int64_t nmax = hi, mmax = 2 * hi;
        std::vector<int> cands;
        for (int cls = 0; cls < jsvd::class_count(); ++cls)
            if (jsvd::class_fits(cls, mmax, nmax) && jsvd::class_fits(cls, lo + 1, lo + 1))
                cands.push_back(cls);
        lo = hi;
        if (cands.size() < 2)
            continue;
        int64_t total = 0, vtotal = 0, stotal = 0;
        for (int64_t k = 0; k < kCount; ++k)
        {
            total += ms[k] * ns[k];
            vtotal += ns[k] * ns[k];
            stotal += ns[k];
        }
        std::vector<double> host(total);
        for (int64_t k = 0, off = 0; k < kCount; ++k)
            for (int64_t i = 0; i < ms[k]; ++i)
                for (int64_t j = 0; j < ns[k]; ++j)
                    host[off++] = std::sin(0.013 * double((i + 1) * (j + 1))) +
                                  std::cos(0.017 * double((i + 3) * (j + 5 + k)));
        JSVD_CHECK(ctx->a.reserve(total));
        JSVD_CHECK(ctx->u.reserve(total));
        JSVD_CHECK(ctx->v.reserve(vtotal));
        JSVD_CHECK(ctx->s.reserve(stotal));
        JSVD_HIP(hipMemcpyAsync(ctx->a.ptr, host.data(), sizeof(double) * total, hipMemcpyHostToDevice, st));
        double t_default = -1.0, t_best = 1e300;
        int best = -1;
        {
            std::lock_guard<std::mutex> tl(g_tune_mu);
            scope.bucket = b;
            scope.saved = g_tuned[b];
        }
        for (int cls : cands)
        {
            {
                std::lock_guard<std::mutex> tl(g_tune_mu);
                g_tuned[b] = cls;
            }
            jsvd_plan *plan = nullptr;
            JSVD_CHECK(jsvd_plan_create(kCount, ms.data(), ns.data(), &bench, &plan));
            std::unique_ptr<jsvd_plan, void (*)(jsvd_plan *)> guard(plan, jsvd_plan_destroy);
            JSVD_CHECK(jsvd_plan_execute_f64(plan, ctx->a.ptr, ctx->u.ptr, ctx->s.ptr, ctx->v.ptr, nullptr, st));
            float ms_total = 0.0f;
            JSVD_HIP(hipEventRecord(e0, st));
            for (int r = 0; r < c.layout_bench_reps; ++r)
                JSVD_CHECK(jsvd_plan_execute_f64(plan, ctx->a.ptr, ctx->u.ptr, ctx->s.ptr, ctx->v.ptr, nullptr, st));
            JSVD_HIP(hipEventRecord(e1, st));
            JSVD_HIP(hipEventSynchronize(e1));
            JSVD_HIP(hipEventElapsedTime(&ms_total, e0, e1));
            const double t = double(ms_total) / c.layout_bench_reps;
            out->sample_count += 1;
            if (cls == jsvd::select_class_f64(mmax, nmax))
                t_default = t;
            if (t < t_best)
            {
                t_best = t;
                best = cls;
            }
        }
        {
            std::lock_guard<std::mutex> tl(g_tune_mu);
            g_tuned[b] = best; // the winner stays installed
            scope.bucket = -1;
        }
        if (t_default > 0.0 && t_best > 0.0)
            out->estimated_best_speedup = std::max(out->estimated_best_speedup, t_default / t_best);
    }
    return JSVD_OK;
}

jsvd_status jsvd_autotune_batch(int64_t count, const int64_t *m, const int64_t *n, const jsvd_config *cfg,
                                const double *d_a, double *d_u, double *d_sigma, double *d_v, int32_t reps,
                                void *stream, jsvd_batch_tune_report *out)
{
    g_last_error.clear();
    if (count <= 0 || !m || !n || !d_a || !d_u || !d_sigma || !d_v || !out)
        return fail(JSVD_EINVAL, "Input batch description is invalid.");
    if (reps < 1 || reps > 1000)
        return fail(JSVD_EINVAL, "reps must be in [1, 1000].");
    for (int64_t k = 0; k < count; ++k)
        JSVD_CHECK(validate_shape(static_cast<uint64_t>(m[k]), static_cast<uint64_t>(n[k])));
    jsvd_config c;
    jsvd_config_default(&c);
    if (cfg)
        c = *cfg;
    JSVD_CHECK(validate_config(c));
    hipStream_t st = static_cast<hipStream_t>(stream);

    // Matrices per bucket and the classes that fit every one of them.
    std::vector<std::vector<int>> cands(kTuneBuckets);
    std::vector<char> present(kTuneBuckets, 0);
    for (int64_t k = 0; k < count; ++k)
    {
        const int b = tune_bucket(n[k]);
        if (b >= 0)
            present[b] = 1;
    }
    for (int b = 0; b < kTuneBuckets; ++b)
    {
        if (!present[b])
            continue;
        for (int cls = 0; cls < jsvd::class_count(); ++cls)
        {
            bool fits = true;
            for (int64_t k = 0; k < count && fits; ++k)
                if (tune_bucket(n[k]) == b)
                    fits = jsvd::class_fits(cls, m[k], n[k]);
            if (fits)
                cands[b].push_back(cls);
        }
    }

    struct Scope
    {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        int bucket = -1, saved = -1;
        ~Scope()
        {
            if (e0)
                (void)hipEventDestroy(e0);
            if (e1)
                (void)hipEventDestroy(e1);
            if (bucket >= 0)
            {
                std::lock_guard<std::mutex> tl(g_tune_mu);
                g_tuned[bucket] = saved;
            }
        }
    } scope;
    JSVD_HIP(hipEventCreate(&scope.e0));
    JSVD_HIP(hipEventCreate(&scope.e1));
    out->samples = 0;
    // One warm-up, then the mean of `reps` executions of a plan built from
    // the current table.
    auto time_plan = [&](double &ms) -> jsvd_status {
        jsvd_plan *plan = nullptr;
        JSVD_CHECK(jsvd_plan_create(count, m, n, &c, &plan));
        std::unique_ptr<jsvd_plan, void (*)(jsvd_plan *)> guard(plan, jsvd_plan_destroy);
        JSVD_CHECK(jsvd_plan_execute_f64(plan, d_a, d_u, d_sigma, d_v, nullptr, st));
        JSVD_HIP(hipEventRecord(scope.e0, st));
        for (int r = 0; r < reps; ++r)
            JSVD_CHECK(jsvd_plan_execute_f64(plan, d_a, d_u, d_sigma, d_v, nullptr, st));
        JSVD_HIP(hipEventRecord(scope.e1, st));
        JSVD_HIP(hipEventSynchronize(scope.e1));
        float total = 0.0f;
        JSVD_HIP(hipEventElapsedTime(&total, scope.e0, scope.e1));
        ms = double(total) / reps;
        out->samples += 1;
        return JSVD_OK;
    };
    // The first timing runs while the clocks ramp up and is discarded.
    double t_cur = 0.0;
    JSVD_CHECK(time_plan(t_cur));
    JSVD_CHECK(time_plan(t_cur));
    out->ms_initial = t_cur;
    // Coordinate descent over the buckets, smallest column counts first: each
    // bucket keeps its current choice unless a candidate (the default -1 or a
    // class fitting all its matrices) is at least 1% faster on the whole plan,
    // so that timing noise does not flip choices.
    constexpr double kMargin = 0.99;
    for (int b = 0; b < kTuneBuckets; ++b)
    {
        if (!present[b])
            continue;
        int best;
        {
            std::lock_guard<std::mutex> tl(g_tune_mu);
            best = g_tuned[b];
            scope.bucket = b;
            scope.saved = best;
        }
        std::vector<int> trial{-1};
        trial.insert(trial.end(), cands[b].begin(), cands[b].end());
        for (int cls : trial)
        {
            if (cls == best)
                continue;
            {
                std::lock_guard<std::mutex> tl(g_tune_mu);
                g_tuned[b] = cls;
            }
            double t = 0.0;
            JSVD_CHECK(time_plan(t));
            if (t < kMargin * t_cur)
            {
                t_cur = t;
                best = cls;
            }
        }
        {
            std::lock_guard<std::mutex> tl(g_tune_mu);
            g_tuned[b] = best;
            scope.bucket = -1;
        }
    }
    out->ms_tuned = t_cur;
    {
        std::lock_guard<std::mutex> tl(g_tune_mu);
        for (int b = 0; b < kTuneBuckets; ++b)
            out->classes[b] = g_tuned[b];
    }
    return JSVD_OK;
}

int32_t jsvd_tune_class(int32_t bucket, int32_t cls)
{
    if (bucket < 0 || bucket >= kTuneBuckets)
        return -2;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    const int32_t prev = g_tuned[bucket];
    if (cls >= -1 && cls < jsvd::class_count())
        g_tuned[bucket] = cls;
    return prev;
}

int32_t jsvd_class_count(void)
{
    return jsvd::class_count();
}

} // extern "C"

// ---------------------------------------------------------------------------
// jsvd_solve_device_noalloc_f64: the device-resident batch of
// jsvd_plan_create + jsvd_plan_execute_f64 for callers that cannot allocate
// (code under stream capture, serving loops).  Its work-lists and counters
// live in the static array below; nothing is allocated on the device or the
// host, and every device step is a kernel on the caller's stream.
// ---------------------------------------------------------------------------
namespace
{
    // At most this many matrices per call; the scratch holds two work-lists of
    // that length (the batch-order staging copy and the bucket-ordered list
    // the solver reads) and the work-list counters.
    constexpr int64_t kNoAllocMaxMatrices = 32768;
    constexpr int kNoAllocBuckets = 16; // >= jsvd::class_count()
    constexpr int kNoAllocWidths = 257; // column counts 0 .. 256 for the longest-first order
    constexpr size_t kNoAllocListBytes = size_t(kNoAllocMaxMatrices) * sizeof(jsvd::WorkItem);
    constexpr size_t kNoAllocCounterBytes = 2 * kNoAllocBuckets * sizeof(unsigned int);
    constexpr size_t kNoAllocScratchBytes = 2 * kNoAllocListBytes + 256;
    static_assert(kNoAllocCounterBytes <= 256, "counters fit the tail of the scratch");
    // Work-list entries per upload launch: they travel in the kernel
    // arguments (3848 bytes), so no host-to-device copy is needed.
    constexpr int kNoAllocChunk = 96;
} // namespace

// 2 x 1.25 MiB + 256 B on every device the library runs on, held until the
// process ends.  Host code never uses this name as an address: it is only the
// symbol handed to hipGetSymbolAddress, whose per-device result is what the
// kernels receive.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace
{
    struct NoAllocChunk
    {
        int32_t count;
        int32_t base;
        jsvd::WorkItem items[kNoAllocChunk];
    };

    __global__ __launch_bounds__(128) void noalloc_store_items(NoAllocChunk chunk, jsvd::WorkItem *staging)
    {
        const int i = static_cast<int>(threadIdx.x);
        if (i < chunk.count)
            staging[chunk.base + i] = chunk.items[i];
    }

    struct NoAllocOrder
    {
        int32_t count;
        int32_t buckets;
        int32_t base[kNoAllocBuckets]; // first list entry of each bucket
    };

    // One workgroup: counting sort of the staged entries into bucket order
    // (the bucket is in the entry's pad field), widest matrices first inside
    // a bucket (the longest-first order of the plan's work-lists).  Entries of
    // equal width land in no fixed order, which changes no result: each
    // matrix is solved whole by one workgroup whatever its position.
    __global__ __launch_bounds__(1024) void noalloc_order(const jsvd::WorkItem *__restrict__ staging,
                                                          jsvd::WorkItem *__restrict__ list, NoAllocOrder args)
    {
        __shared__ int cursor[kNoAllocBuckets * kNoAllocWidths];
        const int tid = static_cast<int>(threadIdx.x);
        for (int i = tid; i < args.buckets * kNoAllocWidths; i += 1024)
            cursor[i] = 0;
        __syncthreads();
        auto bin = [](const jsvd::WorkItem &w) {
            return w.pad * kNoAllocWidths + (kNoAllocWidths - 1 - min(w.n, kNoAllocWidths - 1));
        };
        for (int k = tid; k < args.count; k += 1024)
            atomicAdd(&cursor[bin(staging[k])], 1);
        __syncthreads();
        if (tid < args.buckets)
        {
            int run = args.base[tid];
            for (int w = 0; w < kNoAllocWidths; ++w)
            {
                const int c = cursor[tid * kNoAllocWidths + w];
                cursor[tid * kNoAllocWidths + w] = run;
                run += c;
            }
        }
        __syncthreads();
        for (int k = tid; k < args.count; k += 1024)
        {
            const jsvd::WorkItem w = staging[k];
            list[atomicAdd(&cursor[bin(w)], 1)] = w;
        }
    }

    // Per device: the scratch array's address (hipGetSymbolAddress, once,
    // with that device current) and the CU count for the launch grids.
    struct NoAllocDevice
    {
        bool ready = false;
        unsigned char *ptr = nullptr;
        int cus = 0;
    };
    std::mutex g_noalloc_mu;
    NoAllocDevice g_noalloc[64];

    jsvd_status noalloc_device(int dev, NoAllocDevice *out)
    {
        if (dev < 0 || dev >= 64)
            return fail(JSVD_EINVAL, "jsvd_solve_device_noalloc_f64 supports devices 0..63.");
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        NoAllocDevice &d = g_noalloc[dev];
        if (!d.ready)
        {
            void *p = nullptr;
            JSVD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            int cus = 0;
            JSVD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            for (int c = 0; c < jsvd::class_count(); ++c)
                if (!jsvd::class_uses_vlog(c))
                    JSVD_HIP(hipFuncSetAttribute(jsvd::lds_class_f64_symbol(c),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 int(jsvd::kMaxLdsBytes)));
            d.ptr = static_cast<unsigned char *>(p);
            d.cus = cus;
            d.ready = true;
        }
        *out = d;
        return JSVD_OK;
    }
} // namespace

extern "C" jsvd_status jsvd_solve_device_noalloc_f64(int64_t count, const int64_t *m, const int64_t *n,
                                                     const jsvd_config *cfg_in, const double *d_a, double *d_u,
                                                     double *d_sigma, double *d_v, int32_t *d_sweeps,
                                                     void *stream_ptr)
{
    g_last_error.clear();
    discard_stale_error();
    if (count < 0 || (count > 0 && (!m || !n)))
        return fail(JSVD_EINVAL, "Input batch description is invalid.");
    if (count > kNoAllocMaxMatrices)
        return fail(JSVD_EINVAL, "jsvd_solve_device_noalloc_f64 takes at most 32768 matrices (the size of its "
                                 "static work-list); larger batches need jsvd_plan_create.");
    jsvd_config cfg;
    jsvd_config_default(&cfg);
    if (cfg_in)
        cfg = *cfg_in;
    JSVD_CHECK(validate_config(cfg));
    if (cfg.precision != JSVD_PRECISION_FP64)
        return fail(JSVD_EINVAL, "jsvd_solve_device_noalloc_f64 solves in FP64 only.");
    if (count == 0)
        return JSVD_OK;
    if (!d_a || !d_u || !d_sigma || !d_v)
        return fail(JSVD_EINVAL, "device buffer pointer is null.");

    // Pass 1: shapes, the class of every matrix (the plan's choice, so each
    // result equals the plan's bitwise), and per-class counts, work and LDS.
    bool four_wave = false;
    for (int64_t k = 0; k < count; ++k)
    {
        if (m[k] <= 0 || n[k] <= 0)
            return fail(JSVD_EINVAL, "Input matrix shape must be non-zero.");
        JSVD_CHECK(validate_shape(static_cast<uint64_t>(m[k]), static_cast<uint64_t>(n[k])));
        four_wave = four_wave || (tuned_class(m[k], n[k]) < 0 && jsvd::select_class_f64(m[k], n[k]) == 4);
    }
    const int classes = jsvd::class_count();
    int64_t per_class[kNoAllocBuckets] = {};
    double work[kNoAllocBuckets] = {};
    int lds[kNoAllocBuckets] = {};
    for (int64_t k = 0; k < count; ++k)
    {
        const int cls = preferred_class(m[k], n[k], four_wave);
        if (cls < 0 || jsvd::class_uses_vlog(cls))
            return fail(JSVD_EINVAL, "jsvd_solve_device_noalloc_f64 takes matrices that fit the LDS-resident "
                                     "work-list kernel with A and V in LDS (every m <= 160 with n <= 128 does, "
                                     "and every m <= 256 with n <= 64); larger ones need jsvd_plan_create.");
        per_class[cls] += 1;
        work[cls] += double(n[k]) * n[k] * (m[k] + n[k]);
        lds[cls] = std::max<int>(lds[cls], static_cast<int>(jsvd::class_lds_bytes(cls, m[k], n[k])));
    }
    // Buckets: one per class present, most work first (as the plan orders
    // its launches); all run on the caller's stream.
    int order[kNoAllocBuckets], buckets = 0, bucket_of[kNoAllocBuckets];
    for (int c = 0; c < classes; ++c)
        if (per_class[c] > 0)
            order[buckets++] = c;
    std::stable_sort(order, order + buckets, [&](int x, int y) { return work[x] > work[y]; });
    NoAllocOrder ord{};
    ord.count = static_cast<int32_t>(count);
    ord.buckets = buckets;
    for (int b = 0, first = 0; b < buckets; ++b)
    {
        bucket_of[order[b]] = b;
        ord.base[b] = first;
        first += static_cast<int>(per_class[order[b]]);
    }

    int dev = 0;
    JSVD_HIP(hipGetDevice(&dev));
    NoAllocDevice nd;
    JSVD_CHECK(noalloc_device(dev, &nd));
    auto *staging = reinterpret_cast<jsvd::WorkItem *>(nd.ptr);
    auto *list = reinterpret_cast<jsvd::WorkItem *>(nd.ptr + kNoAllocListBytes);
    auto *counters = reinterpret_cast<unsigned int *>(nd.ptr + 2 * kNoAllocListBytes);
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);

    // Pass 2: the entries in batch order (offsets of the packed layout),
    // uploaded through kernel arguments, then ordered on the device.
    NoAllocChunk chunk{};
    int64_t ao = 0, vo = 0, so = 0;
    for (int64_t k = 0; k < count; ++k)
    {
        jsvd::WorkItem &w = chunk.items[chunk.count++];
        w.m = static_cast<int32_t>(m[k]);
        w.n = static_cast<int32_t>(n[k]);
        w.index = static_cast<int32_t>(k);
        w.pad = bucket_of[preferred_class(m[k], n[k], four_wave)];
        w.a_off = ao;
        w.v_off = vo;
        w.s_off = so;
        ao += m[k] * n[k];
        vo += n[k] * n[k];
        so += n[k];
        if (chunk.count == kNoAllocChunk || k + 1 == count)
        {
            hipLaunchKernelGGL(noalloc_store_items, dim3(1), dim3(128), 0, stream, chunk, staging);
            JSVD_HIP(hipGetLastError());
            chunk.base += chunk.count;
            chunk.count = 0;
        }
    }
    hipLaunchKernelGGL(noalloc_order, dim3(1), dim3(1024), 0, stream, staging, list, ord);
    JSVD_HIP(hipGetLastError());
    // The solver's counters start at zero (a fill kernel, not a memset).
    JSVD_HIP(jsvd::launch_fill_i32(reinterpret_cast<int32_t *>(counters), 0, 2 * buckets, stream));
    for (int b = 0; b < buckets; ++b)
    {
        const int cls = order[b];
        int occ = 0;
        JSVD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, jsvd::lds_class_f64_symbol(cls),
                                                              jsvd::class_block(cls), lds[cls]));
        if (occ <= 0)
            return fail(JSVD_EDEVICE, "work-list kernel cannot be resident with the requested LDS size.");
        jsvd::ClassArgs args{};
        args.items = list + ord.base[b];
        args.count = static_cast<int>(per_class[cls]);
        args.a = d_a;
        args.u = d_u;
        args.sigma = d_sigma;
        args.v = d_v;
        args.sweeps = d_sweeps;
        args.eps = cfg.epsilon;
        args.max_sweeps = cfg.max_sweeps;
        args.counters = counters + 2 * b;
        args.priority = occ == 1 ? 1 : 0;
        const int grid = static_cast<int>(std::min<int64_t>(int64_t(occ) * nd.cus, per_class[cls]));
        JSVD_HIP(jsvd::launch_lds_class_f64(cls, grid, lds[cls], args, stream));
    }
    return JSVD_OK;
}
