// jsvd_kernels.hpp -- launch-side declarations of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace jsvd
{
    // One entry of the device work-list: a matrix and where its operands live
    // in the packed batch buffers (offsets in elements; U shares A's offset).
    struct WorkItem
    {
        int32_t m;
        int32_t n;
        int32_t index; // position in the caller's batch (sweeps output slot)
        int32_t pad;
        int64_t a_off;
        int64_t v_off;
        int64_t s_off;
    };

    // Dynamic LDS bytes the LDS-resident FP64 kernel needs for an m x n matrix.
    inline int64_t lds_bytes_f64(int64_t m, int64_t n)
    {
        return (m * n + n * n + n) * 8;
    }
    // Same for the mixed-precision kernel (FP32 A and V, FP64 sigma scratch).
    inline int64_t lds_bytes_mixed(int64_t m, int64_t n)
    {
        return (m * n + n * n) * 4 + n * 8;
    }

    // Dynamic LDS budget per workgroup: 160 KiB on gfx950 minus room for the
    // kernels' static LDS and the 16-byte alignment of the dynamic block.
    constexpr int64_t kMaxLdsBytes = 160 * 1024 - 1024;

    // stop (optional): per matrix, 0 when the last sweep rotated no pair
    // (converged), 1 when max_sweeps sweeps ran and the last still rotated.
    hipError_t launch_lds_batch_f64(int block, int grid, int lds_bytes, const WorkItem *items, int count,
                                    const double *a, double *u, double *sigma, double *v, int32_t *sweeps,
                                    int32_t *stop, double eps, int max_sweeps, unsigned int *counters,
                                    hipStream_t stream);
    const void *lds_batch_f64_symbol(int block);

    // Arguments of one launch of the shape-class work-list kernel.
    struct ClassArgs
    {
        const WorkItem *items;
        int count;
        const double *a;      // packed row-major inputs (offset a_off)
        const double *b_cm;   // optional packed column-major start matrices (offset a_off); replaces a
        const double *q_init; // optional packed column-major initial V (offset v_off); identity if null
        double *u;
        double *sigma;
        double *v;
        int32_t *sweeps;
        int32_t *stop;        // optional: 0 converged, 1 stopped at max_sweeps (slot w.index)
        int add_sweeps;       // 1: add this solve's sweeps to the slot instead of storing them
        int u_all_nonzero;    // 1: U = A / sigma wherever sigma > 0 (0: the reference's sigma > eps rule)
        double eps;
        int max_sweeps;
        double *vlog;         // rotation-log classes: per-workgroup tangent log
        int64_t vlog_stride;  // doubles per workgroup log
        unsigned int *counters;
        int priority;         // 1: the workgroup's waves issue at raised priority
    };

    // Shape classes of the LDS-resident kernel with compile-time lane groups.
    // select_class_f64 returns -1 when no class fits (the generic kernel then
    // takes the matrix).
    int select_class_f64(int64_t m, int64_t n);
    int class_count();
    bool class_fits(int cls, int64_t m, int64_t n);
    bool class_uses_vlog(int cls);
    int64_t class_lds_bytes(int cls, int64_t m, int64_t n);
    int class_block(int cls);
    const void *lds_class_f64_symbol(int cls);
    hipError_t launch_lds_class_f64(int cls, int grid, int lds_bytes, const ClassArgs &args, hipStream_t stream);

    // Mixed-precision stages (FP32 sweeps, FP64 re-orthonormalisation of V,
    // B = A Q on FP64 MFMA); the FP64 refinement is the class kernel.
    struct MixedArgs
    {
        const WorkItem *items;
        int count;
        const double *a;          // packed row-major A (offset a_off)
        double *q;                // packed column-major Q scratch (offset v_off)
        double *u;                // U buffer, holds B column-major between stages
        int32_t *sweeps;
        double eps;
        int max_sweeps;
        unsigned int *counters;   // six: two per stage
    };
    int select_f32_class(int64_t m, int64_t n);
    bool f32_class_fits(int cls, int64_t m, int64_t n);
    int64_t f32_class_lds_bytes(int cls, int64_t m, int64_t n);
    int f32_class_block(int cls);
    const void *f32_class_symbol(int cls);
    hipError_t launch_mixed_sweeps_f32(int cls, int grid, int lds_bytes, const MixedArgs &args, hipStream_t stream);
    int64_t orthonormalize_lds_bytes(int64_t n);
    const void *orthonormalize_symbol();
    int orthonormalize_block();
    hipError_t launch_orthonormalize_f64(int grid, int lds_bytes, const MixedArgs &args, hipStream_t stream);
    const void *gemm_aq_symbol();
    int gemm_aq_block();
    hipError_t launch_gemm_aq_f64(int grid, const MixedArgs &args, hipStream_t stream);

    // QR preconditioner of the mixed path: A P = Q1 R and R^T = Q2 R2
    // (Householder QR in FP64, twice), Jacobi on X = R2^T, then U = Q1 Ux
    // and V = P Q2 Vx.  The same QR kernel runs both factorisations: on A
    // (items with A's shapes) and in place on X = R^T (square items).
    struct PrecondArgs
    {
        const WorkItem *items;  // the batch's own shapes and offsets
        int count;
        const double *a;        // packed row-major input of the QR (a_off)
        double *refl;           // packed column-major m x n reflectors of Q1 (a_off)
        double *tau;            // n per matrix (s_off)
        int32_t *perm;          // n per matrix (s_off): column k of A P is column perm[k] of A; null: not written
        double *x;              // packed row-major n x n output R^T of the QR (v_off)
        const double *refl2;    // packed column-major n x n reflectors of Q2 (v_off)
        const double *tau2;     // n per matrix (s_off)
        // Blocked QR: the 16 x 16 triangular factor T of each panel (Q1 =
        // prod (I - V_p T_p V_p^T)), written by the QR at tq and read by the
        // back-transform from tq (Q1) and tq2 (Q2); matrix k's panels start
        // at 16 s_off + 240 k (index k), 256 doubles each.
        double *tq;
        const double *tq2;
        const double *ux;       // packed row-major n x n left vectors of X (v_off)
        const double *vx;       // packed row-major n x n right vectors of X (v_off)
        const double *sigma;    // sigma of X = sigma of A (s_off)
        double *u;              // output U (a_off)
        double *v;              // output V (v_off)
        double eps;             // U rule: columns with sigma <= eps are zero
        unsigned int *counters; // four: two per kernel
    };
    bool precond_fits(int64_t m, int64_t n);
    int qrcp_variant(int64_t m, int64_t n); // QR kernel shape for an m x n matrix, -1 if none
    bool qrcp_blocked(int variant);         // blocked QR on MFMA (writes T; needs the blocked back-transform)
    int64_t qrcp_t_elems(int64_t s_elems, int64_t count); // T buffer of a blocked-QR plan
    int64_t qrcp_lds_bytes(int variant);
    int64_t precond_finish_lds_bytes(int variant, int64_t m, int64_t n);
    const void *qrcp_symbol(int variant);
    int qrcp_block(int variant);
    const void *precond_finish_symbol(int variant, int64_t m); // for the bucket's QR variant and largest m
    int precond_finish_block(int variant);
    hipError_t launch_qrcp_f64(int variant, int grid, int lds_bytes, const PrecondArgs &args, hipStream_t stream);
    // Q = orthonormal factor of V32 (polar factor by a polynomial in
    // Q^T Q - I) and B = X Q on FP64 MFMA for a preconditioned bucket (ma:
    // its square entries, q, u = B buffer, counters 2-3).
    const void *polar_bq_symbol(int64_t n); // instance for a bucket's largest n (<= 128)
    int polar_bq_lds_bytes(int64_t n);
    int polar_bq_block(); // threads per workgroup
    hipError_t launch_polar_bq_f64(int64_t n, int grid, const PrecondArgs &pa, const MixedArgs &ma, hipStream_t stream);
    hipError_t launch_precond_finish_f64(int variant, int64_t m, int grid, int lds_bytes, const PrecondArgs &args,
                                         hipStream_t stream);

    // SVD adjoint: gA from (U, s, V) and upstream (gU, gs, gV); any upstream
    // pointer may be null (zero).  One workgroup per work-list entry.
    int64_t backward_lds_bytes(int64_t m, int64_t n, int elem);
    const void *backward_symbol(int elem);
    hipError_t launch_backward_f64(const WorkItem *items, int count, int lds_bytes, const double *u, const double *s,
                                   const double *v, const double *gu, const double *gs, const double *gv, double *ga,
                                   hipStream_t stream);
    hipError_t launch_backward_f32(const WorkItem *items, int count, int lds_bytes, const float *u, const float *s,
                                   const float *v, const float *gu, const float *gs, const float *gv, float *ga,
                                   hipStream_t stream);

    // Error metrics of computed SVDs (packed batch, row-major operands): per
    // matrix ||U^T U - D||_F (D = diag(sigma > eps)), ||V^T V - I||_F and
    // ||A - U diag(sigma) V^T||_F / ||A||_F, three doubles per matrix in
    // `out`.  The work is cut into 32 x 32 output tiles (kind 0: U^T U,
    // 1: V^T V, 2: the reconstruction), listed per matrix in that order;
    // first[k] .. first[k+1]-1 are matrix k's tiles and `partial` holds two
    // doubles per tile.
    struct MetricTile
    {
        int32_t matrix; // index into the batch-order work-list
        int32_t kind;
        int32_t ti, tj; // tile row and column
    };
    hipError_t launch_metrics_f64(const MetricTile *tiles, int64_t tile_count, const int64_t *first,
                                  const WorkItem *mats, int count, const double *a, const double *u,
                                  const double *sigma, const double *v, double eps, double *partial, double *out,
                                  hipStream_t stream);

    // Persistent plain-FP64 sweep kernel over one column-major matrix.
    // fault: one word, zero at launch; spin_limit: wall-clock ticks a grid
    // barrier may wait before the launch is marked faulted (stop reason 2).
    hipError_t launch_persistent_sweeps_f64(int block, int grid, double *a_cm, double *v_cm, int64_t m, int n,
                                            double eps, int max_sweeps, int32_t *flags, unsigned int *barrier,
                                            int32_t *sweeps_out, int32_t *stop_out, unsigned int *fault,
                                            uint64_t spin_limit, hipStream_t stream);
    const void *persistent_sweeps_f64_symbol(int block);
    // Matrices too large for LDS with m up to wave_sweeps_max_rows(): one
    // wave per column pair (workgroups of 8 or 16 waves), solved by groups of
    // `members` workgroups, each group with its own barrier, so one launch
    // runs `groups` independent solves at once.
    struct WaveMatrix
    {
        double *a;      // column-major m x n workspace (leading dimension m); null in the ring form
        double *v;      // column-major n x n workspace, identity at launch; null in the ring form
        int32_t *flags; // max_sweeps sweep flags, zero at launch
        int32_t m;
        int32_t n;
        int32_t index;  // sweeps output slot
        int32_t pad;
        int64_t a_off, v_off, s_off; // the matrix's offsets in the packed batch buffers (ring form)
    };
    struct WaveGroupArgs
    {
        const WaveMatrix *mats; // longest first; group g takes g, g + groups, ...
        int count;
        int members;            // workgroups per group
        double eps;
        int max_sweeps;
        unsigned int *bars;     // one flag word per workgroup (members per group), zero at launch
        int32_t *sweeps;        // optional
        int32_t *stop;          // optional: 0 converged, 1 stopped at max_sweeps, 2 barrier timed out
        unsigned int *fault;    // one word, zero at launch: set when a barrier wait hit spin_limit
        uint64_t spin_limit;    // wall-clock ticks a barrier may wait
        // Ring form only: column mailboxes, ring_mailbox_words(m, n, members)
        // 64-bit words per group, zero at launch; the packed row-major batch
        // buffers the kernel reads A from and writes U, sigma and V to.
        unsigned long long *mbox;
        int64_t mbox_group;
        const double *in_a;
        double *out_u, *out_s, *out_v;
        // Ring form: number of groups; with xcd_spread = X > 0 the grid holds
        // X * ceil(groups / X) * members workgroups, workgroup b runs on XCD
        // b % X (the dispatcher deals workgroups out to the XCDs in turn),
        // group q's members sit on XCD q % X and slots without a group exit
        // at once; xcc gets one word per member (its XCC id).
        int ring_groups;
        int xcd_spread;
        unsigned int *xcc;
        // Ring form: two sets of LDS slots used by round parity (the launch
        // then takes twice ring_lds_bytes), so a round needs one workgroup
        // barrier instead of two.
        int ring_db;
    };
    int wave_sweeps_max_rows();
    // Ring form of the grouped solves (one wave per column pair, every
    // pair's columns held in registers for the whole solve; columns move to
    // the neighbouring pair positions through LDS inside a workgroup and
    // through tagged mailbox words between neighbouring workgroups, so a
    // round has no group barrier).  A launch whose largest matrix has m rows
    // and n columns (every matrix m >= n) fits when m <= 512, or m <= 1024
    // and n <= 512; its groups have ceil(half / ring_waves(m, n)) workgroups.
    bool ring_fits(int64_t m, int64_t n);
    int ring_waves(int64_t m, int64_t n);
    int ring_lds_bytes(int64_t m, int64_t n);
    int64_t ring_mailbox_words(int64_t m, int64_t n, int members);
    const void *wave_ring_sweeps_f64_symbol(int64_t m, int64_t n);
    hipError_t launch_wave_ring_sweeps_f64(int64_t m, int64_t n, int groups, const WaveGroupArgs &args,
                                           hipStream_t stream);
    // Kernel instance for matrices of up to m rows (all matrices of one
    // launch share it: ceil(m / 64) row registers per lane) and workgroups
    // of `waves` (8 or 16) waves.
    const void *wave_group_sweeps_f64_symbol(int64_t m, int waves);
    hipError_t launch_wave_group_sweeps_f64(int64_t m, int waves, int groups, const WaveGroupArgs &args,
                                            hipStream_t stream);
    // The same solves, one workgroup of sixteen waves per matrix (no
    // cross-workgroup barrier; `members` and `bars` are unused): the
    // throughput form for many matrices.  Matrices of more than
    // wave_sweeps_max_rows() rows are accepted here (rows streamed; `m` is
    // then the launch's largest row count).
    hipError_t launch_wave_block_sweeps_f64(int64_t m, int grid, const WaveGroupArgs &args, hipStream_t stream);

    // Blocked Jacobi over one column-major matrix whose n columns are padded
    // with zero columns to nv = nb * block and whose m rows are padded with
    // zero rows to lda (a multiple of 64).  Per round: partial Grams of each
    // block pair over `chunks` row chunks of A, the Gram solve, and the
    // panel update of A and V.  Round 0 of a sweep rotates
    // every column pair of each block pair; later rounds rotate only the
    // pairs across the two blocks.
    struct BlockedArgs
    {
        int block;          // column block width: 16 or 32
        int nb;             // number of column blocks
        int64_t lda;        // leading dimension (padded rows) of A, multiple of 64
        int nv;             // padded column count = rows and columns of V
        double *a;          // lda x nv column-major
        double *v;          // nv x nv column-major
        int64_t chunk_rows; // rows per Gram chunk, multiple of 64
        int chunks;
        double eps;
        double *gram;          // pairs * chunks * (2 block)^2
        double *q;             // pairs * (2 block)^2
        int32_t *pair_rotated; // pairs
        // Sweep state, four ints: [0] raised when the running sweep rotated
        // a pair, [1] done (every later launch returns at once), [2]
        // converged, [3] sweeps executed.  Zero before the first sweep.
        int32_t *state;
    };
    hipError_t launch_blocked_round(const BlockedArgs &args, int round, hipStream_t stream);
    // Before sweep `sweep` of a blocked solve (closes the previous one).
    hipError_t launch_blocked_sweep_step(int32_t *state, int sweep, hipStream_t stream);
    // After the last enqueued sweep: sweeps and stop reason of the matrix.
    hipError_t launch_blocked_sweep_close(int32_t *state, int enqueued, int32_t *sweeps_out, int32_t *stop_out,
                                          hipStream_t stream);

    // out (cols x rows, row-major) = transpose of in (rows x cols, row-major);
    // when row_scale is given, input row r is divided by row_scale[r] (and
    // written as 0 when row_scale[r] <= eps).
    hipError_t launch_transpose_f64(const double *in, double *out, int64_t rows, int64_t cols,
                                    const double *row_scale, double eps, hipStream_t stream);
    // Same with explicit leading dimensions: in[r * ld_in + c] ->
    // out[c * ld_out + r].
    hipError_t launch_transpose_ld_f64(const double *in, int64_t ld_in, double *out, int64_t ld_out, int64_t rows,
                                       int64_t cols, const double *row_scale, double eps, hipStream_t stream);

    // sigma[j] = ||column j|| of a column-major m x n matrix whose columns
    // are lda elements apart.
    hipError_t launch_column_norms_f64(const double *a_cm, int64_t lda, double *sigma, int64_t m, int n,
                                       hipStream_t stream);

    // V <- I (n x n), flags[0..count) <- 0, sweeps slot <- 0.
    hipError_t launch_identity_f64(double *v, int n, hipStream_t stream);
    hipError_t launch_fill_i32(int32_t *p, int32_t value, int64_t count, hipStream_t stream);
    // count 64-bit words at p from network (big-endian) to host byte order,
    // in place (*.mat payloads staged as read from the file).
    hipError_t launch_bswap64(void *p, int64_t count, hipStream_t stream);
} // namespace jsvd
