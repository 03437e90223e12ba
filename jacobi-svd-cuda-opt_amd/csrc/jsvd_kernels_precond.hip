// jsvd_kernels_precond.hip -- the QR preconditioner of the mixed-precision
// path (no reference counterpart: the reference is plain FP64 Jacobi).
//
// One-sided Jacobi applied to X = R^T, where A P = Q1 R is a QR
// factorisation, converges in about half the sweeps it needs on A, and a
// second QR, R^T = Q2 R2, with Jacobi on X = R2^T saves about a quarter of
// the rest (Drmac and Veselic, "New fast and accurate Jacobi SVD algorithm",
// SIAM J. Matrix Anal. Appl. 29 (2008)).  Measured with the reference's pair
// order on the 128 x 128 matrices of kappa 1e3..1e8 at a cosine threshold of
// 1e-7 (the FP32 stage's): 15-27 sweeps on A, 10-11 on R^T, 7-8 on R2^T;
// column pivoting in the first QR saves under half a sweep more, so the QR
// here does not pivot (P = I, the permutation output is the identity).  With
// X = Ux S Vx^T,
//     A = Q1 R P^T = Q1 R2^T Q2^T P^T = (Q1 Ux) S (P Q2 Vx)^T,
// so U = Q1 Ux and V = P Q2 Vx.  Kernels, one workgroup per matrix each:
//  * qr_blocked_f64 / finish_blocked_f64 (65 to 128 columns, at most 128
//    rows): blocked QR and back-transform on FP64 MFMA (see their section
//    below).
//  * polar_bq_f64: the polar factor Q of the FP32 stage's V, B = X Q and two
//    first-order simultaneous Jacobi steps on B (see the kernel).
//  * qr_f64: Householder QR in FP64, every column in the registers of a lane
//    group, two workgroup barriers per step (see the kernel).  Writes
//    R^T (row-major n x n), the reflectors (column-major m x n, explicit
//    zeros above the unit diagonal), tau and the permutation.  It runs on A,
//    then in place on X = R^T (every column is read before any is written).
//  * precond_finish_f64: the reflectors of Q1 in LDS; lane groups of eight
//    carry one column of [Ux; 0] in registers through H_0 ... H_{n-1}
//    (U = Q1 Ux) and apply the U rule (zero columns where sigma <= eps);
//    then the reflectors of Q2 carry the columns of Vx (V = P Q2 Vx).
#include <utility>

#include "jsvd_device.hpp"
#include "jsvd_kernels.hpp"

namespace jsvd
{
    namespace
    {
        constexpr int kFinishBlock = 512;
        constexpr int kFinishWaves = 8; // waves of the blocked back-transform
        constexpr int kFinishGroup = 8;
        constexpr int kFinishRows = 32; // most rows per lane: m <= 8 * 32 = 256

        template <int G>
        __device__ __forceinline__ double lane_group_sum(double x)
        {
            return dev::group_sum_dpp<G>(x);
        }

        __device__ __forceinline__ void claim(unsigned int *counter, int *slot)
        {
            if (threadIdx.x == 0)
                *slot = (int)atomicAdd(counter, 1u);
            __syncthreads();
        }

        __device__ __forceinline__ void finish_list(unsigned int *counters)
        {
            if (threadIdx.x == 0)
            {
                const unsigned int done = atomicAdd(&counters[1], 1u);
                if (done == gridDim.x - 1)
                {
                    __hip_atomic_store(&counters[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&counters[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }

        // Householder QR with every column held in registers by a group of G
        // lanes (lane l holds rows l + G t, t < RW).  Step k: every wave
        // derives the reflector's tau from column k's row-k value and squared
        // norm below it, which the previous step's update left in LDS; column
        // k's group writes the reflector to LDS; after one barrier every later
        // column's group applies it and leaves its new row-(k+1) value and
        // squared norm below it; a second barrier ends the step.
        template <int BS, int G, int RW>
        __global__ __launch_bounds__(BS) void qr_f64(PrecondArgs args)
        {
            constexpr int GROUPS = BS / G;
            extern __shared__ double2 lds_raw[];
            double *vcol = reinterpret_cast<double *>(lds_raw); // reflector of the step, m values
            __shared__ double s_head[GROUPS], s_tail[GROUPS], s_tau[GROUPS];
            __shared__ int s_item;
            const int tid = threadIdx.x, gid = tid / G, lane = tid & (G - 1);
            for (;;)
            {
                claim(&args.counters[0], &s_item);
                const int item = s_item;
                if (item >= args.count)
                    break;
                const WorkItem w = args.items[item];
                const int m = w.m, n = w.n;
                const bool live = gid < n;
                const double *ag = args.a + w.a_off;
                double col[RW];
#pragma unroll
                for (int t = 0; t < RW; ++t)
                {
                    const int i = lane + G * t;
                    col[t] = (live && i < m) ? ag[int64_t(i) * n + gid] : 0.0;
                }
                {
                    double h = 0.0, tl = 0.0;
#pragma unroll
                    for (int t = 0; t < RW; ++t)
                    {
                        const int i = lane + G * t;
                        if (i == 0)
                            h = col[t];
                        else
                            tl = fma(col[t], col[t], tl);
                    }
                    dev::group_sum2_dpp<G>(h, tl);
                    if (live && lane == 0)
                    {
                        s_head[gid] = h;
                        s_tail[gid] = tl;
                    }
                }
                __syncthreads();
                for (int k = 0; k < n; ++k)
                {
                    const double alpha = s_head[k], x2 = s_tail[k];
                    // LAPACK dlarfg: beta = -sign(alpha) ||(alpha, x)||,
                    // tau = (beta - alpha) / beta, v = x / (alpha - beta).
                    double tau = 0.0, beta = alpha, scale = 0.0;
                    if (x2 != 0.0)
                    {
                        beta = -copysign(sqrt(fma(alpha, alpha, x2)), alpha);
                        tau = (beta - alpha) / beta;
                        scale = 1.0 / (alpha - beta);
                    }
                    // The reflector covers all G RW rows: zeros above row k
                    // and in the padding rows (whose column values are zero),
                    // so the loops below need no row tests (which would turn
                    // every LDS access into a branch).
                    if (gid == k)
                    {
#pragma unroll
                        for (int t = 0; t < RW; ++t)
                        {
                            const int i = lane + G * t;
                            const double v = i < k ? 0.0 : (i == k ? 1.0 : col[t] * scale);
                            vcol[i] = v;
                            col[t] = i < k ? col[t] : (i == k ? beta : v);
                        }
                        if (lane == 0)
                            s_tau[k] = tau;
                    }
                    __syncthreads();
                    if (live && gid > k)
                    {
                        double d0 = 0.0, d1 = 0.0;
#pragma unroll
                        for (int t = 0; t < RW; ++t)
                        {
                            if (t & 1)
                                d1 = fma(vcol[lane + G * t], col[t], d1);
                            else
                                d0 = fma(vcol[lane + G * t], col[t], d0);
                        }
                        const double d = lane_group_sum<G>(d0 + d1) * tau;
                        double hh = 0.0, tt = 0.0;
#pragma unroll
                        for (int t = 0; t < RW; ++t)
                        {
                            const int i = lane + G * t;
                            const double x = fma(-d, vcol[i], col[t]);
                            col[t] = x;
                            hh = i == k + 1 ? x : hh;
                            tt = i > k + 1 ? fma(x, x, tt) : tt;
                        }
                        dev::group_sum2_dpp<G>(hh, tt);
                        if (lane == 0)
                        {
                            s_head[gid] = hh;
                            s_tail[gid] = tt;
                        }
                    }
                    __syncthreads();
                }
                // Outputs of column j = gid: row j of X = R^T (R's column j:
                // rows 0 .. j), the reflector (zeros above, 1 on the diagonal)
                // and tau; the column order is A's own (identity permutation).
                if (live)
                {
                    const int j = gid;
                    double *xr = args.x + w.v_off + int64_t(j) * n;
                    double *rc = args.refl + w.a_off + int64_t(j) * m;
#pragma unroll
                    for (int t = 0; t < RW; ++t)
                    {
                        const int i = lane + G * t;
                        if (i < n)
                            xr[i] = i <= j ? col[t] : 0.0;
                        if (i < m)
                            rc[i] = i < j ? 0.0 : (i == j ? 1.0 : col[t]);
                    }
                    if (lane == 0)
                    {
                        if (args.perm)
                            args.perm[w.s_off + j] = j;
                        args.tau[w.s_off + j] = s_tau[j];
                    }
                }
                __syncthreads();
            }
            finish_list(args.counters);
        }

        // The reflectors H_k (unit diagonal, column-major with `rows` rows at
        // rg) and their tau into LDS: column k at Hs + k LD with LD = G R >=
        // rows, zero padding rows (branch-free loops in apply_reflectors).
        template <int R>
        __device__ __forceinline__ void load_reflectors(double *Hs, double *taus, const double *rg, const double *tg,
                                                        int rows, int n)
        {
            constexpr int LD = kFinishGroup * R;
            for (int idx = threadIdx.x; idx < LD * n; idx += kFinishBlock)
            {
                const int k = idx / LD, i = idx - k * LD;
                Hs[idx] = i < rows ? rg[k * rows + i] : 0.0;
            }
            for (int j = threadIdx.x; j < n; j += kFinishBlock)
                taus[j] = tg[j];
        }

        // dst = H_0 H_1 ... H_{n-1} [src; 0] for the row-major n x n src,
        // one column per lane group (lane l holds rows l + G t), written
        // row-major `rows` x n.  U (PERM false): row i to row i, columns j
        // with sigma[j] <= eps as zeros (the reference's U rule).  V (PERM
        // true): row i to row row_of[i], every column kept.
        template <int R, bool PERM>
        __device__ __forceinline__ void apply_reflectors(const double *Hs, const double *taus, const double *src,
                                                         double *dst, int rows, int n, const int32_t *row_of,
                                                         const double *sigma, double eps)
        {
            constexpr int G = kFinishGroup, GROUPS = kFinishBlock / G, LD = G * R;
            const int gid = threadIdx.x / G, lane = threadIdx.x & (G - 1);
            for (int j = gid; j < n; j += GROUPS)
            {
                double y[R];
#pragma unroll
                for (int t = 0; t < R; ++t)
                {
                    const int i = lane + G * t;
                    y[t] = i < n ? src[i * n + j] : 0.0;
                }
                for (int k = n - 1; k >= 0; --k)
                {
                    const double *h = Hs + k * LD;
                    // Four partial sums: a single chain would wait on every
                    // FMA's latency.
                    double dd[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int t = 0; t < R; ++t)
                        dd[t & 3] = fma(h[lane + G * t], y[t], dd[t & 3]);
                    const double d = lane_group_sum<G>((dd[0] + dd[1]) + (dd[2] + dd[3])) * taus[k];
#pragma unroll
                    for (int t = 0; t < R; ++t)
                        y[t] = fma(-d, h[lane + G * t], y[t]);
                }
                const bool keep = PERM || sigma[j] > eps;
#pragma unroll
                for (int t = 0; t < R; ++t)
                {
                    const int i = lane + G * t;
                    if (i < rows)
                        dst[(PERM ? row_of[i] : i) * n + j] = keep ? y[t] : 0.0;
                }
            }
        }

        template <int R>
        __global__ __launch_bounds__(kFinishBlock) void precond_finish_f64(PrecondArgs args)
        {
            extern __shared__ double2 lds_raw[];
            double *Hs = reinterpret_cast<double *>(lds_raw);
            __shared__ int s_item;
            for (;;)
            {
                claim(&args.counters[2], &s_item);
                const int item = s_item;
                if (item >= args.count)
                    break;
                const WorkItem w = args.items[item];
                const int m = w.m, n = w.n;
                double *taus = Hs + kFinishGroup * R * n;
                // U = Q1 [Ux; 0] with the U rule.
                load_reflectors<R>(Hs, taus, args.refl + w.a_off, args.tau + w.s_off, m, n);
                __syncthreads();
                apply_reflectors<R, false>(Hs, taus, args.ux + w.v_off, args.u + w.a_off, m, n, nullptr,
                                    args.sigma + w.s_off, args.eps);
                __syncthreads();
                // V = P Q2 Vx: row i of Q2 Vx is row perm[i] of V.  Vx is
                // orthogonal (accumulated rotations), so V is orthogonal for
                // every input, rank-deficient ones included.
                load_reflectors<R>(Hs, taus, args.refl2 + w.v_off, args.tau2 + w.s_off, n, n);
                __syncthreads();
                apply_reflectors<R, true>(Hs, taus, args.vx + w.v_off, args.v + w.v_off, n, n, args.perm + w.s_off,
                                    nullptr, 0.0);
                __syncthreads();
            }
            finish_list(args.counters + 2);
        }

        typedef double f64x4 __attribute__((ext_vector_type(4)));

        // Waves of the polar kernel: two row halves by four column quarters.
        constexpr int kPolarWaves = 8;
        // First-order simultaneous Jacobi steps after B = X Q.
        constexpr int kFirstOrderSteps = 2;

        // One n x n product C = A B (n <= NT) on FP64 MFMA
        // (v_mfma_f64_16x16x4f64) by one workgroup of eight waves: wave w owns
        // the (NT/2) x (NT/4) block (w / 4, w % 4) of C as TX x TY tiles of
        // 16 x 16 (TX = NT / 32, TY = NT / 64), accumulated in `acc` (two
        // waves per SIMD, each with half the accumulators of a quadrant, so
        // they stay in registers).  K runs in blocks of
        // 16; in step s of a block, lane l supplies A[i0 + l%16][k_l] and
        // B[k_l][j0 + l%16] with k_l = k0 + 4 (l / 16) + s (any assignment of
        // the block's k to the four lane groups sums the same products, and
        // this one gives each lane four consecutive k, so row-major A and
        // column-major B operands are read as 32 contiguous bytes).  a(i, k)
        // and b(k, j) return the operands (zero outside n x n); lane l then
        // holds C[i0 + l/16 + 4 r][j0 + l%16] in acc[x][y][r].
        // v itself, as a value the compiler cannot see through: lane indices
        // laundered at each product are not hoisted (with everything derived
        // from them) to the top of the kernel, where they would hold
        // registers through every other product.
        __device__ __forceinline__ int opaque_int(int v)
        {
            asm volatile("" : "+v"(v));
            return v;
        }

        template <int NT, typename FA, typename FB>
        __device__ __forceinline__ void mfma_nn(f64x4 (&acc)[NT / 32][NT / 64], FA a, FB b, int n, int wr, int wc,
                                                int lr, int lk)
        {
            constexpr int TX = NT / 32, TY = NT / 64;
            wr = opaque_int(wr);
            wc = opaque_int(wc);
            lr = opaque_int(lr);
            lk = opaque_int(lk);
#pragma unroll
            for (int x = 0; x < TX; ++x)
#pragma unroll
                for (int y = 0; y < TY; ++y)
                    acc[x][y] = f64x4{0.0, 0.0, 0.0, 0.0};
            for (int k0 = 0; k0 < n; k0 += 16)
            {
                double av[TX][4], bv[TY][4];
#pragma unroll
                for (int s = 0; s < 4; ++s)
                {
                    const int k = k0 + 4 * lk + s;
#pragma unroll
                    for (int t = 0; t < TX; ++t)
                        av[t][s] = a(wr + 16 * t + lr, k);
#pragma unroll
                    for (int t = 0; t < TY; ++t)
                        bv[t][s] = b(k, wc + 16 * t + lr);
                }
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int x = 0; x < TX; ++x)
#pragma unroll
                        for (int y = 0; y < TY; ++y)
                            acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x][s], bv[y][s], acc[x][y], 0, 0, 0);
            }
        }

        // Q <- orthonormal factor of V32 (column-major n x n at q), then
        // B = X Q (X row-major at x, B column-major at b); one workgroup of
        // four waves per matrix, every product on FP64 MFMA.  With
        // E = Q^T Q - I (V32 is orthogonal to about 1e-6),
        //   Q (Q^T Q)^(-1/2) = Q (I - E/2 + 3 E^2 / 8 - ...),
        // whose truncation after E^2 leaves about (5/16) |E|^3 in Q^T Q - I:
        // one pass forms E, F = E E and Q <- Q M with M = I - E/2 + 3 F / 8
        // (three products instead of the four of two Newton-Schulz steps),
        // and a second pass runs only when max |E| > 2^-14 (truncation then
        // above about 7e-14; the FP32 sweeps leave |E| of 1e-6 .. 1e-4).  The
        // NT x (NT + 1) LDS buffer (odd stride: rows on distinct banks) holds
        // E, then M, then the new Q, each the B operand of the next product;
        // Q and X are read from L2.  Then two simultaneous first-order Jacobi
        // steps rotate B and Q by W (five more products each, described at the
        // step); X, no longer needed, holds one of its operands.
        template <int NT>
        __global__ __launch_bounds__(64 * kPolarWaves) void polar_bq_f64(PrecondArgs pa, MixedArgs ma)
        {
            constexpr int TX = NT / 32, TY = NT / 64, LD = NT + 1, BS = 64 * kPolarWaves;
            extern __shared__ double2 polar_raw[];
            double *S = reinterpret_cast<double *>(polar_raw);
            __shared__ double s_max[kPolarWaves];
            __shared__ int s_item;
            const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
            int lr = lane & 15, lk = lane >> 4;
            int wr = (wave >> 2) * (NT / 2), wc = (wave & 3) * (NT / 4);
            auto fresh = [&]() {
                wr = opaque_int(wr);
                wc = opaque_int(wc);
                lr = opaque_int(lr);
                lk = opaque_int(lk);
            };
            f64x4 acc[TX][TY];
            for (;;)
            {
                claim(&ma.counters[2], &s_item);
                const int item = s_item;
                if (item >= ma.count)
                    break;
                const WorkItem w = ma.items[item]; // square entry: m = n, a_off = v_off
                const int n = w.n;
                double *q = ma.q + w.v_off;
                const double *x = pa.x + w.v_off;
                double *b = ma.u + w.a_off;
                // Column-major Q: Q[r][c] = q[c n + r].
                auto qt = [=](int i, int k) { return (i < n && k < n) ? q[i * n + k] : 0.0; }; // (Q^T)[i][k]
                auto qc = [=](int k, int j) { return (k < n && j < n) ? q[j * n + k] : 0.0; }; // Q[k][j]
                auto qr = [=](int i, int k) { return (i < n && k < n) ? q[k * n + i] : 0.0; }; // Q[i][k]
                auto sm = [=](int k, int j) { return S[k * LD + j]; };                        // LDS operand
                for (int pass = 0; pass < 2; ++pass)
                {
                    // E = Q^T Q - I into LDS (zero outside n x n), and max |E|.
                    mfma_nn<NT>(acc, qt, qc, n, wr, wc, lr, lk);
                    fresh();
                    double emax = 0.0;
#pragma unroll
                    for (int xx = 0; xx < TX; ++xx)
#pragma unroll
                        for (int yy = 0; yy < TY; ++yy)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                            {
                                const int i = wr + 16 * xx + lk + 4 * r, j = wc + 16 * yy + lr;
                                const double e = (i < n && j < n) ? acc[xx][yy][r] - (i == j ? 1.0 : 0.0) : 0.0;
                                S[i * LD + j] = e;
                                emax = fmax(emax, fabs(e));
                            }
                    for (int o = 32; o > 0; o >>= 1)
                        emax = fmax(emax, __shfl_xor(emax, o, 64));
                    if (lane == 0)
                        s_max[wave] = emax;
                    __syncthreads();
                    double e_all = s_max[0];
#pragma unroll
                    for (int k = 1; k < kPolarWaves; ++k)
                        e_all = fmax(e_all, s_max[k]);
                    // F = E E (both operands from LDS), then this wave's part of
                    // M = I - E/2 + 3 F / 8 in registers.
                    mfma_nn<NT>(acc, sm, sm, n, wr, wc, lr, lk);
                    fresh();
#pragma unroll
                    for (int xx = 0; xx < TX; ++xx)
#pragma unroll
                        for (int yy = 0; yy < TY; ++yy)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                            {
                                const int i = wr + 16 * xx + lk + 4 * r, j = wc + 16 * yy + lr;
                                const double e = S[i * LD + j];
                                acc[xx][yy][r] = fma(0.375, acc[xx][yy][r], fma(-0.5, e, i == j ? 1.0 : 0.0));
                            }
                    __syncthreads(); // every wave has read E
                    fresh();
#pragma unroll
                    for (int xx = 0; xx < TX; ++xx)
#pragma unroll
                        for (int yy = 0; yy < TY; ++yy)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                            {
                                const int i = wr + 16 * xx + lk + 4 * r, j = wc + 16 * yy + lr;
                                S[i * LD + j] = (i < n && j < n) ? acc[xx][yy][r] : 0.0;
                            }
                    __syncthreads();
                    // Q <- Q M; the new Q goes to global memory (the refinement's
                    // starting V) and to LDS (the operand of B = X Q).
                    mfma_nn<NT>(acc, qr, sm, n, wr, wc, lr, lk);
                    __syncthreads(); // every wave has read Q and M
                    fresh();
#pragma unroll
                    for (int xx = 0; xx < TX; ++xx)
#pragma unroll
                        for (int yy = 0; yy < TY; ++yy)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                            {
                                const int i = wr + 16 * xx + lk + 4 * r, j = wc + 16 * yy + lr;
                                const bool in = i < n && j < n;
                                S[i * LD + j] = in ? acc[xx][yy][r] : 0.0;
                                if (in)
                                    q[j * n + i] = acc[xx][yy][r];
                            }
                    // The second pass reads the new Q from global memory: its
                    // stores must be visible to the whole workgroup.
                    __threadfence_block();
                    __syncthreads();
                    if (!(e_all > 0x1p-14))
                        break;
                }
                // B = X Q (X row-major, B column-major).
                mfma_nn<NT>(acc, [=](int i, int k) { return (i < n && k < n) ? x[i * n + k] : 0.0; }, sm, n, wr, wc,
                            lr, lk);
                auto each = [&](auto f) {
                    fresh();
#pragma unroll
                    for (int xx = 0; xx < TX; ++xx)
#pragma unroll
                        for (int yy = 0; yy < TY; ++yy)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                f(acc[xx][yy][r], wr + 16 * xx + lk + 4 * r, wc + 16 * yy + lr);
                };
                each([&](double v, int i, int j) {
                    if (i < n && j < n)
                        b[j * n + i] = v;
                });
                __threadfence_block();
                __syncthreads();
                // A simultaneous first-order Jacobi step on B: with
                // G = B^T B, the rotation generator E_ij = G_ij / (G_jj - G_ii)
                // (antisymmetric, formed from G's upper triangle) and
                // W = I + E + E^2/2 + E^3/6 (orthogonal to O(E^4)), B <- B W
                // and Q <- Q W take the columns' cosines from the FP32
                // stage's ~5e-6 to ~1e-10 and a second step below the FP64
                // threshold, which leaves the FP64 refinement (~3.6 sweeps
                // without the steps) its closing sweep without rotations on
                // most matrices.  Entries above 2^-14 (close
                // singular values) are left to the sweeps.
                for (int step = 0; step < kFirstOrderSteps; ++step)
                {
                    auto bt = [=](int i, int k) { return (i < n && k < n) ? b[i * n + k] : 0.0; }; // (B^T)[i][k]
                    auto bc = [=](int k, int j) { return (k < n && j < n) ? b[j * n + k] : 0.0; }; // B[k][j]
                    mfma_nn<NT>(acc, bt, bc, n, wr, wc, lr, lk);
                    each([&](double v, int i, int j) { S[i * LD + j] = (i < n && j < n) ? v : 0.0; });
                    __syncthreads();
                    // E in place: the thread of pair i < j reads G_ij, G_ii, G_jj
                    // and writes E_ij and E_ji = -E_ij; the diagonal (read by
                    // every pair) is cleared after a barrier.
                    for (int idx = tid; idx < n * n; idx += BS)
                    {
                        const int i = idx / n, j = idx - i * n;
                        if (i < j)
                        {
                            const double g = S[i * LD + j], den = S[j * LD + j] - S[i * LD + i];
                            const double e = (den != 0.0 && fabs(g) * 16384.0 <= fabs(den)) ? g / den : 0.0;
                            S[i * LD + j] = e;
                            S[j * LD + i] = -e;
                        }
                    }
                    __syncthreads();
                    for (int i = tid; i < n; i += BS)
                        S[i * LD + i] = 0.0;
                    __syncthreads();
                    // P1 = E (I + E/3); X (no longer needed) takes I + P1/2.
                    mfma_nn<NT>(acc, sm,
                                [=](int k, int j) { return S[k * LD + j] * (1.0 / 3.0) + (k == j ? 1.0 : 0.0); }, n, wr,
                                wc, lr, lk);
                    double *xs = pa.x + w.v_off;
                    each([&](double v, int i, int j) {
                        if (i < n && j < n)
                            xs[i * n + j] = fma(0.5, v, i == j ? 1.0 : 0.0);
                    });
                    __threadfence_block();
                    __syncthreads();
                    // W = I + E (I + P1/2), into LDS once every wave has read E.
                    mfma_nn<NT>(acc, sm, [=](int k, int j) { return (k < n && j < n) ? xs[k * n + j] : 0.0; }, n, wr,
                                wc, lr, lk);
                    __syncthreads();
                    each([&](double v, int i, int j) { S[i * LD + j] = (i < n && j < n) ? v + (i == j ? 1.0 : 0.0) : 0.0; });
                    __syncthreads();
                    // B <- B W and Q <- Q W, each stored once every wave has read it.
                    mfma_nn<NT>(acc, [=](int i, int k) { return (i < n && k < n) ? b[k * n + i] : 0.0; }, sm, n, wr, wc,
                                lr, lk);
                    __syncthreads();
                    each([&](double v, int i, int j) {
                        if (i < n && j < n)
                            b[j * n + i] = v;
                    });
                    mfma_nn<NT>(acc, qr, sm, n, wr, wc, lr, lk);
                    __syncthreads();
                    each([&](double v, int i, int j) {
                        if (i < n && j < n)
                            q[j * n + i] = v;
                    });
                    __threadfence_block();
                    __syncthreads(); // LDS and the global operands free for the next matrix
                }
            }
            if (threadIdx.x == 0)
            {
                const unsigned int done = atomicAdd(&ma.counters[3], 1u);
                if (done == gridDim.x - 1)
                {
                    __hip_atomic_store(&ma.counters[2], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&ma.counters[3], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }

        // ------------------------------------------------------------------
        // Blocked Householder QR and back-transform on FP64 MFMA, one
        // workgroup of four waves per matrix, rows padded with zeros to
        // 16 MT, columns in blocks of 16.
        //
        // The matrix lives in registers as v_mfma_f64_16x16x4f64 result
        // tiles: column block b (columns 16 b .. 16 b + 15) belongs to wave
        // b % 4 as its slot b / 4, and lane l = 16 q + c holds rows
        // 16 t + q + 4 r of column 16 b + c as element r of tile t.  The
        // product's operands are A[i][k] from lane 16 k + i and B[k][j] from
        // lane 16 k + j, so element r of a tile is also the B operand of the
        // K-step over rows 16 t + 4 r .. 16 t + 4 r + 3: a product V^T Y reads
        // Y straight from the registers that hold it.
        //
        // QR (LAPACK dgeqrf with dlarft's compact WY form): one wave factors
        // panel p (the 16 columns of block p, rows 16 p ..) in registers,
        // Householder step by step (dlarfg's beta, tau and v), forming T from
        // the reflectors' inner products as it goes; then every wave applies
        // Q_p^T = I - V T^T V^T to its blocks: W = V^T C, W <- T^T W,
        // C <- C - V W, all on MFMA with V and T in LDS, and the owner of
        // panel p + 1 factors it.  V and T are double-buffered, so one
        // workgroup barrier per panel suffices.
        template <int MT>
        struct BlkShape
        {
            static constexpr int ROWS = 16 * MT; // padded rows
            static constexpr int LDV = ROWS + 4; // LDS stride of a V column (spreads the 16 columns over banks)
            static constexpr int S = 4 * MT;     // rows per lane of a column
        };

        __device__ __forceinline__ double readlane_f64(double x, int l)
        {
            const long long bits = __double_as_longlong(x);
            const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits & 0xffffffffll), l);
            const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), l);
            return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                        static_cast<long long>(static_cast<unsigned int>(lo)));
        }

        __device__ __forceinline__ void wave_lds_sync()
        {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }

        // Y <- (I - V T V^T) Y (TRANS false) or (I - V T^T V^T) Y (TRANS
        // true) for one column block y of MT tiles.  V: LDS, column-major with
        // stride LDV, rows from 16 t0 on (V's rows above are zero); T: LDS,
        // 16 x 16 row-major.  The instance for T0, the even tile count at or
        // below t0, leaves the tiles below T0 out; the tile between T0 and t0
        // (if any) takes zero operands, selected rather than branched around
        // so that no tile's registers need copies.
        template <int MT, bool TRANS, int T0 = 0>
        __device__ __forceinline__ void apply_block_reflector(f64x4 (&y)[MT], const double *V, const double *T, int t0,
                                                              int c, int q)
        {
            if constexpr (T0 + 2 < MT)
            {
                if (t0 >= T0 + 2)
                {
                    apply_block_reflector<MT, TRANS, T0 + 2>(y, V, T, t0, c, q);
                    return;
                }
            }
            constexpr int LDV = BlkShape<MT>::LDV;
            // W = V^T Y: A[a][row] = V[row][a] from lane (c = a, q = row % 4),
            // one accumulator per K-step r (four independent chains), at most
            // one tile's operands loaded ahead.
            f64x4 wr[4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
                wr[r] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int t = T0; t < MT; ++t)
            {
                const bool on = t >= t0;
                const int tt = on ? t - t0 : 0;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                {
                    const double a = V[c * LDV + 16 * tt + 4 * r + q];
                    wr[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(on ? a : 0.0, y[t][r], wr[r], 0, 0, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            const f64x4 w = (wr[0] + wr[1]) + (wr[2] + wr[3]);
            // W2 = T^T W or T W: A[a][k] = T[k][a] or T[a][k], B = W's elements.
            f64x4 w2 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int r = 0; r < 4; ++r)
                w2 = __builtin_amdgcn_mfma_f64_16x16x4f64(TRANS ? T[(4 * r + q) * 16 + c] : T[c * 16 + 4 * r + q], w[r],
                                                          w2, 0, 0, 0);
            // Y <- Y - V W2: A[i][k] = -V[16 t + i][4 r + k]; the tiles'
            // products are independent, one K-step of all tiles at a time.
#pragma unroll
            for (int r = 0; r < 4; ++r)
            {
#pragma unroll
                for (int t = T0; t < MT; ++t)
                {
                    const bool on = t >= t0;
                    const int tt = on ? t - t0 : 0;
                    const double a = V[(4 * r + q) * LDV + 16 * tt + c];
                    y[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(on ? -a : 0.0, w2[r], y[t], 0, 0, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }

        // Lane J of each row of 16 lanes, to every lane of the row (DPP
        // row_newbcast, two 32-bit moves).
        template <int J>
        __device__ __forceinline__ double row_bcast(double x)
        {
            const long long bits = __double_as_longlong(x);
            const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(bits & 0xffffffffll), 0x150 + J, 0xF, 0xF, false);
            const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(bits >> 32), 0x150 + J, 0xF, 0xF, false);
            return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                        static_cast<long long>(static_cast<unsigned int>(lo)));
        }

        // Householder step J of a panel held by one wave: lane (c, q) holds
        // rows q + 4 s of panel column c in x[s]; trow is row c of T and
        // tau_c the tau of column c.  Every lane forms the reflector vector
        // of its own column and row J's lane broadcasts column J's to its
        // row of 16 lanes (the lanes of one q hold the same rows), so the
        // step has no divergent branch and no LDS round trip.
        template <int J, int S>
        __device__ __forceinline__ void qr_step(double (&x)[S], double (&trow)[16], double &tau_c, int c, int q)
        {
            constexpr int sj = J / 4, qj = J % 4;
            // Head (row J) and squared tail (rows below J) of each column.
            double h = q == qj ? x[sj] : 0.0;
            double tl0 = q > qj ? x[sj] * x[sj] : 0.0, tl1 = 0.0;
#pragma unroll
            for (int s = sj + 1; s < S; ++s)
            {
                if (s & 1)
                    tl1 = fma(x[s], x[s], tl1);
                else
                    tl0 = fma(x[s], x[s], tl0);
            }
            h = dev::xor32_sum(dev::xor16_sum(h));
            const double tl = dev::xor32_sum(dev::xor16_sum(tl0 + tl1));
            const double alpha = readlane_f64(h, J), x2 = readlane_f64(tl, J);
            // LAPACK dlarfg: beta = -sign(alpha) ||(alpha, x)||,
            // tau = (beta - alpha) / beta, v = x / (alpha - beta); tau = 0
            // (no reflection) when x = 0.
            // |beta| = sqrt(s2), s2 = alpha^2 + x2, from the refined 1/sqrt
            // estimate of s2 scaled by an even power of two into [0.5, 2);
            // then d = alpha - beta (no cancellation: alpha and -beta share
            // a sign), scale = 1/d from the refined reciprocal of d scaled
            // into [0.5, 1), and tau = -d / beta = |d| / |beta|.
            double beta = alpha, tj = 0.0, scale = 0.0;
            if (x2 != 0.0)
            {
                const double s2 = fma(alpha, alpha, x2);
                const int k = __builtin_amdgcn_frexp_exp(s2) & ~1;
                const double sc = __builtin_amdgcn_ldexp(s2, -k);
                const double y = dev::rsqrt_refined(sc);
                const double bmag = __builtin_amdgcn_ldexp(sc * y, k / 2);
                beta = -copysign(bmag, alpha);
                const double d = alpha - beta;
                const int kd = __builtin_amdgcn_frexp_exp(d);
                const double r = dev::divide_refined(1.0, fabs(__builtin_amdgcn_ldexp(d, -kd)));
                scale = copysign(__builtin_amdgcn_ldexp(r, -kd), d);
                tj = fabs(d) * __builtin_amdgcn_ldexp(y, -(k / 2));
            }
            // v (zeros above row J, 1 on it) of this lane's column, column J's
            // broadcast; column J keeps its R entries, takes beta on row J and
            // v below.
            double vb[S];
            {
                const int o = q + 4 * sj;
                const double xs = x[sj] * scale;
                const double v = o > J ? xs : (o == J ? 1.0 : 0.0);
                vb[sj] = row_bcast<J>(v);
                const double nx = o > J ? v : (o == J ? beta : x[sj]);
                x[sj] = c == J ? nx : x[sj];
            }
            // Registers below row J's: column J's entries become v = x scale,
            // the other columns' are multiplied by 1 (unchanged).
            const double g = c == J ? scale : 1.0;
#pragma unroll
            for (int s = sj + 1; s < S; ++s)
            {
                x[s] *= g;
                vb[s] = row_bcast<J>(x[s]);
            }
            double d0 = 0.0, d1 = 0.0;
#pragma unroll
            for (int s = sj; s < S; ++s)
            {
                if (s & 1)
                    d1 = fma(vb[s], x[s], d1);
                else
                    d0 = fma(vb[s], x[s], d0);
            }
            const double d = dev::xor32_sum(dev::xor16_sum(d0 + d1));
            // Column J of T (dlarft, forward, columnwise): T[c][J] =
            // -tau_J sum_l T[c][l] G[l][J] for c < J, with G[l][J] = v_l^T v_J
            // the d of column l (its rows from J on hold v_l); T[J][J] = tau_J.
            double acc = 0.0;
#pragma unroll
            for (int l = 0; l < J; ++l)
                acc = fma(trow[l], readlane_f64(d, l), acc);
            trow[J] = c < J ? -tj * acc : (c == J ? tj : 0.0);
            tau_c = c == J ? tj : tau_c;
            // Later columns: x <- x - tau v (v^T x).
            const double f = c > J ? tj * d : 0.0;
#pragma unroll
            for (int s = sj; s < S; ++s)
                x[s] = fma(-f, vb[s], x[s]);
        }

        template <int S, int... J>
        __device__ __forceinline__ void qr_steps(double (&x)[S], double (&trow)[16], double &tau_c, int c, int q,
                                                 std::integer_sequence<int, J...>)
        {
            ((qr_step<J, S>(x, trow, tau_c, c, q), __builtin_amdgcn_sched_barrier(0)), ...);
        }

        // Householder factor of a staged panel over its first SP registers
        // per lane (rows q + 4 s, s < SP, relative to 16 p; the rows past them
        // are zero and lie beyond the matrix, so nothing of them is written).
        template <int MT, int SP>
        __device__ __forceinline__ void factor_panel(int p, double *Vp, double *Tp, const PrecondArgs &args,
                                                     const WorkItem &w, int64_t t_off, int c, int q)
        {
            constexpr int LDV = BlkShape<MT>::LDV;
            const int m = w.m, n = w.n, col = 16 * p + c;
            const bool live = col < n;
            double *xg = args.x + w.v_off;
            double *rg = args.refl + w.a_off;
            double x[SP];
#pragma unroll
            for (int s = 0; s < SP; ++s)
                x[s] = Vp[c * LDV + q + 4 * s];
            double trow[16];
#pragma unroll
            for (int j = 0; j < 16; ++j)
                trow[j] = 0.0;
            double tau_c = 0.0;
            qr_steps<SP>(x, trow, tau_c, c, q, std::make_integer_sequence<int, 16>{});
            wave_lds_sync();
            // V for the other waves, T, and the outputs of column col: R's
            // rows 16 p .. as X = R^T (zeros below the diagonal), the
            // reflector (zeros above, unit diagonal), tau.
#pragma unroll
            for (int s = 0; s < SP; ++s)
            {
                const int o = q + 4 * s;
                Vp[c * LDV + o] = o < c ? 0.0 : (o == c ? 1.0 : x[s]);
            }
            if (q == 0)
            {
#pragma unroll
                for (int j = 0; j < 16; ++j)
                {
                    Tp[c * 16 + j] = trow[j];
                    args.tq[t_off + 256 * p + c * 16 + j] = trow[j];
                }
            }
            if (live)
            {
#pragma unroll
                for (int s = 0; s < SP; ++s)
                {
                    const int o = q + 4 * s, row = 16 * p + o;
                    if (row < n)
                        xg[int64_t(col) * n + row] = o <= c ? x[s] : 0.0;
                    if (row < m)
                        rg[int64_t(col) * m + row] = o < c ? 0.0 : (o == c ? 1.0 : x[s]);
                }
                if (q == 0)
                    args.tau[w.s_off + col] = tau_c;
            }
            wave_lds_sync();
        }

        // factor_panel over the smallest even tile count K >= tiles.
        template <int MT, int K>
        __device__ __forceinline__ void factor_panel_fit(int tiles, int p, double *Vp, double *Tp,
                                                         const PrecondArgs &args, const WorkItem &w, int64_t t_off,
                                                         int c, int q)
        {
            static_assert(K % 2 == 0, "even tile counts");
            if constexpr (K > 2)
            {
                if (tiles <= K - 2)
                {
                    factor_panel_fit<MT, K - 2>(tiles, p, Vp, Tp, args, w, t_off, c, q);
                    return;
                }
            }
            factor_panel<MT, 4 * K>(p, Vp, Tp, args, w, t_off, c, q);
        }

        // Panel p of the QR, by its owning wave: the block's tiles from t = p
        // on are staged in LDS (rows relative to 16 p, zeros beyond the padded
        // matrix) into registers x[s] (rows q + 4 s of column c), factored by
        // 16 Householder steps, and leave behind the explicit V (unit
        // diagonal, zeros above) in Vp, T in Tp and the outputs: R's rows (as
        // X = R^T), the reflectors, tau and T.  Tiles above p hold R rows that
        // earlier panels finished.
        template <int MT, int NBW>
        __device__ __forceinline__ void qr_panel(const f64x4 (&acc)[NBW][MT], int slot, int p, double *Vp, double *Tp,
                                                 const PrecondArgs &args, const WorkItem &w, int64_t t_off, int c,
                                                 int q)
        {
            constexpr int LDV = BlkShape<MT>::LDV;
            const int m = w.m, n = w.n, col = 16 * p + c;
            const bool live = col < n;
            double *xg = args.x + w.v_off;
            double *rg = args.refl + w.a_off;
            // Stage the block from tile p on; write the finished R rows above.
#pragma unroll
            for (int t = 0; t < MT; ++t)
            {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                {
                    const int row = 16 * t + q + 4 * r;
                    double val = acc[0][t][r];
                    if constexpr (NBW > 1)
                        val = slot == 1 ? acc[1][t][r] : val;
                    if (t >= p)
                        Vp[c * LDV + row - 16 * p] = val;
                    else if (live)
                    {
                        if (row < n)
                            xg[int64_t(col) * n + row] = val;
                        rg[int64_t(col) * m + row] = 0.0;
                    }
                }
                if (t < p)
                {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        Vp[c * LDV + 16 * (MT - 1 - t) + q + 4 * r] = 0.0; // rows past the padded matrix
                }
            }
            wave_lds_sync();
            // Rows from 16 ceil(min(m, rows) / 16) on are zero padding: the
            // factor holds the tiles left (rounded up to an even count).
            const int tiles = (min(m, BlkShape<MT>::ROWS) + 15) / 16 - p;
            factor_panel_fit<MT, MT>(tiles, p, Vp, Tp, args, w, t_off, c, q);
        }

        template <int MT, int NBW>
        __global__ __launch_bounds__(256) void qr_blocked_f64(PrecondArgs args)
        {
            constexpr int LDV = BlkShape<MT>::LDV;
            __shared__ double Vbuf[2][16 * LDV];
            __shared__ double Tbuf[2][256];
            __shared__ int s_item;
            const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
            const int c = lane & 15, q = lane >> 4;
            for (;;)
            {
                claim(&args.counters[0], &s_item);
                const int item = s_item;
                if (item >= args.count)
                    break;
                const WorkItem w = args.items[item];
                const int m = w.m, n = w.n, nb = (n + 15) / 16;
                const int64_t t_off = 16 * w.s_off + 240 * int64_t(w.index);
                const double *ag = args.a + w.a_off;
                f64x4 acc[NBW][MT];
#pragma unroll
                for (int sl = 0; sl < NBW; ++sl)
                {
                    const int col = 16 * (4 * sl + wave) + c;
#pragma unroll
                    for (int t = 0; t < MT; ++t)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                        {
                            const int row = 16 * t + q + 4 * r;
                            acc[sl][t][r] = (row < m && col < n) ? ag[int64_t(row) * n + col] : 0.0;
                        }
                }
                // Panel p + 1 is factored by its owner after every wave has
                // applied panel p to all its blocks (finished blocks are never
                // read again and padding columns stay zero, so every slot is
                // updated without a test).
                for (int p = -1; p + 1 < nb; ++p)
                {
                    if (p >= 0)
                    {
#pragma unroll
                        for (int sl = 0; sl < NBW; ++sl)
                            apply_block_reflector<MT, true>(acc[sl], Vbuf[p & 1], Tbuf[p & 1], p, c, q);
                    }
                    const int np = p + 1;
                    if (wave == (np & 3))
                        qr_panel<MT, NBW>(acc, np >> 2, np, Vbuf[np & 1], Tbuf[np & 1], args, w, t_off, c, q);
                    __syncthreads();
                }
            }
            finish_list(args.counters);
        }

        // Y <- Q Y for the Q = Q_0 Q_1 ... Q_{nb-1} of a blocked QR (panel p:
        // columns 16 p .. of the column-major `rows` x n reflectors at rg, T
        // at tg + 256 p), over the column blocks y of this wave; panels are
        // staged in LDS by the whole workgroup (BS threads), last panel first.
        template <int MT, int NBW, int BS>
        __device__ __forceinline__ void apply_q_blocked(f64x4 (&y)[NBW][MT], const double *rg, const double *tg,
                                                        int rows, int n, double (*Vbuf)[16 * BlkShape<MT>::LDV],
                                                        double (*Tbuf)[256], int wave, int c, int q)
        {
            constexpr int LDV = BlkShape<MT>::LDV, ROWS = BlkShape<MT>::ROWS;
            const int tid = threadIdx.x, nb = (n + 15) / 16;
            auto stage = [&](int p) {
                double *Vp = Vbuf[p & 1];
                for (int idx = tid; idx < 16 * ROWS; idx += BS)
                {
                    const int cc = idx / ROWS, o = idx - cc * ROWS, row = 16 * p + o, col = 16 * p + cc;
                    Vp[cc * LDV + o] = (row < rows && col < n) ? rg[int64_t(col) * rows + row] : 0.0;
                }
                for (int e = tid; e < 256; e += BS)
                    Tbuf[p & 1][e] = tg[256 * p + e];
            };
            stage(nb - 1);
            __syncthreads();
            for (int p = nb - 1; p >= 0; --p)
            {
                if (p > 0)
                    stage(p - 1);
                const double *V = Vbuf[p & 1];
#pragma unroll
                for (int sl = 0; sl < NBW; ++sl)
                    apply_block_reflector<MT, false>(y[sl], V, Tbuf[p & 1], p, c, q);
                __syncthreads();
            }
        }

        // Back-transform of a blocked-QR bucket: U = Q1 [Ux; 0] with the U
        // rule, then V = Q2 Vx (the QR does not pivot: P = I).  WAVES waves,
        // column block b in wave b % WAVES as its slot b / WAVES: eight waves
        // with one block each keep a wave's registers to one block, so two
        // matrices share a CU.
        template <int MT, int NBW, int WAVES>
        __global__ __launch_bounds__(64 * WAVES) void finish_blocked_f64(PrecondArgs args)
        {
            constexpr int LDV = BlkShape<MT>::LDV;
            __shared__ double Vbuf[2][16 * LDV];
            __shared__ double Tbuf[2][256];
            __shared__ int s_item;
            const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
            const int c = lane & 15, q = lane >> 4;
            for (;;)
            {
                claim(&args.counters[2], &s_item);
                const int item = __builtin_amdgcn_readfirstlane(s_item);
                if (item >= args.count)
                    break;
                const WorkItem w = args.items[item];
                const int m = w.m, n = w.n;
                const int64_t t_off = 16 * w.s_off + 240 * int64_t(w.index);
                f64x4 y[NBW][MT];
                for (int part = 0; part < 2; ++part)
                {
                    // part 0: Ux (n x n) -> U (m x n); part 1: Vx -> V (n x n).
                    const double *src = (part == 0 ? args.ux : args.vx) + w.v_off;
                    const int rows = part == 0 ? m : n;
                    // Row 16 t + 4 r + q of a lane: the tests compare q with a
                    // uniform bound, so no per-element row index is held.
#pragma unroll
                    for (int sl = 0; sl < NBW; ++sl)
                    {
                        const int col = 16 * (WAVES * sl + wave) + c;
                        const double *ps = src + q * n + col;
#pragma unroll
                        for (int t = 0; t < MT; ++t)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                            {
                                const int k = 16 * t + 4 * r;
                                y[sl][t][r] = (q < n - k && col < n) ? ps[k * n] : 0.0;
                            }
                    }
                    // One call site for both parts: the y blocks stay in one
                    // register set.
                    const double *rg = part == 0 ? args.refl + w.a_off : args.refl2 + w.v_off;
                    const double *tg = (part == 0 ? args.tq : args.tq2) + t_off;
                    apply_q_blocked<MT, NBW, 64 * WAVES>(y, rg, tg, rows, n, Vbuf, Tbuf, wave, c, q);
                    double *dst = (part == 0 ? args.u + w.a_off : args.v + w.v_off);
#pragma unroll
                    for (int sl = 0; sl < NBW; ++sl)
                    {
                        const int col = 16 * (WAVES * sl + wave) + c;
                        if (col < n)
                        {
                            const bool keep = part == 1 || args.sigma[w.s_off + col] > args.eps;
                            double *pd = dst + q * n + col;
#pragma unroll
                            for (int t = 0; t < MT; ++t)
#pragma unroll
                                for (int r = 0; r < 4; ++r)
                                {
                                    const int k = 16 * t + 4 * r;
                                    if (q < rows - k)
                                        pd[k * n] = keep ? y[sl][t][r] : 0.0;
                                }
                        }
                    }
                }
            }
            finish_list(args.counters + 2);
        }
    } // namespace

    namespace
    {
        // QR shapes; a matrix takes the first that fits.  Unblocked
        // (mt = 0): qr_f64<block, group, rows> for n <= block / group columns
        // and m <= group rows rows.  Blocked (mt > 0): qr_blocked_f64<mt, nbw>
        // for m <= 16 mt rows and n <= 64 nbw columns (so 65 to 128 columns
        // of at most 128 rows; taller matrices of at most 64 columns take the
        // last unblocked shape).  Up to 64 columns the
        // unblocked kernel is faster (measured on the 4096 uneven batch:
        // 2.97 vs 3.15 ms for the mixed execution, the blocked panels' 16
        // dependent steps then outweigh their cheaper updates); from 65 to
        // 128 columns the blocked one (1024 x 128x128: 10.0 vs 11.7 ms).
        struct QrSpec
        {
            int block, group, rows, mt, nbw;
        };
        constexpr QrSpec kQr[] = {{256, 4, 8, 0, 0}, {256, 4, 16, 0, 0}, {256, 4, 32, 0, 0},
                                  {256, 0, 0, 8, 2}, {512, 8, 32, 0, 0}};
        constexpr int kQrCount = sizeof(kQr) / sizeof(kQr[0]);
        template <int I>
        const void *qr_symbol_t()
        {
            if constexpr (kQr[I].mt > 0)
                return reinterpret_cast<const void *>(&qr_blocked_f64<kQr[I].mt, kQr[I].nbw>);
            else
                return reinterpret_cast<const void *>(&qr_f64<kQr[I].block, kQr[I].group, kQr[I].rows>);
        }
        template <int I>
        void qr_launch_t(int grid, int lds, const PrecondArgs &args, hipStream_t stream)
        {
            if constexpr (kQr[I].mt > 0)
                qr_blocked_f64<kQr[I].mt, kQr[I].nbw><<<grid, kQr[I].block, 0, stream>>>(args);
            else
                qr_f64<kQr[I].block, kQr[I].group, kQr[I].rows><<<grid, kQr[I].block, lds, stream>>>(args);
        }
        template <int I>
        const void *finish_blocked_symbol_t()
        {
            if constexpr (kQr[I].mt > 0)
                return reinterpret_cast<const void *>(
                    &finish_blocked_f64<kQr[I].mt, (4 * kQr[I].nbw + kFinishWaves - 1) / kFinishWaves, kFinishWaves>);
            else
                return nullptr;
        }
        template <int I>
        void finish_blocked_launch_t(int grid, const PrecondArgs &args, hipStream_t stream)
        {
            if constexpr (kQr[I].mt > 0)
                finish_blocked_f64<kQr[I].mt, (4 * kQr[I].nbw + kFinishWaves - 1) / kFinishWaves, kFinishWaves>
                    <<<grid, 64 * kFinishWaves, 0, stream>>>(args);
        }
        const void *(*const kQrSymbol[])() = {qr_symbol_t<0>, qr_symbol_t<1>, qr_symbol_t<2>, qr_symbol_t<3>,
                                             qr_symbol_t<4>};
        void (*const kQrLaunch[])(int, int, const PrecondArgs &, hipStream_t) = {
            qr_launch_t<0>, qr_launch_t<1>, qr_launch_t<2>, qr_launch_t<3>, qr_launch_t<4>};
        const void *(*const kFinishBlockedSymbol[])() = {finish_blocked_symbol_t<0>, finish_blocked_symbol_t<1>,
                                                         finish_blocked_symbol_t<2>, finish_blocked_symbol_t<3>,
                                                         finish_blocked_symbol_t<4>};
        void (*const kFinishBlockedLaunch[])(int, const PrecondArgs &, hipStream_t) = {
            finish_blocked_launch_t<0>, finish_blocked_launch_t<1>, finish_blocked_launch_t<2>,
            finish_blocked_launch_t<3>, finish_blocked_launch_t<4>};
        static_assert(sizeof(kQrLaunch) / sizeof(kQrLaunch[0]) == kQrCount, "one launcher per shape");
        static_assert(sizeof(kQrSymbol) / sizeof(kQrSymbol[0]) == kQrCount, "one symbol per shape");
        static_assert(sizeof(kFinishBlockedLaunch) / sizeof(kFinishBlockedLaunch[0]) == kQrCount,
                      "one back-transform launcher per shape");
    } // namespace

    int qrcp_variant(int64_t m, int64_t n)
    {
        for (int v = 0; v < kQrCount; ++v)
        {
            const QrSpec &k = kQr[v];
            const bool fits = k.mt > 0 ? (m <= 16 * k.mt && n <= 64 * k.nbw && n <= m)
                                       : (n <= k.block / k.group && m <= int64_t(k.group) * k.rows);
            if (fits)
                return v;
        }
        return -1;
    }

    bool qrcp_blocked(int variant)
    {
        return variant >= 0 && variant < kQrCount && kQr[variant].mt > 0;
    }

    int64_t qrcp_t_elems(int64_t s_elems, int64_t count)
    {
        return 16 * s_elems + 240 * count;
    }

    int64_t qrcp_lds_bytes(int variant)
    {
        // Unblocked: the reflector of a step over every register row
        // (G RW >= m).  Blocked: static LDS only.
        return kQr[variant].mt > 0 ? 0 : int64_t(kQr[variant].group) * kQr[variant].rows * 8;
    }

    namespace
    {
        // Rows per lane of the back-transform: 8, 16 or 32 (m <= 64, 128, 256).
        int finish_rows(int64_t m)
        {
            return m <= 8 * kFinishGroup ? 8 : (m <= 16 * kFinishGroup ? 16 : 32);
        }
    } // namespace

    int64_t precond_finish_lds_bytes(int variant, int64_t m, int64_t n)
    {
        // Blocked: static LDS only.  Unblocked: reflectors (LD = 8 x rows per
        // lane, zero padded) and tau; those of Q2 (n <= m rows) reuse the
        // space.
        if (qrcp_blocked(variant))
            return 0;
        return (int64_t(kFinishGroup) * finish_rows(m) * n + n) * 8;
    }

    bool precond_fits(int64_t m, int64_t n)
    {
        const int v = qrcp_variant(m, n);
        return n <= 128 && v >= 0 && m <= int64_t(kFinishGroup) * kFinishRows &&
               precond_finish_lds_bytes(v, m, n) <= kMaxLdsBytes;
    }

    const void *qrcp_symbol(int variant)
    {
        return variant >= 0 && variant < kQrCount ? kQrSymbol[variant]() : nullptr;
    }
    int qrcp_block(int variant)
    {
        return kQr[variant].block;
    }
    const void *precond_finish_symbol(int variant, int64_t m)
    {
        if (qrcp_blocked(variant))
            return kFinishBlockedSymbol[variant]();
        const int r = finish_rows(m);
        return r == 8    ? reinterpret_cast<const void *>(&precond_finish_f64<8>)
               : r == 16 ? reinterpret_cast<const void *>(&precond_finish_f64<16>)
                         : reinterpret_cast<const void *>(&precond_finish_f64<32>);
    }
    int precond_finish_block(int variant)
    {
        return qrcp_blocked(variant) ? 64 * kFinishWaves : kFinishBlock;
    }

    const void *polar_bq_symbol(int64_t n)
    {
        return n <= 64 ? reinterpret_cast<const void *>(&polar_bq_f64<64>)
                       : reinterpret_cast<const void *>(&polar_bq_f64<128>);
    }

    int polar_bq_lds_bytes(int64_t n)
    {
        const int nt = n <= 64 ? 64 : 128;
        return nt * (nt + 1) * 8;
    }

    int polar_bq_block()
    {
        return 64 * kPolarWaves;
    }

    hipError_t launch_polar_bq_f64(int64_t n, int grid, const PrecondArgs &pa, const MixedArgs &ma, hipStream_t stream)
    {
        if (n <= 64)
            polar_bq_f64<64><<<grid, polar_bq_block(), polar_bq_lds_bytes(n), stream>>>(pa, ma);
        else
            polar_bq_f64<128><<<grid, polar_bq_block(), polar_bq_lds_bytes(n), stream>>>(pa, ma);
        return hipGetLastError();
    }

    hipError_t launch_qrcp_f64(int variant, int grid, int lds_bytes, const PrecondArgs &args, hipStream_t stream)
    {
        if (variant < 0 || variant >= kQrCount)
            return hipErrorInvalidValue;
        kQrLaunch[variant](grid, lds_bytes, args, stream);
        return hipGetLastError();
    }

    hipError_t launch_precond_finish_f64(int variant, int64_t m, int grid, int lds_bytes, const PrecondArgs &args,
                                         hipStream_t stream)
    {
        if (qrcp_blocked(variant))
        {
            kFinishBlockedLaunch[variant](grid, args, stream);
            return hipGetLastError();
        }
        const int r = finish_rows(m);
        if (r == 8)
            precond_finish_f64<8><<<grid, kFinishBlock, lds_bytes, stream>>>(args);
        else if (r == 16)
            precond_finish_f64<16><<<grid, kFinishBlock, lds_bytes, stream>>>(args);
        else
            precond_finish_f64<32><<<grid, kFinishBlock, lds_bytes, stream>>>(args);
        return hipGetLastError();
    }
} // namespace jsvd
