// jsvd_kernels_precond.hip -- the QR preconditioner of the mixed-precision
// path (no reference counterpart: the reference is plain FP64 Jacobi).
//
// One-sided Jacobi applied to X = R^T, where A P = Q1 R is a QR
// factorisation, converges in about half the sweeps it needs on A (Drmac and
// Veselic, "New fast and accurate Jacobi SVD algorithm", SIAM J. Matrix Anal.
// Appl. 29 (2008), use column pivoting; measured on the 128 x 128 matrices of
// kappa 1e3..1e8 with the reference's pair order, Jacobi on R^T takes 10-11
// sweeps with pivoting, 11-12 without and 17-26 on A, so the QR here does not
// pivot: P = I, and the permutation output is the identity).  With
// X = Ux S Vx^T,
//     A = (Q1 Vx) S (P Ux)^T,
// so U = Q1 Vx and V = P Ux.  Two kernels, one workgroup per matrix each:
//  * qr_f64: Householder QR in FP64, every column in the registers of a lane
//    group, two workgroup barriers per step (see the kernel).  Writes
//    X = R^T (row-major n x n), the reflectors (column-major m x n, explicit
//    zeros above the unit diagonal), tau and the permutation.
//  * precond_finish_f64: the reflectors in LDS; lane groups of eight carry
//    one column of [Vx; 0] in registers through H_{n-1} ... H_0 (U = Q1 Vx),
//    apply the U rule (zero columns where sigma <= eps), and write V = P Ux
//    by rows.  A column of Ux that is exactly zero (sigma = 0) is completed to
//    an orthonormal basis from unit vectors, so V stays orthogonal for
//    rank-deficient inputs.
#include "jsvd_device.hpp"
#include "jsvd_kernels.hpp"

namespace jsvd
{
    namespace
    {
        constexpr int kFinishBlock = 512;
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
                        args.perm[w.s_off + j] = j;
                        args.tau[w.s_off + j] = s_tau[j];
                    }
                }
                __syncthreads();
            }
            finish_list(args.counters);
        }

        template <int R>
        __global__ __launch_bounds__(kFinishBlock) void precond_finish_f64(PrecondArgs args)
        {
            extern __shared__ double2 lds_raw[];
            double *lds = reinterpret_cast<double *>(lds_raw);
            __shared__ int s_item;
            constexpr int G = kFinishGroup, GROUPS = kFinishBlock / G, LD = G * R;
            const int tid = threadIdx.x, gid = tid / G, lane = tid & (G - 1);
            for (;;)
            {
                claim(&args.counters[2], &s_item);
                const int item = s_item;
                if (item >= args.count)
                    break;
                const WorkItem w = args.items[item];
                const int m = w.m, n = w.n;
                // Reflectors column-major with LD >= m rows, zero padding rows
                // (branch-free loops below).
                double *Hs = lds;
                double *taus = Hs + LD * n;
                const double *rg = args.refl + w.a_off;
                for (int idx = tid; idx < LD * n; idx += kFinishBlock)
                {
                    const int k = idx / LD, i = idx - k * LD;
                    Hs[idx] = i < m ? rg[k * m + i] : 0.0;
                }
                for (int j = tid; j < n; j += kFinishBlock)
                    taus[j] = args.tau[w.s_off + j];
                __syncthreads();
                const double *vx = args.vx + w.v_off;
                const double *sg = args.sigma + w.s_off;
                double *ug = args.u + w.a_off;
                // U = H_0 H_1 ... H_{n-1} [Vx; 0], one column per lane group;
                // lane l holds rows l + G t.
                for (int j = gid; j < n; j += GROUPS)
                {
                    double y[R];
#pragma unroll
                    for (int t = 0; t < R; ++t)
                    {
                        const int i = lane + G * t;
                        y[t] = i < n ? vx[i * n + j] : 0.0;
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
                    // The reference's U rule: a column whose sigma is not above
                    // eps is written as zeros.
                    const bool keep = sg[j] > args.eps;
#pragma unroll
                    for (int t = 0; t < R; ++t)
                    {
                        const int i = lane + G * t;
                        if (i < m)
                            ug[i * n + j] = keep ? y[t] : 0.0;
                    }
                }
                // V = P Ux (row i of Ux is row perm[i] of V), assembled in LDS
                // over the reflectors, which are no longer needed.
                __syncthreads();
                const double *ux = args.ux + w.v_off;
                const int32_t *pm = args.perm + w.s_off;
                double *Vs = lds; // n x n row-major (m >= n, so it fits)
                for (int idx = tid; idx < n * n; idx += kFinishBlock)
                {
                    const int i = idx / n, j = idx - i * n;
                    Vs[pm[i] * n + j] = ux[idx];
                }
                int zero = 0;
                for (int j = tid; j < n; j += kFinishBlock)
                    zero |= sg[j] == 0.0;
                zero = __syncthreads_or(zero);
                // An exactly zero singular value leaves a zero column in Ux.
                // Wave 0 fills each such column with the unit vector least
                // covered by the other columns (smallest sum_c V[i][c]^2),
                // projected off them twice and normalised, so V stays
                // orthogonal for rank-deficient inputs.
                if (zero && tid < 64)
                {
                    double *r = Vs + n * n; // n scratch values past V (taus are no longer needed)
                    for (int j = 0; j < n; ++j)
                    {
                        if (!(sg[j] == 0.0))
                            continue;
                        double best = 1e300;
                        int bi = n;
                        for (int i = tid; i < n; i += 64)
                        {
                            double cov = 0.0;
                            for (int c = 0; c < n; ++c)
                                if (c != j)
                                    cov = fma(Vs[i * n + c], Vs[i * n + c], cov);
                            if (cov < best)
                            {
                                best = cov;
                                bi = i;
                            }
                        }
                        for (int o = 32; o > 0; o >>= 1)
                        {
                            const double ob = __shfl_xor(best, o, 64);
                            const int oi = __shfl_xor(bi, o, 64);
                            if (ob < best || (ob == best && oi < bi))
                            {
                                best = ob;
                                bi = oi;
                            }
                        }
                        for (int i = tid; i < n; i += 64)
                            r[i] = i == bi ? 1.0 : 0.0;
                        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        for (int pass = 0; pass < 2; ++pass)
                            for (int c = 0; c < n; ++c)
                            {
                                if (c == j)
                                    continue;
                                double dot = 0.0;
                                for (int i = tid; i < n; i += 64)
                                    dot = fma(Vs[i * n + c], r[i], dot);
                                for (int o = 32; o > 0; o >>= 1)
                                    dot += __shfl_xor(dot, o, 64);
                                for (int i = tid; i < n; i += 64)
                                    r[i] = fma(-dot, Vs[i * n + c], r[i]);
                                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                                __builtin_amdgcn_wave_barrier();
                            }
                        double nr = 0.0;
                        for (int i = tid; i < n; i += 64)
                            nr = fma(r[i], r[i], nr);
                        for (int o = 32; o > 0; o >>= 1)
                            nr += __shfl_xor(nr, o, 64);
                        const double inv = nr > 0.0 ? 1.0 / sqrt(nr) : 0.0;
                        for (int i = tid; i < n; i += 64)
                            Vs[i * n + j] = r[i] * inv;
                        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                    }
                }
                __syncthreads();
                double *vg = args.v + w.v_off;
                for (int idx = tid; idx < n * n; idx += kFinishBlock)
                    vg[idx] = Vs[idx];
                __syncthreads();
            }
            finish_list(args.counters + 2);
        }

        typedef double f64x4 __attribute__((ext_vector_type(4)));

        // One n x n product C = A B (n <= NT) on FP64 MFMA
        // (v_mfma_f64_16x16x4f64) by one workgroup of four waves: wave w owns
        // the (NT/2) x (NT/2) quadrant (w / 2, w % 2) of C as T x T tiles of
        // 16 x 16 (T = NT / 32), accumulated in `acc`.  K runs in blocks of
        // 16; in step s of a block, lane l supplies A[i0 + l%16][k_l] and
        // B[k_l][j0 + l%16] with k_l = k0 + 4 (l / 16) + s (any assignment of
        // the block's k to the four lane groups sums the same products, and
        // this one gives each lane four consecutive k, so row-major A and
        // column-major B operands are read as 32 contiguous bytes).  a(i, k)
        // and b(k, j) return the operands (zero outside n x n); lane l then
        // holds C[i0 + l/16 + 4 r][j0 + l%16] in acc[x][y][r].
        template <int NT, typename FA, typename FB>
        __device__ __forceinline__ void mfma_nn(f64x4 (&acc)[NT / 32][NT / 32], FA a, FB b, int n, int wr, int wc,
                                                int lr, int lk)
        {
            constexpr int T = NT / 32;
#pragma unroll
            for (int x = 0; x < T; ++x)
#pragma unroll
                for (int y = 0; y < T; ++y)
                    acc[x][y] = f64x4{0.0, 0.0, 0.0, 0.0};
            for (int k0 = 0; k0 < n; k0 += 16)
            {
                double av[T][4], bv[T][4];
#pragma unroll
                for (int t = 0; t < T; ++t)
#pragma unroll
                    for (int s = 0; s < 4; ++s)
                    {
                        const int k = k0 + 4 * lk + s;
                        av[t][s] = a(wr + 16 * t + lr, k);
                        bv[t][s] = b(k, wc + 16 * t + lr);
                    }
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int x = 0; x < T; ++x)
#pragma unroll
                        for (int y = 0; y < T; ++y)
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
        // Q and X are read from L2.
        template <int NT>
        __global__ __launch_bounds__(256) void polar_bq_f64(PrecondArgs pa, MixedArgs ma)
        {
            constexpr int T = NT / 32, LD = NT + 1;
            extern __shared__ double2 polar_raw[];
            double *S = reinterpret_cast<double *>(polar_raw);
            __shared__ double s_max[4];
            __shared__ int s_item;
            const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
            const int lr = lane & 15, lk = lane >> 4;
            const int wr = (wave >> 1) * (NT / 2), wc = (wave & 1) * (NT / 2);
            f64x4 acc[T][T];
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
                    double emax = 0.0;
#pragma unroll
                    for (int xx = 0; xx < T; ++xx)
#pragma unroll
                        for (int yy = 0; yy < T; ++yy)
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
                    const double e_all = fmax(fmax(s_max[0], s_max[1]), fmax(s_max[2], s_max[3]));
                    // F = E E (both operands from LDS), then this wave's part of
                    // M = I - E/2 + 3 F / 8 in registers.
                    mfma_nn<NT>(acc, sm, sm, n, wr, wc, lr, lk);
#pragma unroll
                    for (int xx = 0; xx < T; ++xx)
#pragma unroll
                        for (int yy = 0; yy < T; ++yy)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                            {
                                const int i = wr + 16 * xx + lk + 4 * r, j = wc + 16 * yy + lr;
                                const double e = S[i * LD + j];
                                acc[xx][yy][r] = fma(0.375, acc[xx][yy][r], fma(-0.5, e, i == j ? 1.0 : 0.0));
                            }
                    __syncthreads(); // every wave has read E
#pragma unroll
                    for (int xx = 0; xx < T; ++xx)
#pragma unroll
                        for (int yy = 0; yy < T; ++yy)
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
#pragma unroll
                    for (int xx = 0; xx < T; ++xx)
#pragma unroll
                        for (int yy = 0; yy < T; ++yy)
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
#pragma unroll
                for (int xx = 0; xx < T; ++xx)
#pragma unroll
                    for (int yy = 0; yy < T; ++yy)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                        {
                            const int i = wr + 16 * xx + lk + 4 * r, j = wc + 16 * yy + lr;
                            if (i < n && j < n)
                                b[j * n + i] = acc[xx][yy][r];
                        }
                __syncthreads(); // LDS free for the next matrix
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
    } // namespace

    namespace
    {
        // QR shapes: (workgroup, lanes per column, rows per lane); a matrix
        // takes the first that fits (n <= BS / G columns, m <= G RW rows).
        struct QrSpec
        {
            int block, group, rows;
        };
        constexpr QrSpec kQr[] = {{256, 4, 8}, {256, 4, 16}, {256, 4, 32}, {512, 4, 32}, {512, 8, 32}};
        constexpr int kQrCount = sizeof(kQr) / sizeof(kQr[0]);
        template <int I>
        const void *qr_symbol_t()
        {
            return reinterpret_cast<const void *>(&qr_f64<kQr[I].block, kQr[I].group, kQr[I].rows>);
        }
        template <int I>
        void qr_launch_t(int grid, int lds, const PrecondArgs &args, hipStream_t stream)
        {
            qr_f64<kQr[I].block, kQr[I].group, kQr[I].rows><<<grid, kQr[I].block, lds, stream>>>(args);
        }
        const void *(*const kQrSymbol[])() = {qr_symbol_t<0>, qr_symbol_t<1>, qr_symbol_t<2>, qr_symbol_t<3>,
                                             qr_symbol_t<4>};
        void (*const kQrLaunch[])(int, int, const PrecondArgs &, hipStream_t) = {
            qr_launch_t<0>, qr_launch_t<1>, qr_launch_t<2>, qr_launch_t<3>, qr_launch_t<4>};
        static_assert(sizeof(kQrLaunch) / sizeof(kQrLaunch[0]) == kQrCount, "one launcher per shape");
    } // namespace

    int qrcp_variant(int64_t m, int64_t n)
    {
        for (int v = 0; v < kQrCount; ++v)
            if (n <= kQr[v].block / kQr[v].group && m <= int64_t(kQr[v].group) * kQr[v].rows)
                return v;
        return -1;
    }

    int64_t qrcp_lds_bytes(int variant)
    {
        // The reflector of a step over every register row (G RW >= m).
        return int64_t(kQr[variant].group) * kQr[variant].rows * 8;
    }

    namespace
    {
        // Rows per lane of the back-transform: 8, 16 or 32 (m <= 64, 128, 256).
        int finish_rows(int64_t m)
        {
            return m <= 8 * kFinishGroup ? 8 : (m <= 16 * kFinishGroup ? 16 : 32);
        }
    } // namespace

    int64_t precond_finish_lds_bytes(int64_t m, int64_t n)
    {
        // Reflectors (LD = 8 x rows per lane, zero padded) and tau, later
        // V n x n and one scratch column.
        return (int64_t(kFinishGroup) * finish_rows(m) * n + n) * 8;
    }

    bool precond_fits(int64_t m, int64_t n)
    {
        return n <= 128 && qrcp_variant(m, n) >= 0 && m <= int64_t(kFinishGroup) * kFinishRows &&
               precond_finish_lds_bytes(m, n) <= kMaxLdsBytes;
    }

    const void *qrcp_symbol(int variant)
    {
        return variant >= 0 && variant < kQrCount ? kQrSymbol[variant]() : nullptr;
    }
    int qrcp_block(int variant)
    {
        return kQr[variant].block;
    }
    const void *precond_finish_symbol(int64_t m)
    {
        const int r = finish_rows(m);
        return r == 8    ? reinterpret_cast<const void *>(&precond_finish_f64<8>)
               : r == 16 ? reinterpret_cast<const void *>(&precond_finish_f64<16>)
                         : reinterpret_cast<const void *>(&precond_finish_f64<32>);
    }
    int precond_finish_block()
    {
        return kFinishBlock;
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

    hipError_t launch_polar_bq_f64(int64_t n, int grid, const PrecondArgs &pa, const MixedArgs &ma, hipStream_t stream)
    {
        if (n <= 64)
            polar_bq_f64<64><<<grid, 256, polar_bq_lds_bytes(n), stream>>>(pa, ma);
        else
            polar_bq_f64<128><<<grid, 256, polar_bq_lds_bytes(n), stream>>>(pa, ma);
        return hipGetLastError();
    }

    hipError_t launch_qrcp_f64(int variant, int grid, int lds_bytes, const PrecondArgs &args, hipStream_t stream)
    {
        if (variant < 0 || variant >= kQrCount)
            return hipErrorInvalidValue;
        kQrLaunch[variant](grid, lds_bytes, args, stream);
        return hipGetLastError();
    }

    hipError_t launch_precond_finish_f64(int64_t m, int grid, int lds_bytes, const PrecondArgs &args,
                                         hipStream_t stream)
    {
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
