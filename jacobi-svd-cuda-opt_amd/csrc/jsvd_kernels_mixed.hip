// jsvd_kernels_mixed.hip -- mixed-precision one-sided Jacobi (no reference
// counterpart: the reference is FP64 only).
//
// Matrices that fit the QR preconditioner (jsvd_kernels_precond.hip: two
// Householder QRs, the polar factor with B = X Q, the back-transform) run
// steps 1 and 4 below on X = R2^T in place of A, with the preconditioner's
// kernels instead of steps 2 and 3.  The others take four work-list
// launches on one stream per part of a batch (the plan splits a large bucket
// of wide matrices into two parts, each on a stream of its own):
//  1. mixed_sweeps_compact_f32: FP32 sweeps with A and V in the compact
//     LDS layout of the FP64 class kernel (top columns held in registers,
//     half + 2 bottom slots in LDS: two 128 x 128 matrices per CU).  The dot
//     p.q and the column updates run in FP32 with FP32 accumulation; the
//     squared column norms are summed afresh at each sweep's start and then
//     carried through the rotations in FP64; the rotation parameters (zeta,
//     t, c, s) and the off-diagonal test are computed in FP64.  The FP32
//     phase stops at the FP32 rounding floor, tol32 = max(eps, 8 sqrt(m)
//     2^-24).  V32 is written out as FP64.
//  2. orthonormalize_f64: modified Gram-Schmidt of V32's columns in FP64,
//     giving an orthogonal Q (V32 is orthogonal only to ~1e-6).
//  3. gemm_aq_f64: B = A Q in FP64 on MFMA (v_mfma_f64_16x16x4f64), written
//     column-major into the U buffer as scratch.
//  4. The FP64 class kernel on B with V starting from Q (FP64 refinement
//     sweeps under the caller's eps, at most refine_sweeps), which writes U,
//     sigma and V = Q W.  B's columns are already orthogonal to ~1e-6, so a
//     few FP64 sweeps recover full FP64 accuracy.
#include "jsvd_device.hpp"
#include "jsvd_kernels.hpp"

#include <algorithm>

namespace jsvd
{
    namespace
    {
        template <int CTRL>
        __device__ __forceinline__ float dpp_f32(float x)
        {
            return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, true));
        }

        // FP32 group all-reduce over G lanes (same DPP ladder as the FP64 one).
        template <int G>
        __device__ __forceinline__ float group_sum_f32(float x)
        {
            if constexpr (G >= 2)
                x += dpp_f32<0xB1>(x);
            if constexpr (G >= 4)
                x += dpp_f32<0x4E>(x);
            if constexpr (G >= 8)
                x += dpp_f32<0x141>(x);
            if constexpr (G >= 16)
                x += dpp_f32<0x140>(x);
            if constexpr (G >= 32)
                x += __shfl_xor(x, 16, 64);
            if constexpr (G >= 64)
                x += __shfl_xor(x, 32, 64);
            return x;
        }

        struct F32Spec
        {
            int block, group, rmax4, vmax4;
        };
        // FP32 shape classes: lanes hold row quads (one 16-byte LDS access per
        // lane); columns are zero-padded to a multiple of 4G rows.
        constexpr F32Spec kF32Classes[] = {
            {64, 4, 4, 2},     // n <= 32, m <= 64
            {256, 8, 4, 2},    // n <= 64, m <= 128
            {512, 8, 4, 4},    // n <= 128, m <= 128
            {512, 16, 4, 4},   // n <= 256 (two passes over 128), m <= 256
        };
        constexpr int kF32Count = sizeof(kF32Classes) / sizeof(kF32Classes[0]);

        // Pairs per lane group and round of an FP32 class: its widest matrix
        // (4 G vmax4 columns, so half = 2 G vmax4 pairs) over its BS / G
        // groups, one or two.
        __host__ __device__ constexpr int f32_passes(int block, int group, int vmax4)
        {
            return (2 * group * vmax4 + block / group - 1) / (block / group) < 2 ? 1 : 2;
        }

        __host__ __device__ __forceinline__ int padded4(int rows, int G)
        {
            return ((rows + 4 * G - 1) / (4 * G)) * (4 * G);
        }

        // One column pair in FP32 with its top column held in registers (xp
        // for A, yp for V, updated in place) and its bottom column in LDS (aq,
        // vq): the FP32 dot p.q (FP32 accumulation) and FP32 updates; the
        // off-diagonal test and the rotation parameters (zeta, t, c, s) in
        // FP64.  The squared norms come from the columns' tracked norms (pn
        // for the top column, *qn for the bottom one), which a rotation by
        // the FP32 (c, s) carries over in FP64:
        //   |p'|^2 = c^2 |p|^2 - 2 c s p.q + s^2 |q|^2,
        //   |q'|^2 = s^2 |p|^2 + 2 c s p.q + c^2 |q|^2
        // (one dot per pair instead of three; the sweeps start from norms
        // summed afresh).
        template <int G, int RMAX4, int VMAX4>
        __device__ __forceinline__ int pair_f32_held(float4 (&xp)[RMAX4], float4 (&yp)[VMAX4],
                                                     float4 *__restrict__ aq, float4 *__restrict__ vq, int rows4,
                                                     int vrows4, int lane, double tol, double &pn,
                                                     double *__restrict__ qn)
        {
            float4 xq[RMAX4];
            float pq = 0.0f;
#pragma unroll
            for (int j = 0; j < RMAX4; ++j)
            {
                if (j < rows4)
                {
                    xq[j] = aq[lane + j * G];
                    pq += xp[j].x * xq[j].x + xp[j].y * xq[j].y + xp[j].z * xq[j].z + xp[j].w * xq[j].w;
                }
            }
            const double dpp = pn, dqq = *qn;
            const double dpq = group_sum_f32<G>(pq);
            if (!dev::rotation_needed(dpp, dqq, dpq, tol))
                return 0;
            float4 yq[VMAX4];
#pragma unroll
            for (int j = 0; j < VMAX4; ++j)
            {
                if (j < vrows4)
                    yq[j] = vq[lane + j * G];
            }
            // FP64 parameters from the refined hardware estimates (the FP64
            // kernels' evaluation: about half the dependent operations of
            // IEEE division and square root).
            double cd, sd;
            dev::rotation_from_tangent(dev::rotation_tangent(dpp, dqq, dpq), cd, sd);
            const float c = static_cast<float>(cd), s = static_cast<float>(sd);
            if (s != 0.0f)
            {
#pragma unroll
                for (int j = 0; j < RMAX4; ++j)
                {
                    if (j < rows4)
                    {
                        float4 np, nq;
                        np.x = c * xp[j].x - s * xq[j].x;
                        np.y = c * xp[j].y - s * xq[j].y;
                        np.z = c * xp[j].z - s * xq[j].z;
                        np.w = c * xp[j].w - s * xq[j].w;
                        nq.x = s * xp[j].x + c * xq[j].x;
                        nq.y = s * xp[j].y + c * xq[j].y;
                        nq.z = s * xp[j].z + c * xq[j].z;
                        nq.w = s * xp[j].w + c * xq[j].w;
                        xp[j] = np;
                        aq[lane + j * G] = nq;
                    }
                }
#pragma unroll
                for (int j = 0; j < VMAX4; ++j)
                {
                    if (j < vrows4)
                    {
                        float4 np, nq;
                        np.x = c * yp[j].x - s * yq[j].x;
                        np.y = c * yp[j].y - s * yq[j].y;
                        np.z = c * yp[j].z - s * yq[j].z;
                        np.w = c * yp[j].w - s * yq[j].w;
                        nq.x = s * yp[j].x + c * yq[j].x;
                        nq.y = s * yp[j].y + c * yq[j].y;
                        nq.z = s * yp[j].z + c * yq[j].z;
                        nq.w = s * yp[j].w + c * yq[j].w;
                        yp[j] = np;
                        vq[lane + j * G] = nq;
                    }
                }
                const double c2 = static_cast<double>(c) * c, s2 = static_cast<double>(s) * s;
                const double x = 2.0 * static_cast<double>(c) * s * dpq;
                pn = fmax(c2 * dpp - x + s2 * dqq, 0.0);
                if (lane == 0)
                    *qn = fmax(s2 * dpp + x + c2 * dqq, 0.0);
            }
            return 1;
        }

        // FP32 sweeps in the compact layout of the FP64 class kernel: each
        // lane group holds the top column (A and V) of its circle position in
        // registers and follows it along the ring top[1] -> ... -> top[half-1]
        // -> bottom[half-1] -> ... -> bottom[0] -> top[1]; LDS holds only the
        // bottom row, half + 2 slots of A and of V, the bottom column at
        // position P in global round g living in slot (g + P) mod (half + 2).
        // Same pairs in the same order as the reference circle, and half the
        // LDS of holding all columns (two 128 x 128 matrices per CU).
        template <int BS, int G, int RMAX4, int VMAX4>
        __global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(G < 16 ? 4 : 1))) void mixed_sweeps_compact_f32(MixedArgs args)
        {
            extern __shared__ float4 lds_quads[];
            float *lds32 = reinterpret_cast<float *>(lds_quads);
            __shared__ int s_item;
            constexpr int GROUPS = BS / G;
            // One pass (every class up to 128 x 128) holds RMAX4 + VMAX4
            // quads of top columns in registers.
            constexpr int PASSES = f32_passes(BS, G, VMAX4);
            const int tid = threadIdx.x, gid = tid / G, lane = tid & (G - 1);
            for (;;)
            {
                if (tid == 0)
                    s_item = (int)atomicAdd(&args.counters[0], 1u);
                __syncthreads();
                const int item = s_item;
                if (item >= args.count)
                    break;
                const WorkItem w = args.items[item];
                const int m = w.m, n = w.n;
                const int lda = padded4(m, G), ldv = padded4(n, G);
                const int even = n + (n & 1), half = even >> 1, ring = even - 1;
                const int rounds = n < 2 ? 0 : even - 1;
                const int ns = half + 2;
                float *As = lds32;          // ns slots of lda rows
                float *Vs = As + ns * lda;  // ns slots of ldv rows
                float4 *A4 = reinterpret_cast<float4 *>(As);
                float4 *V4 = reinterpret_cast<float4 *>(Vs);
                double *Ns = reinterpret_cast<double *>(Vs + ns * ldv); // ns squared slot norms
                const int lda4 = lda / 4, ldv4 = ldv / 4;
                const int rows4 = lda / (4 * G), vrows4 = ldv / (4 * G);
                const double *ag = args.a + w.a_off;
                // Slots 0 .. half-1 <- columns col(0 .. half-1) of A and of the
                // identity (zeros for the dummy column and padding).
                auto fill = [&](auto col) {
                    for (int idx = tid; idx < half * lda; idx += BS)
                    {
                        const int sl = idx / lda, i = idx - sl * lda, j = col(sl);
                        As[idx] = (i < m && j < n) ? static_cast<float>(ag[i * n + j]) : 0.0f;
                    }
                    for (int idx = tid; idx < half * ldv; idx += BS)
                    {
                        const int sl = idx / ldv, i = idx - sl * ldv, j = col(sl);
                        Vs[idx] = (i == j && j < n) ? 1.0f : 0.0f;
                    }
                };
                int pos[PASSES], ct[PASSES], cb[PASSES], fresh[PASSES];
                double tn[PASSES]; // squared norm of the held top column
                float4 ta[PASSES][RMAX4];
                float4 tv[PASSES][VMAX4];
                auto load_top = [&](int i, int sl) {
#pragma unroll
                    for (int j = 0; j < RMAX4; ++j)
                        if (j < rows4)
                            ta[i][j] = A4[sl * lda4 + lane + j * G];
#pragma unroll
                    for (int j = 0; j < VMAX4; ++j)
                        if (j < vrows4)
                            tv[i][j] = V4[sl * ldv4 + lane + j * G];
                };
                // The top row (position k holds column k) is staged through
                // slots 0 .. half-1 into registers, then the bottom row
                // (position P holds column even-1-P) fills slots 0 .. half-1.
                fill([](int sl) { return sl; });
                __syncthreads();
#pragma unroll
                for (int i = 0; i < PASSES; ++i)
                {
                    const int k = gid + i * GROUPS;
                    pos[i] = k;
                    ct[i] = k;
                    cb[i] = even - 1 - k;
                    fresh[i] = 0;
                    if (k < half && k < n)
                        load_top(i, k);
                }
                __syncthreads();
                fill([even](int sl) { return even - 1 - sl; });
                __syncthreads();
                const double tol = fmax(args.eps, 8.0 * sqrt(static_cast<double>(m)) * 5.9604644775390625e-08);
                int z = even - 1; // bottom[0]'s column this round
                int gs = 0;       // global round mod ns
                int executed = 0;
                // Squared column norms summed afresh from the FP32 columns:
                // every slot's (the bottom row's columns and the tops that a
                // group loads later) and every held top's.
                auto sq_norm = [&](const float4 *col) {
                    float acc = 0.0f;
#pragma unroll
                    for (int j = 0; j < RMAX4; ++j)
                        if (j < rows4)
                        {
                            const float4 x = col[lane + j * G];
                            acc += x.x * x.x + x.y * x.y + x.z * x.z + x.w * x.w;
                        }
                    return static_cast<double>(group_sum_f32<G>(acc));
                };
                auto fresh_norms = [&]() {
                    for (int sl = gid; sl < ns; sl += GROUPS)
                    {
                        const double nn = sq_norm(A4 + sl * lda4);
                        if (lane == 0)
                            Ns[sl] = nn;
                    }
#pragma unroll
                    for (int i = 0; i < PASSES; ++i)
                    {
                        float acc = 0.0f;
#pragma unroll
                        for (int j = 0; j < RMAX4; ++j)
                            if (j < rows4)
                                acc += ta[i][j].x * ta[i][j].x + ta[i][j].y * ta[i][j].y + ta[i][j].z * ta[i][j].z +
                                       ta[i][j].w * ta[i][j].w;
                        tn[i] = static_cast<double>(group_sum_f32<G>(acc));
                    }
                };
                for (int sweep = 0; sweep < args.max_sweeps; ++sweep)
                {
                    int any = 0;
                    fresh_norms();
                    __syncthreads();
                    for (int r = 0; r < rounds; ++r)
                    {
                        const int z1 = z - 1 < 1 ? ring : z - 1;
                        const int z2 = z - 2 < 1 ? z - 2 + ring : z - 2;
                        const int ls = gs == 0 ? ns - 1 : gs - 1;     // fresh top loads
                        const int ws = gs < 2 ? gs + ns - 2 : gs - 2; // wrap stores
#pragma unroll
                        for (int i = 0; i < PASSES; ++i)
                        {
                            const int k = gid + i * GROUPS;
                            if (k < half)
                            {
                                const int p = ct[i], q = cb[i];
                                if (fresh[i] & (p < n))
                                {
                                    load_top(i, ls);
                                    tn[i] = Ns[ls];
                                }
                                int bsl = gs + pos[i];
                                bsl = half == 1 ? 0 : (bsl >= ns ? bsl - ns : bsl);
                                int rot = 0;
                                if ((p < n) & (q < n))
                                    rot = pair_f32_held<G, RMAX4, VMAX4>(ta[i], tv[i], A4 + bsl * lda4,
                                                                         V4 + bsl * ldv4, rows4, vrows4, lane, tol,
                                                                         tn[i], Ns + bsl);
                                any |= rot;
                                // Leaving position half-1: the top column becomes
                                // bottom[half-1] in slot ws; the group restarts at
                                // position 1 with the previous bottom[0].
                                const int wrap = (k > 0) & (pos[i] == half - 1);
                                if (wrap & (p < n))
                                {
#pragma unroll
                                    for (int j = 0; j < RMAX4; ++j)
                                        if (j < rows4)
                                            A4[ws * lda4 + lane + j * G] = ta[i][j];
#pragma unroll
                                    for (int j = 0; j < VMAX4; ++j)
                                        if (j < vrows4)
                                            V4[ws * ldv4 + lane + j * G] = tv[i][j];
                                    if (lane == 0)
                                        Ns[ws] = tn[i];
                                }
                                const int q1 = q - 1 < 1 ? ring : q - 1;
                                const int q2 = q - 2 < 1 ? q - 2 + ring : q - 2;
                                cb[i] = k == 0 ? q1 : (wrap ? z2 : q2);
                                ct[i] = wrap ? z : p;
                                pos[i] = wrap ? 1 : (k > 0 ? pos[i] + 1 : 0);
                                fresh[i] = wrap;
                            }
                        }
                        z = z1;
                        gs = gs + 1 == ns ? 0 : gs + 1;
                        __syncthreads();
                    }
                    executed = sweep + 1;
                    if (!__syncthreads_or(any))
                        break;
                }
                // V32 -> packed column-major FP64 scratch: every group writes
                // its held top column (loading it first if it is still in its
                // slot) and its bottom column from the slot of its position.
                double *qg = args.q + w.v_off;
                const int ls = gs == 0 ? ns - 1 : gs - 1;
#pragma unroll
                for (int i = 0; i < PASSES; ++i)
                {
                    const int k = gid + i * GROUPS;
                    if (k < half)
                    {
                        if (fresh[i] && ct[i] < n)
                            load_top(i, ls);
                        int bsl = gs + pos[i];
                        bsl = half == 1 ? 0 : (bsl >= ns ? bsl - ns : bsl);
                        if (ct[i] < n)
                        {
#pragma unroll
                            for (int j = 0; j < VMAX4; ++j)
                            {
                                const int r0 = 4 * (lane + j * G);
                                if (j < vrows4)
                                {
                                    const float4 x = tv[i][j];
                                    double *col = qg + (int64_t)ct[i] * n;
                                    if (r0 < n)
                                        col[r0] = x.x;
                                    if (r0 + 1 < n)
                                        col[r0 + 1] = x.y;
                                    if (r0 + 2 < n)
                                        col[r0 + 2] = x.z;
                                    if (r0 + 3 < n)
                                        col[r0 + 3] = x.w;
                                }
                            }
                        }
                        if (cb[i] < n)
                        {
                            double *col = qg + (int64_t)cb[i] * n;
                            for (int row = lane; row < n; row += G)
                                col[row] = static_cast<double>(Vs[bsl * ldv + row]);
                        }
                    }
                }
                if (tid == 0 && args.sweeps)
                    args.sweeps[w.index] = executed;
                __syncthreads();
            }
            if (tid == 0)
            {
                const unsigned int done = atomicAdd(&args.counters[1], 1u);
                if (done == gridDim.x - 1)
                {
                    __hip_atomic_store(&args.counters[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&args.counters[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }

        // Modified Gram-Schmidt of the n columns of Q (column-major, n x n) in
        // FP64, in LDS, with deferred normalisation: step j projects the unit
        // vector q_j / ||q_j|| out of every later column (lane groups of 16,
        // one column each at a time), and the group that updates column j + 1
        // also computes its norm for the next step, so a step needs a single
        // workgroup barrier; every column is scaled to unit length at the end.
        template <int BS>
        __global__ __launch_bounds__(BS) void orthonormalize_f64(MixedArgs args)
        {
            extern __shared__ double ldsq[];
            __shared__ int s_item;
            constexpr int G = 16, GROUPS = BS / G;
            const int tid = threadIdx.x, lane = tid & (G - 1), gid = tid / G;
            for (;;)
            {
                if (tid == 0)
                    s_item = (int)atomicAdd(&args.counters[2], 1u);
                __syncthreads();
                const int item = s_item;
                if (item >= args.count)
                    break;
                const WorkItem w = args.items[item];
                const int n = w.n;
                double *Q = ldsq;
                double *inv = ldsq + n * n; // 1 / ||q_j|| of the projected q_j
                double *qg = args.q + w.v_off;
                for (int idx = tid; idx < n * n; idx += BS)
                    Q[idx] = qg[idx];
                __syncthreads();
                if (gid == 0)
                {
                    double d = 0.0;
                    for (int i = lane; i < n; i += G)
                        d += Q[i] * Q[i];
                    d = dev::group_sum_dpp<G>(d);
                    if (lane == 0)
                        inv[0] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
                }
                __syncthreads();
                for (int j = 0; j < n; ++j)
                {
                    const double *qj = Q + j * n;
                    const double ij2 = inv[j] * inv[j];
                    for (int k = j + 1 + gid; k < n; k += GROUPS)
                    {
                        double *qk = Q + k * n;
                        double d = 0.0;
                        for (int i = lane; i < n; i += G)
                            d += qj[i] * qk[i];
                        d = dev::group_sum_dpp<G>(d) * ij2;
                        double nk = 0.0;
                        for (int i = lane; i < n; i += G)
                        {
                            const double x = qk[i] - d * qj[i];
                            qk[i] = x;
                            nk += x * x;
                        }
                        if (k == j + 1)
                        {
                            nk = dev::group_sum_dpp<G>(nk);
                            if (lane == 0)
                                inv[k] = nk > 0.0 ? 1.0 / sqrt(nk) : 0.0;
                        }
                    }
                    __syncthreads();
                }
                for (int idx = tid; idx < n * n; idx += BS)
                    qg[idx] = Q[idx] * inv[idx / n];
                __syncthreads();
            }
            if (tid == 0)
            {
                const unsigned int done = atomicAdd(&args.counters[3], 1u);
                if (done == gridDim.x - 1)
                {
                    __hip_atomic_store(&args.counters[2], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&args.counters[3], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }

        typedef double f64x4 __attribute__((ext_vector_type(4)));

        // B = A Q on FP64 MFMA.  A: packed row-major m x n; Q: packed
        // column-major n x n; B: column-major m x n written into the U buffer.
        // Each wave computes 16 x 16 output tiles with v_mfma_f64_16x16x4f64:
        // lane l supplies A[i0 + l%16][k0 + l/16] and Q[k0 + l/16][j0 + l%16]
        // and receives D[i0 + l/16 + 4r][j0 + l%16], r = 0..3 (output layout
        // measured on gfx950).
        template <int BS>
        __global__ __launch_bounds__(BS) void gemm_aq_f64(MixedArgs args)
        {
            __shared__ int s_item;
            constexpr int WAVES = BS / 64;
            const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
            const int lr = l & 15, lk = l >> 4;
            for (;;)
            {
                if (tid == 0)
                    s_item = (int)atomicAdd(&args.counters[4], 1u);
                __syncthreads();
                const int item = s_item;
                if (item >= args.count)
                    break;
                const WorkItem w = args.items[item];
                const int m = w.m, n = w.n;
                const double *A = args.a + w.a_off;
                const double *Q = args.q + w.v_off;
                double *B = args.u + w.a_off;
                const int ti_n = (m + 15) / 16, tj_n = (n + 15) / 16;
                for (int tile = wave; tile < ti_n * tj_n; tile += WAVES)
                {
                    const int i0 = (tile / tj_n) * 16, j0 = (tile % tj_n) * 16;
                    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
                    const int ia = i0 + lr, jb = j0 + lr;
                    for (int k0 = 0; k0 < n; k0 += 4)
                    {
                        const int k = k0 + lk;
                        const double av = (ia < m && k < n) ? A[(int64_t)ia * n + k] : 0.0;
                        const double bv = (jb < n && k < n) ? Q[(int64_t)jb * n + k] : 0.0;
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                    {
                        const int row = i0 + lk + 4 * r;
                        if (row < m && jb < n)
                            B[(int64_t)jb * m + row] = acc[r];
                    }
                }
                __syncthreads();
            }
            if (tid == 0)
            {
                const unsigned int done = atomicAdd(&args.counters[5], 1u);
                if (done == gridDim.x - 1)
                {
                    __hip_atomic_store(&args.counters[4], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&args.counters[5], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }

        template <int I>
        const void *f32_symbol_t()
        {
            constexpr F32Spec c = kF32Classes[I];
            return reinterpret_cast<const void *>(&mixed_sweeps_compact_f32<c.block, c.group, c.rmax4, c.vmax4>);
        }
        template <int I>
        void f32_launch_t(int grid, int lds_bytes, const MixedArgs &args, hipStream_t stream)
        {
            constexpr F32Spec c = kF32Classes[I];
            mixed_sweeps_compact_f32<c.block, c.group, c.rmax4, c.vmax4><<<grid, c.block, lds_bytes, stream>>>(args);
        }
        using F32Launch = void (*)(int, int, const MixedArgs &, hipStream_t);
        const F32Launch kF32Launch[] = {f32_launch_t<0>, f32_launch_t<1>, f32_launch_t<2>, f32_launch_t<3>};
        const void *(*const kF32Symbol[])() = {f32_symbol_t<0>, f32_symbol_t<1>, f32_symbol_t<2>, f32_symbol_t<3>};
        static_assert(sizeof(kF32Launch) / sizeof(kF32Launch[0]) == kF32Count);
    } // namespace

    constexpr int kOrthoBlock = 512;
    constexpr int kGemmBlock = 256;

    int64_t f32_class_lds_bytes(int cls, int64_t m, int64_t n)
    {
        // half + 2 slots of A and of V (compact layout) and their squared
        // norms (FP64).
        const int g = kF32Classes[cls].group;
        const int64_t slots = (n + 1) / 2 + 2;
        return (padded4(static_cast<int>(m), g) + padded4(static_cast<int>(n), g)) * slots * 4 + slots * 8;
    }

    bool f32_class_fits(int cls, int64_t m, int64_t n)
    {
        if (cls < 0 || cls >= kF32Count)
            return false;
        const F32Spec &c = kF32Classes[cls];
        const int64_t half = (n + 1) / 2;
        const int64_t groups = c.block / c.group;
        const int64_t passes = f32_passes(c.block, c.group, c.vmax4);
        if (half > passes * groups)
            return false;
        if (padded4(static_cast<int>(m), c.group) / (4 * c.group) > c.rmax4 ||
            padded4(static_cast<int>(n), c.group) / (4 * c.group) > c.vmax4)
            return false;
        return f32_class_lds_bytes(cls, m, n) <= kMaxLdsBytes;
    }

    int select_f32_class(int64_t m, int64_t n)
    {
        for (int i = 0; i < kF32Count; ++i)
            if (f32_class_fits(i, m, n))
                return i;
        return -1;
    }

    int f32_class_block(int cls)
    {
        return kF32Classes[cls].block;
    }

    const void *f32_class_symbol(int cls)
    {
        return (cls >= 0 && cls < kF32Count) ? kF32Symbol[cls]() : nullptr;
    }

    hipError_t launch_mixed_sweeps_f32(int cls, int grid, int lds_bytes, const MixedArgs &args, hipStream_t stream)
    {
        if (cls < 0 || cls >= kF32Count)
            return hipErrorInvalidValue;
        kF32Launch[cls](grid, lds_bytes, args, stream);
        return hipGetLastError();
    }

    int64_t orthonormalize_lds_bytes(int64_t n)
    {
        return (n * n + n) * 8;
    }

    const void *orthonormalize_symbol()
    {
        return reinterpret_cast<const void *>(&orthonormalize_f64<kOrthoBlock>);
    }

    int orthonormalize_block()
    {
        return kOrthoBlock;
    }

    hipError_t launch_orthonormalize_f64(int grid, int lds_bytes, const MixedArgs &args, hipStream_t stream)
    {
        orthonormalize_f64<kOrthoBlock><<<grid, kOrthoBlock, lds_bytes, stream>>>(args);
        return hipGetLastError();
    }

    const void *gemm_aq_symbol()
    {
        return reinterpret_cast<const void *>(&gemm_aq_f64<kGemmBlock>);
    }

    int gemm_aq_block()
    {
        return kGemmBlock;
    }

    hipError_t launch_gemm_aq_f64(int grid, const MixedArgs &args, hipStream_t stream)
    {
        gemm_aq_f64<kGemmBlock><<<grid, kGemmBlock, 0, stream>>>(args);
        return hipGetLastError();
    }
} // namespace jsvd
