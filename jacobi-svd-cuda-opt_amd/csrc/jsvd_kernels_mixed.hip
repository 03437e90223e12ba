// jsvd_kernels_mixed.hip -- mixed-precision one-sided Jacobi (no reference
// counterpart: the reference is FP64 only).
//
// Per matrix, four work-list launches on one stream:
//  1. mixed_sweeps_f32: FP32 sweeps with A and V resident in LDS.  Dots and
//     column updates run in FP32 with FP32 accumulation; the rotation
//     parameters (zeta, t, c, s) and the off-diagonal test are computed in
//     FP64 from the FP32 dots.  The FP32 phase stops at the FP32 rounding
//     floor, tol32 = max(eps, 8 sqrt(m) 2^-24).  V32 is written out as FP64.
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
            {1024, 16, 2, 2},  // n <= 128, m <= 128
            {512, 16, 4, 4},   // n <= 256 (two passes over 128), m <= 256
        };
        constexpr int kF32Count = sizeof(kF32Classes) / sizeof(kF32Classes[0]);

        __host__ __device__ __forceinline__ int padded4(int rows, int G)
        {
            return ((rows + 4 * G - 1) / (4 * G)) * (4 * G);
        }

        template <int G, int RMAX4, int VMAX4>
        __device__ __forceinline__ int pair_f32(float4 *__restrict__ ap, float4 *__restrict__ aq,
                                                float4 *__restrict__ vp, float4 *__restrict__ vq, int rows4,
                                                int vrows4, int lane, double tol)
        {
            float4 xp[RMAX4], xq[RMAX4];
            float pp = 0.0f, qq = 0.0f, pq = 0.0f;
#pragma unroll
            for (int j = 0; j < RMAX4; ++j)
            {
                if (j < rows4)
                {
                    const int i = lane + j * G;
                    xp[j] = ap[i];
                    xq[j] = aq[i];
                    pp += xp[j].x * xp[j].x + xp[j].y * xp[j].y + xp[j].z * xp[j].z + xp[j].w * xp[j].w;
                    qq += xq[j].x * xq[j].x + xq[j].y * xq[j].y + xq[j].z * xq[j].z + xq[j].w * xq[j].w;
                    pq += xp[j].x * xq[j].x + xp[j].y * xq[j].y + xp[j].z * xq[j].z + xp[j].w * xq[j].w;
                }
            }
            const double dpp = group_sum_f32<G>(pp);
            const double dqq = group_sum_f32<G>(qq);
            const double dpq = group_sum_f32<G>(pq);
            // FP64 off-diagonal test and rotation parameters.
            const double rhs = tol * sqrt(fmax(dpp, 0.0) * fmax(dqq, 0.0));
            const double apq = fabs(dpq);
            if (!(apq > rhs && apq > 1.0e-300))
                return 0;
            float4 yp[VMAX4], yq[VMAX4];
#pragma unroll
            for (int j = 0; j < VMAX4; ++j)
            {
                if (j < vrows4)
                {
                    const int i = lane + j * G;
                    yp[j] = vp[i];
                    yq[j] = vq[i];
                }
            }
            const double zeta = (dqq - dpp) / (2.0 * dpq);
            const double root = sqrt(1.0 + zeta * zeta);
            // One division: 1 / (|zeta| + root) equals both branches of the
            // sign-split form bit for bit (also for zeta = -0, which takes +).
            const double tmag = 1.0 / (fabs(zeta) + root);
            const double t = (zeta >= 0.0) ? tmag : -tmag;
            const double cd = 1.0 / sqrt(1.0 + t * t);
            const double sd = t * cd;
            const float c = static_cast<float>(cd), s = static_cast<float>(sd);
            if (s != 0.0f)
            {
#pragma unroll
                for (int j = 0; j < RMAX4; ++j)
                {
                    if (j < rows4)
                    {
                        const int i = lane + j * G;
                        float4 np, nq;
                        np.x = c * xp[j].x - s * xq[j].x;
                        np.y = c * xp[j].y - s * xq[j].y;
                        np.z = c * xp[j].z - s * xq[j].z;
                        np.w = c * xp[j].w - s * xq[j].w;
                        nq.x = s * xp[j].x + c * xq[j].x;
                        nq.y = s * xp[j].y + c * xq[j].y;
                        nq.z = s * xp[j].z + c * xq[j].z;
                        nq.w = s * xp[j].w + c * xq[j].w;
                        ap[i] = np;
                        aq[i] = nq;
                    }
                }
#pragma unroll
                for (int j = 0; j < VMAX4; ++j)
                {
                    if (j < vrows4)
                    {
                        const int i = lane + j * G;
                        float4 np, nq;
                        np.x = c * yp[j].x - s * yq[j].x;
                        np.y = c * yp[j].y - s * yq[j].y;
                        np.z = c * yp[j].z - s * yq[j].z;
                        np.w = c * yp[j].w - s * yq[j].w;
                        nq.x = s * yp[j].x + c * yq[j].x;
                        nq.y = s * yp[j].y + c * yq[j].y;
                        nq.z = s * yp[j].z + c * yq[j].z;
                        nq.w = s * yp[j].w + c * yq[j].w;
                        vp[i] = np;
                        vq[i] = nq;
                    }
                }
            }
            return 1;
        }

        template <int BS, int G, int RMAX4, int VMAX4>
        __global__ __launch_bounds__(BS) void mixed_sweeps_f32(MixedArgs args)
        {
            // float4 keeps the dynamic LDS block 16-byte aligned for the
            // row-quad accesses.
            extern __shared__ float4 lds_quads[];
            float *lds32 = reinterpret_cast<float *>(lds_quads);
            __shared__ int s_item;
            constexpr int GROUPS = BS / G;
            const int tid = threadIdx.x;
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
                float *A = lds32;
                float *V = lds32 + lda * n;
                const double *ag = args.a + w.a_off;
                for (int idx = tid; idx < lda * n; idx += BS)
                {
                    const int j = idx / lda, i = idx - j * lda;
                    A[idx] = i < m ? static_cast<float>(ag[i * n + j]) : 0.0f;
                }
                for (int idx = tid; idx < ldv * n; idx += BS)
                {
                    const int j = idx / ldv, i = idx - j * ldv;
                    V[idx] = (i == j) ? 1.0f : 0.0f;
                }
                __syncthreads();
                const double tol = fmax(args.eps, 8.0 * sqrt(static_cast<double>(m)) * 5.9604644775390625e-08);
                const int even = n + (n & 1);
                const int half = even >> 1;
                const int rounds = n < 2 ? 0 : even - 1;
                const int gid = tid / G, lane = tid & (G - 1);
                const int rows4 = lda / (4 * G), vrows4 = ldv / (4 * G);
                float4 *A4 = reinterpret_cast<float4 *>(A);
                float4 *V4 = reinterpret_cast<float4 *>(V);
                const int lda4 = lda / 4, ldv4 = ldv / 4;
                int executed = 0;
                for (int sweep = 0; sweep < args.max_sweeps; ++sweep)
                {
                    int any = 0;
                    for (int r = 0; r < rounds; ++r)
                    {
                        for (int k = gid; k < half; k += GROUPS)
                        {
                            const int p = dev::column_at(k, r, n, even);
                            const int q = dev::column_at(even - 1 - k, r, n, even);
                            if (p < 0 || q < 0)
                                continue;
                            any |= pair_f32<G, RMAX4, VMAX4>(A4 + p * lda4, A4 + q * lda4, V4 + p * ldv4,
                                                             V4 + q * ldv4, rows4, vrows4, lane, tol);
                        }
                        __syncthreads();
                    }
                    executed = sweep + 1;
                    if (!__syncthreads_or(any))
                        break;
                }
                // V32 -> packed column-major FP64 scratch.
                double *qg = args.q + w.v_off;
                for (int idx = tid; idx < n * n; idx += BS)
                {
                    const int j = idx / n, i = idx - j * n;
                    qg[idx] = static_cast<double>(V[j * ldv + i]);
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
        // FP64, in LDS.  Column j is normalised, then projected out of every
        // later column by lane groups of 16.
        template <int BS>
        __global__ __launch_bounds__(BS) void orthonormalize_f64(MixedArgs args)
        {
            extern __shared__ double ldsq[];
            __shared__ int s_item;
            __shared__ double red[BS / 64];
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
                double *qg = args.q + w.v_off;
                for (int idx = tid; idx < n * n; idx += BS)
                    Q[idx] = qg[idx];
                __syncthreads();
                for (int j = 0; j < n; ++j)
                {
                    double *qj = Q + j * n;
                    double part = 0.0;
                    for (int i = tid; i < n; i += BS)
                        part += qj[i] * qj[i];
                    part = dev::wave_sum(part);
                    if ((tid & 63) == 0)
                        red[tid >> 6] = part;
                    __syncthreads();
                    double nrm2 = red[0];
#pragma unroll
                    for (int wv = 1; wv < BS / 64; ++wv)
                        nrm2 += red[wv];
                    const double inv = nrm2 > 0.0 ? 1.0 / sqrt(nrm2) : 0.0;
                    __syncthreads();
                    for (int i = tid; i < n; i += BS)
                        qj[i] *= inv;
                    __syncthreads();
                    for (int k = j + 1 + gid; k < n; k += GROUPS)
                    {
                        double *qk = Q + k * n;
                        double d = 0.0;
                        for (int i = lane; i < n; i += G)
                            d += qj[i] * qk[i];
                        d = dev::group_sum_dpp<G>(d);
                        for (int i = lane; i < n; i += G)
                            qk[i] -= d * qj[i];
                    }
                    __syncthreads();
                }
                for (int idx = tid; idx < n * n; idx += BS)
                    qg[idx] = Q[idx];
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
            return reinterpret_cast<const void *>(&mixed_sweeps_f32<c.block, c.group, c.rmax4, c.vmax4>);
        }
        template <int I>
        void f32_launch_t(int grid, int lds_bytes, const MixedArgs &args, hipStream_t stream)
        {
            constexpr F32Spec c = kF32Classes[I];
            mixed_sweeps_f32<c.block, c.group, c.rmax4, c.vmax4><<<grid, c.block, lds_bytes, stream>>>(args);
        }
        using F32Launch = void (*)(int, int, const MixedArgs &, hipStream_t);
        const F32Launch kF32Launch[] = {f32_launch_t<0>, f32_launch_t<1>, f32_launch_t<2>, f32_launch_t<3>};
        const void *(*const kF32Symbol[])() = {f32_symbol_t<0>, f32_symbol_t<1>, f32_symbol_t<2>, f32_symbol_t<3>};
        static_assert(sizeof(kF32Launch) / sizeof(kF32Launch[0]) == kF32Count);
    } // namespace

    constexpr int kOrthoBlock = 256;
    constexpr int kGemmBlock = 256;

    int64_t f32_class_lds_bytes(int cls, int64_t m, int64_t n)
    {
        const int g = kF32Classes[cls].group;
        return (padded4(static_cast<int>(m), g) * n + padded4(static_cast<int>(n), g) * n) * 4;
    }

    bool f32_class_fits(int cls, int64_t m, int64_t n)
    {
        if (cls < 0 || cls >= kF32Count)
            return false;
        const F32Spec &c = kF32Classes[cls];
        const int64_t half = (n + 1) / 2;
        const int64_t groups = c.block / c.group;
        const int64_t passes = c.block == 64 ? 1 : 2;
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
        return n * n * 8;
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
