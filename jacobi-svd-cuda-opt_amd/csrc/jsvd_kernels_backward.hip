// jsvd_kernels_backward.hip -- adjoint of the thin SVD A = U diag(s) V^T
// (m >= n, V square), no reference counterpart.
//
// For upstream gradients gU (m x n), gs (n), gV (n x n):
//   J = U^T gU,  K = V^T gV,  F_ij = 1 / (s_j^2 - s_i^2) (i != j)
//   M = [(J - J^T) o F] diag(s) + diag(s) [(K - K^T) o F] + diag(gs)
//   gA = U M V^T + (I - U U^T) gU diag(s)^-1 V^T
// which is rewritten as gA = (U N + gU diag(s)^-1) V^T with
//   N = M - J diag(s)^-1
// so that only four small products are needed.  One workgroup of four
// waves per matrix forms them on FP64 MFMA (v_mfma_f64_16x16x4f64) in
// 16 x 16 tiles, with m and n padded to multiples of 16 by zeros:
//  1. J and K: a wave takes a row of output tiles (one 16-column block of
//     U or V against up to four 16-column blocks of gU or gV), loading
//     32 rows of every operand into registers before the eight MFMAs that
//     use them, so a whole chunk of loads is in flight at once;
//  2. N from J, K, s, gs, elementwise in LDS;
//  3. per 16-row block of A, by one wave: W = U N + gU diag(s)^-1 (16 x n',
//     kept in the wave's own LDS slice), then at once gA = W V^T for that
//     block, so W never exists whole and no barrier separates the two.
// LDS: N (n'^2 doubles) and a region of max(n', 64) x (n' + 1) doubles that
// holds K while N is formed and then the four waves' 16-row W slices; it
// does not depend on m (66 KB at n = 64: two workgroups per CU).  FP32
// inputs are widened to FP64 on load and the result rounded once: F and
// 1/s amplify the rounding of every product by up to the inverse relative
// gap of the singular values, which FP32 products would carry straight into
// the gradient, and the kernel is bound by load latency, not by MFMA
// throughput.  The result is invariant to a common permutation of (U, s, V)
// columns, so the unsorted column order of the solver needs no sorting.
// Distinct nonzero singular values are required (F and 1/s are unbounded
// otherwise), as for any SVD adjoint.
#include "jsvd_kernels.hpp"

#include <algorithm>

namespace jsvd
{
    namespace
    {
        typedef double f64x4 __attribute__((ext_vector_type(4)));

        constexpr int kChunk = 8; // MFMAs per register chunk: 32 rows of the summation index
        constexpr int kTileGroup = 4; // output tiles per operand chunk

        template <typename T>
        __device__ __forceinline__ double load_or_zero(const T *p, int64_t idx, bool ok)
        {
            return ok ? static_cast<double>(p[idx]) : 0.0;
        }

        __device__ __forceinline__ void wave_fence()
        {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }

        template <typename T, int BS>
        __global__ __launch_bounds__(BS) void svd_backward(const WorkItem *__restrict__ items, int count,
                                                           const T *__restrict__ U, const T *__restrict__ S,
                                                           const T *__restrict__ V, const T *__restrict__ gU,
                                                           const T *__restrict__ gS, const T *__restrict__ gV,
                                                           T *__restrict__ gA)
        {
            constexpr int WAVES = BS / 64;
            extern __shared__ double2 lds_raw[];
            double *lds = reinterpret_cast<double *>(lds_raw);
            const int item = blockIdx.x;
            if (item >= count)
                return;
            const WorkItem w = items[item];
            const int m = w.m, n = w.n;
            const int np = (n + 15) & ~15, mp = (m + 15) & ~15;
            const int tn = np / 16, tm = mp / 16;
            const int ld = np + 1; // odd stride of K and W: A-operand reads spread over the banks
            const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, i = l & 15, g = l >> 4;
            const T *u = U + w.a_off, *gu = gU ? gU + w.a_off : nullptr;
            const T *v = V + w.v_off, *gv = gV ? gV + w.v_off : nullptr;
            const T *s = S + w.s_off, *gs = gS ? gS + w.s_off : nullptr;
            double *Nm = lds;           // np x np row-major: J, then N
            double *Rm = lds + np * np; // K (np x ld), then the waves' W slices (16 x ld each)
            // MFMA operand element (l % 16, l / 16) of A and (l / 16, l % 16) of
            // B; result item r of lane l is tile row l/16 + 4r, column l % 16.

            // 1. J = U^T gU (tile rows 0 .. tn-1) and K = V^T gV (tn .. 2tn-1).
            for (int job = wave; job < 2 * tn; job += WAVES)
            {
                const bool is_k = job >= tn;
                const int ti = is_k ? job - tn : job;
                const T *a = is_k ? v : u, *b = is_k ? gv : gu;
                const int rows = is_k ? n : m;
                double *dst = is_k ? Rm : Nm;
                const int ldd = is_k ? ld : np;
                const int ca = ti * 16 + i;
                for (int tj0 = 0; tj0 < tn; tj0 += kTileGroup)
                {
                    f64x4 acc[kTileGroup];
#pragma unroll
                    for (int t = 0; t < kTileGroup; ++t)
                        acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
                    if (b)
                    {
                        for (int k0 = 0; k0 < rows; k0 += 4 * kChunk)
                        {
                            double av[kChunk], bv[kTileGroup][kChunk];
#pragma unroll
                            for (int c = 0; c < kChunk; ++c)
                            {
                                const int k = k0 + 4 * c + g;
                                av[c] = load_or_zero(a, int64_t(k) * n + ca, k < rows && ca < n);
#pragma unroll
                                for (int t = 0; t < kTileGroup; ++t)
                                {
                                    const int cb = (tj0 + t) * 16 + i;
                                    bv[t][c] = load_or_zero(b, int64_t(k) * n + cb, k < rows && cb < n);
                                }
                            }
#pragma unroll
                            for (int t = 0; t < kTileGroup; ++t)
                                if (tj0 + t < tn)
                                {
#pragma unroll
                                    for (int c = 0; c < kChunk; ++c)
                                        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[c], bv[t][c], acc[t], 0, 0, 0);
                                }
                        }
                    }
#pragma unroll
                    for (int t = 0; t < kTileGroup; ++t)
                        if (tj0 + t < tn)
                        {
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                dst[(ti * 16 + g + 4 * r) * ldd + (tj0 + t) * 16 + i] = acc[t][r];
                        }
                }
            }
            __syncthreads();
            // 2. N in place of J: the thread of (r, c), r <= c, also writes
            // (c, r); padding rows and columns become zero.
            for (int idx = threadIdx.x; idx < np * np; idx += BS)
            {
                const int r = idx / np, c = idx - (idx / np) * np;
                if (r > c)
                    continue;
                double nrc = 0.0, ncr = 0.0;
                if (c < n)
                {
                    const double sr = static_cast<double>(s[r]);
                    if (r == c)
                        nrc = (gs ? static_cast<double>(gs[r]) : 0.0) - Nm[r * np + r] / sr;
                    else
                    {
                        const double sc = static_cast<double>(s[c]);
                        const double f = 1.0 / (sc * sc - sr * sr); // F_rc; F_cr = -F_rc
                        const double jrc = Nm[r * np + c], jcr = Nm[c * np + r];
                        const double krc = Rm[r * ld + c], kcr = Rm[c * ld + r];
                        nrc = ((jrc - jcr) * sc + sr * (krc - kcr)) * f - jrc / sc;
                        ncr = ((jcr - jrc) * sr + sc * (kcr - krc)) * (-f) - jcr / sr;
                    }
                }
                Nm[r * np + c] = nrc;
                if (r != c)
                    Nm[c * np + r] = ncr;
            }
            __syncthreads();
            // 3. Per 16-row block of A: W = U N + gU diag(s)^-1, then gA = W V^T.
            double *Wb = Rm + wave * 16 * ld;
            T *ga = gA + w.a_off;
            for (int ti = wave; ti < tm; ti += WAVES)
            {
                const int ra = ti * 16 + i; // this lane's A-operand row of U
                for (int tj0 = 0; tj0 < tn; tj0 += kTileGroup)
                {
                    f64x4 acc[kTileGroup];
#pragma unroll
                    for (int t = 0; t < kTileGroup; ++t)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                        {
                            const int row = ti * 16 + g + 4 * r, col = (tj0 + t) * 16 + i;
                            acc[t][r] = (gu && row < m && col < n)
                                            ? static_cast<double>(gu[int64_t(row) * n + col]) /
                                                  static_cast<double>(s[col])
                                            : 0.0;
                        }
                    for (int k0 = 0; k0 < np; k0 += 4 * kChunk)
                    {
                        double av[kChunk], bv[kTileGroup][kChunk];
#pragma unroll
                        for (int c = 0; c < kChunk; ++c)
                        {
                            const int k = k0 + 4 * c + g;
                            av[c] = load_or_zero(u, int64_t(ra) * n + k, ra < m && k < n);
#pragma unroll
                            for (int t = 0; t < kTileGroup; ++t)
                                bv[t][c] = (k < np && tj0 + t < tn) ? Nm[k * np + (tj0 + t) * 16 + i] : 0.0;
                        }
#pragma unroll
                        for (int t = 0; t < kTileGroup; ++t)
                            if (tj0 + t < tn)
                            {
#pragma unroll
                                for (int c = 0; c < kChunk; ++c)
                                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[c], bv[t][c], acc[t], 0, 0, 0);
                            }
                    }
#pragma unroll
                    for (int t = 0; t < kTileGroup; ++t)
                        if (tj0 + t < tn)
                        {
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                Wb[(g + 4 * r) * ld + (tj0 + t) * 16 + i] = acc[t][r];
                        }
                }
                wave_fence(); // the slice is written by every lane and read across lanes below
                for (int tj0 = 0; tj0 < tn; tj0 += kTileGroup)
                {
                    f64x4 acc[kTileGroup];
#pragma unroll
                    for (int t = 0; t < kTileGroup; ++t)
                        acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
                    for (int k0 = 0; k0 < np; k0 += 4 * kChunk)
                    {
                        double av[kChunk], bv[kTileGroup][kChunk];
#pragma unroll
                        for (int c = 0; c < kChunk; ++c)
                        {
                            const int k = k0 + 4 * c + g;
                            av[c] = k < np ? Wb[i * ld + k] : 0.0;
#pragma unroll
                            for (int t = 0; t < kTileGroup; ++t)
                            {
                                const int cv = (tj0 + t) * 16 + i;
                                bv[t][c] = load_or_zero(v, int64_t(cv) * n + k, cv < n && k < n);
                            }
                        }
#pragma unroll
                        for (int t = 0; t < kTileGroup; ++t)
                            if (tj0 + t < tn)
                            {
#pragma unroll
                                for (int c = 0; c < kChunk; ++c)
                                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[c], bv[t][c], acc[t], 0, 0, 0);
                            }
                    }
#pragma unroll
                    for (int t = 0; t < kTileGroup; ++t)
                    {
                        const int c = (tj0 + t) * 16 + i;
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                        {
                            const int row = ti * 16 + g + 4 * r;
                            if (tj0 + t < tn && row < m && c < n)
                                ga[int64_t(row) * n + c] = static_cast<T>(acc[t][r]);
                        }
                    }
                }
                wave_fence(); // the slice's reads are done before the next block overwrites it
            }
        }
    } // namespace

    constexpr int kBackwardBlock = 256;

    int64_t backward_lds_bytes(int64_t /*m*/, int64_t n, int /*elem*/)
    {
        const int64_t np = (n + 15) / 16 * 16;
        return (np * np + std::max<int64_t>(np, 16 * (kBackwardBlock / 64)) * (np + 1)) * 8;
    }

    const void *backward_symbol(int elem)
    {
        return elem == 8 ? reinterpret_cast<const void *>(&svd_backward<double, kBackwardBlock>)
                         : reinterpret_cast<const void *>(&svd_backward<float, kBackwardBlock>);
    }

    hipError_t launch_backward_f64(const WorkItem *items, int count, int lds_bytes, const double *u, const double *s,
                                   const double *v, const double *gu, const double *gs, const double *gv, double *ga,
                                   hipStream_t stream)
    {
        if (count == 0)
            return hipSuccess;
        svd_backward<double, kBackwardBlock><<<count, kBackwardBlock, lds_bytes, stream>>>(items, count, u, s, v, gu,
                                                                                          gs, gv, ga);
        return hipGetLastError();
    }

    hipError_t launch_backward_f32(const WorkItem *items, int count, int lds_bytes, const float *u, const float *s,
                                   const float *v, const float *gu, const float *gs, const float *gv, float *ga,
                                   hipStream_t stream)
    {
        if (count == 0)
            return hipSuccess;
        svd_backward<float, kBackwardBlock><<<count, kBackwardBlock, lds_bytes, stream>>>(items, count, u, s, v, gu,
                                                                                         gs, gv, ga);
        return hipGetLastError();
    }
} // namespace jsvd
