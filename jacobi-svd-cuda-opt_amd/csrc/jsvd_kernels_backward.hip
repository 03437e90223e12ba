// jsvd_kernels_backward.hip -- adjoint of the thin SVD A = U diag(s) V^T
// (m >= n, V square), no reference counterpart.
//
// For upstream gradients gU (m x n), gs (n), gV (n x n):
//   J = U^T gU,  K = V^T gV,  F_ij = 1 / (s_j^2 - s_i^2) (i != j)
//   M = [(J - J^T) o F] diag(s) + diag(s) [(K - K^T) o F] + diag(gs)
//   gA = U M V^T + (I - U U^T) gU diag(s)^-1 V^T
// which is rewritten as gA = (U N + gU diag(s)^-1) V^T with
//   N = M - J diag(s)^-1
// so that only four small products are needed.  One workgroup per matrix
// forms them on FP64 MFMA (v_mfma_f64_16x16x4f64) in 16 x 16 tiles, the four
// waves taking tiles round-robin; operands are read straight from HBM/L2
// (U, gU, V, gV) or from LDS (N, W), with zero padding of m and n up to
// multiples of 16.  LDS holds N (np x np) and W (mp x (np + 1), which holds K
// while N is formed): (np^2 + mp (np + 1)) doubles, np/mp = n/m rounded up to
// 16.  FP32 inputs are widened to FP64 on load and the result rounded once.
// The result is invariant to a common permutation of (U, s, V) columns, so
// the unsorted column order of the solver needs no sorting.  Distinct
// nonzero singular values are required (F and 1/s are unbounded otherwise),
// as for any SVD adjoint.
#include "jsvd_kernels.hpp"

namespace jsvd
{
    namespace
    {
        typedef double f64x4 __attribute__((ext_vector_type(4)));

        template <typename T>
        __device__ __forceinline__ double load_or_zero(const T *p, int64_t idx, bool ok)
        {
            return ok ? static_cast<double>(p[idx]) : 0.0;
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
            const int ldw = np + 1; // odd stride: the A-operand reads of W spread over the banks
            const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, i = l & 15, g = l >> 4;
            const T *u = U + w.a_off, *gu = gU ? gU + w.a_off : nullptr;
            const T *v = V + w.v_off, *gv = gV ? gV + w.v_off : nullptr;
            const T *s = S + w.s_off, *gs = gS ? gS + w.s_off : nullptr;
            double *Nm = lds;           // np x np row-major: J, then N
            double *Wm = lds + np * np; // mp x ldw row-major: K (first np rows), then W
            // MFMA operand element (l % 16, l / 16) of A and (l / 16, l % 16) of
            // B; result item r of lane l is tile row l/16 + 4r, column l % 16.

            // J = U^T gU and K = V^T gV.
            for (int t = wave; t < 2 * tn * tn; t += WAVES)
            {
                const bool is_k = t >= tn * tn;
                const int tt = is_k ? t - tn * tn : t;
                const int ti = tt / tn, tj = tt - (tt / tn) * tn;
                const T *a = is_k ? v : u, *b = is_k ? gv : gu;
                const int rows = is_k ? n : m;
                f64x4 acc = {0.0, 0.0, 0.0, 0.0};
                if (b)
                {
                    const int ca = ti * 16 + i, cb = tj * 16 + i;
#pragma unroll 4
                    for (int k0 = 0; k0 < rows; k0 += 4)
                    {
                        const int k = k0 + g;
                        const double av = load_or_zero(a, int64_t(k) * n + ca, k < rows && ca < n);
                        const double bv = load_or_zero(b, int64_t(k) * n + cb, k < rows && cb < n);
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
                    }
                }
                double *dst = is_k ? Wm : Nm;
                const int ld = is_k ? ldw : np;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    dst[(ti * 16 + g + 4 * r) * ld + tj * 16 + i] = acc[r];
            }
            __syncthreads();
            // N in place of J: the thread of (r, c), r <= c, also writes (c, r);
            // padding rows and columns become zero.
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
                        const double krc = Wm[r * ldw + c], kcr = Wm[c * ldw + r];
                        nrc = ((jrc - jcr) * sc + sr * (krc - kcr)) * f - jrc / sc;
                        ncr = ((jcr - jrc) * sr + sc * (kcr - krc)) * (-f) - jcr / sr;
                    }
                }
                Nm[r * np + c] = nrc;
                if (r != c)
                    Nm[c * np + r] = ncr;
            }
            __syncthreads();
            // W = U N + gU diag(s)^-1.
            for (int t = wave; t < tm * tn; t += WAVES)
            {
                const int ti = t / tn, tj = t - (t / tn) * tn;
                f64x4 acc;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                {
                    const int row = ti * 16 + g + 4 * r, col = tj * 16 + i;
                    acc[r] = (gu && row < m && col < n)
                                 ? static_cast<double>(gu[int64_t(row) * n + col]) / static_cast<double>(s[col])
                                 : 0.0;
                }
                const int row = ti * 16 + i;
#pragma unroll 4
                for (int k0 = 0; k0 < np; k0 += 4)
                {
                    const int k = k0 + g;
                    const double av = load_or_zero(u, int64_t(row) * n + k, row < m && k < n);
                    const double bv = Nm[k * np + tj * 16 + i];
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    Wm[(ti * 16 + g + 4 * r) * ldw + tj * 16 + i] = acc[r];
            }
            __syncthreads();
            // gA = W V^T.
            T *ga = gA + w.a_off;
            for (int t = wave; t < tm * tn; t += WAVES)
            {
                const int ti = t / tn, tj = t - (t / tn) * tn;
                f64x4 acc = {0.0, 0.0, 0.0, 0.0};
                const int c = tj * 16 + i;
#pragma unroll 4
                for (int k0 = 0; k0 < np; k0 += 4)
                {
                    const int k = k0 + g;
                    const double av = Wm[(ti * 16 + i) * ldw + k];
                    const double bv = load_or_zero(v, int64_t(c) * n + k, c < n && k < n);
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r)
                {
                    const int row = ti * 16 + g + 4 * r;
                    if (row < m && c < n)
                        ga[int64_t(row) * n + c] = static_cast<T>(acc[r]);
                }
            }
        }
    } // namespace

    constexpr int kBackwardBlock = 256;

    int64_t backward_lds_bytes(int64_t m, int64_t n, int /*elem*/)
    {
        const int64_t np = (n + 15) / 16 * 16, mp = (m + 15) / 16 * 16;
        return (np * np + mp * (np + 1)) * 8;
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
