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
// keeps J, K (then N) and W = U N + gU diag(s)^-1 in LDS; U, gU, V, gV are
// read from HBM (they stay in L2 across the products).  The result is
// invariant to a common permutation of (U, s, V) columns, so the unsorted
// column order of the solver needs no sorting.  Distinct nonzero singular
// values are required (F and 1/s are unbounded otherwise), as for any SVD
// adjoint.
#include "jsvd_kernels.hpp"

namespace jsvd
{
    namespace
    {
        template <typename T, int BS>
        __global__ __launch_bounds__(BS) void svd_backward(const WorkItem *__restrict__ items, int count,
                                                           const T *__restrict__ U, const T *__restrict__ S,
                                                           const T *__restrict__ V, const T *__restrict__ gU,
                                                           const T *__restrict__ gS, const T *__restrict__ gV,
                                                           T *__restrict__ gA)
        {
            extern __shared__ double2 lds_raw[];
            T *lds = reinterpret_cast<T *>(lds_raw);
            const int item = blockIdx.x;
            if (item >= count)
                return;
            const WorkItem w = items[item];
            const int m = w.m, n = w.n, tid = threadIdx.x;
            const T *u = U + w.a_off, *gu = gU ? gU + w.a_off : nullptr;
            const T *v = V + w.v_off, *gv = gV ? gV + w.v_off : nullptr;
            const T *s = S + w.s_off, *gs = gS ? gS + w.s_off : nullptr;
            T *J = lds;         // n x n row-major, becomes N
            T *K = lds + n * n; // n x n row-major
            T *W = K + n * n;   // m x n row-major
            // J = U^T gU and K = V^T gV.
            for (int idx = tid; idx < n * n; idx += BS)
            {
                const int i = idx / n, j = idx - (idx / n) * n;
                T jacc = 0, kacc = 0;
                if (gu)
                    for (int k = 0; k < m; ++k)
                        jacc += u[k * n + i] * gu[k * n + j];
                if (gv)
                    for (int k = 0; k < n; ++k)
                        kacc += v[k * n + i] * gv[k * n + j];
                J[idx] = jacc;
                K[idx] = kacc;
            }
            __syncthreads();
            // N is staged in W's space (m*n >= n*n) because entry (i,j) reads
            // J_ji and K_ji, then copied over J.
            for (int idx = tid; idx < n * n; idx += BS)
            {
                const int i = idx / n, j = idx - (idx / n) * n;
                const T si = s[i], sj = s[j];
                T val;
                if (i == j)
                    val = (gs ? gs[i] : T(0)) - J[idx] / si;
                else
                {
                    const T f = T(1) / (sj * sj - si * si);
                    val = ((J[idx] - J[j * n + i]) * sj + si * (K[idx] - K[j * n + i])) * f - J[idx] / sj;
                }
                W[idx] = val;
            }
            __syncthreads();
            for (int idx = tid; idx < n * n; idx += BS)
                J[idx] = W[idx];
            __syncthreads();
            // W = U N + gU diag(s)^-1.
            for (int idx = tid; idx < m * n; idx += BS)
            {
                const int i = idx / n, j = idx - (idx / n) * n;
                T acc = gu ? gu[idx] / s[j] : T(0);
                for (int k = 0; k < n; ++k)
                    acc += u[i * n + k] * J[k * n + j];
                W[idx] = acc;
            }
            __syncthreads();
            // gA = W V^T.
            T *ga = gA + w.a_off;
            for (int idx = tid; idx < m * n; idx += BS)
            {
                const int i = idx / n, j = idx - (idx / n) * n;
                T acc = 0;
                for (int k = 0; k < n; ++k)
                    acc += W[i * n + k] * v[j * n + k];
                ga[idx] = acc;
            }
        }
    } // namespace

    constexpr int kBackwardBlock = 256;

    int64_t backward_lds_bytes(int64_t m, int64_t n, int elem)
    {
        return (2 * n * n + m * n) * elem;
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
