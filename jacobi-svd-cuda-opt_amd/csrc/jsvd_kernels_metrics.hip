// jsvd_kernels_metrics.hip -- per-matrix error metrics of a computed SVD
// (no reference counterpart: the reference reports sweeps only,
// src/interfaces/cli_support.cu:427-464).
//
// For each matrix of a packed batch, from A, U (m x n), sigma (n) and
// V (n x n), all row-major:
//   orth_u   = ||U^T U - D||_F,  D = diag(sigma_j > eps)  (U columns with
//              sigma_j <= eps are zero by the U rule, so this is the
//              orthogonality of the kept columns)
//   orth_v   = ||V^T V - I||_F
//   residual = ||A - U diag(sigma) V^T||_F / ||A||_F   (0 when A = 0)
// Every product is cut into 32 x 32 output tiles; a tile is one workgroup of
// 256 threads (four outputs per thread) that streams 32-wide slices of its
// operands through padded LDS tiles and writes the sum of squared deviations
// of its outputs (and, for residual tiles, of A) to a partial slot.  A
// second kernel adds each matrix's partials in tile order, so the result
// does not depend on scheduling.
#include "jsvd_kernels.hpp"

namespace jsvd
{
    namespace
    {
        constexpr int kTile = 32;

        // Sum over the workgroup (256 threads) in a fixed order.
        __device__ __forceinline__ double block_sum_256(double x, double *red)
        {
            for (int o = 32; o > 0; o >>= 1)
                x += __shfl_down(x, o, 64);
            const int wave = threadIdx.x >> 6;
            if ((threadIdx.x & 63) == 0)
                red[wave] = x;
            __syncthreads();
            const double total = (red[0] + red[1]) + (red[2] + red[3]);
            __syncthreads();
            return total;
        }

        __global__ __launch_bounds__(256) void metric_tiles_f64(const MetricTile *__restrict__ tiles,
                                                                const WorkItem *__restrict__ mats,
                                                                const double *__restrict__ A,
                                                                const double *__restrict__ U,
                                                                const double *__restrict__ S,
                                                                const double *__restrict__ V, double eps,
                                                                double *__restrict__ partial)
        {
            __shared__ double xs[kTile][kTile + 1];
            __shared__ double ys[kTile][kTile + 1];
            __shared__ double red[4];
            const MetricTile t = tiles[blockIdx.x];
            const WorkItem w = mats[t.matrix];
            const int64_t m = w.m, n = w.n;
            const double *u = U + w.a_off, *v = V + w.v_off, *s = S + w.s_off, *a = A + w.a_off;
            const int tid = threadIdx.x;
            const int ci = tid & 31, cj = tid >> 5; // this thread's outputs: (ci, cj + 8 q)
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            // kind 0: U^T U (k over m rows); kind 1: V^T V (k over n rows);
            // kind 2: U diag(s) V^T (k over n columns).
            const int64_t depth = t.kind == 0 ? m : n;
            const int64_t rows_out = t.kind == 2 ? m : n;
            const int64_t i0 = int64_t(t.ti) * kTile, j0 = int64_t(t.tj) * kTile;
            for (int64_t k0 = 0; k0 < depth; k0 += kTile)
            {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                {
                    // Load element (row r, column c) of the 32 x 32 slices:
                    // xs[kk][i] and ys[kk][j] with kk the summation index.
                    const int e = tid + 256 * q;
                    const int hi = e >> 5, lo = e & 31;
                    double x = 0.0, y = 0.0;
                    if (t.kind == 2)
                    {
                        // lo = kk (contiguous along a row of U and of V).
                        const int64_t kk = k0 + lo;
                        if (kk < n && i0 + hi < m)
                            x = u[(i0 + hi) * n + kk] * s[kk];
                        if (kk < n && j0 + hi < n)
                            y = v[(j0 + hi) * n + kk];
                        xs[lo][hi] = x;
                        ys[lo][hi] = y;
                    }
                    else
                    {
                        // lo = output column (contiguous along a row), hi = kk.
                        const double *src = t.kind == 0 ? u : v;
                        const int64_t kk = k0 + hi;
                        if (kk < depth && i0 + lo < n)
                            x = src[kk * n + i0 + lo];
                        if (kk < depth && j0 + lo < n)
                            y = src[kk * n + j0 + lo];
                        xs[hi][lo] = x;
                        ys[hi][lo] = y;
                    }
                }
                __syncthreads();
#pragma unroll 8
                for (int kk = 0; kk < kTile; ++kk)
                {
                    const double x = xs[kk][ci];
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        acc[q] = fma(x, ys[kk][cj + 8 * q], acc[q]);
                }
                __syncthreads();
            }
            double dev2 = 0.0, ref2 = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                const int64_t i = i0 + ci, j = j0 + cj + 8 * q;
                if (i >= rows_out || j >= n)
                    continue;
                double target;
                if (t.kind == 2)
                {
                    target = a[i * n + j];
                    ref2 = fma(target, target, ref2);
                }
                else if (i != j)
                    target = 0.0;
                else
                    target = (t.kind == 1 || s[j] > eps) ? 1.0 : 0.0;
                const double d = acc[q] - target;
                dev2 = fma(d, d, dev2);
            }
            dev2 = block_sum_256(dev2, red);
            ref2 = block_sum_256(ref2, red);
            if (tid == 0)
            {
                partial[2 * int64_t(blockIdx.x)] = dev2;
                partial[2 * int64_t(blockIdx.x) + 1] = ref2;
            }
        }

        // One thread per matrix: its tiles are first[k] .. first[k+1]-1 in
        // the order U^T U, V^T V, residual.
        __global__ void metric_finish_f64(const MetricTile *__restrict__ tiles, const int64_t *__restrict__ first,
                                          int count, const double *__restrict__ partial, double *__restrict__ out)
        {
            const int k = blockIdx.x * blockDim.x + threadIdx.x;
            if (k >= count)
                return;
            double su = 0.0, sv = 0.0, sr = 0.0, sa = 0.0;
            for (int64_t t = first[k]; t < first[k + 1]; ++t)
            {
                const double d = partial[2 * t];
                switch (tiles[t].kind)
                {
                case 0: su += d; break;
                case 1: sv += d; break;
                default:
                    sr += d;
                    sa += partial[2 * t + 1];
                    break;
                }
            }
            out[3 * int64_t(k)] = sqrt(su);
            out[3 * int64_t(k) + 1] = sqrt(sv);
            out[3 * int64_t(k) + 2] = sa > 0.0 ? sqrt(sr) / sqrt(sa) : sqrt(sr);
        }
    } // namespace

    hipError_t launch_metrics_f64(const MetricTile *tiles, int64_t tile_count, const int64_t *first,
                                  const WorkItem *mats, int count, const double *a, const double *u,
                                  const double *sigma, const double *v, double eps, double *partial, double *out,
                                  hipStream_t stream)
    {
        if (count <= 0)
            return hipSuccess;
        if (tile_count > 0)
            metric_tiles_f64<<<dim3(static_cast<unsigned>(tile_count)), 256, 0, stream>>>(tiles, mats, a, u, sigma,
                                                                                         v, eps, partial);
        metric_finish_f64<<<(count + 255) / 256, 256, 0, stream>>>(tiles, first, count, partial, out);
        return hipGetLastError();
    }
} // namespace jsvd
