// jsvd_kernels_f64.hip -- plain FP64 one-sided Jacobi kernels for gfx950.
//
// Two sweep kernels cover every shape:
//  * lds_batch_f64: a persistent work-list kernel.  Each workgroup claims
//    whole matrices from a device work-list (sorted by cost, largest first)
//    and keeps A and V resident in LDS for the entire solve: load once,
//    all sweeps with workgroup barriers only, convergence decided on-device,
//    U / Sigma / V written once.  Pairs of a round are spread over aligned
//    lane groups; each group reduces its three dots with wave64 shuffles and
//    applies the rotation to the rows it holds in registers.
//  * persistent_sweeps_f64: matrices too large for LDS.  A cooperative
//    persistent kernel with a grid barrier per round; one workgroup per
//    column pair, columns held in registers between the dot and the update,
//    and the sweep loop with its convergence flag kept on the device.
//
// Arithmetic per pair follows the reference kernel
// (src/domain/jacobi_rotation_kernels.cuh:276-353): dots over the rows,
// Rutishauser rotation with the |pq| > eps*sqrt(pp*qq) test, then
// a_p <- c a_p - s a_q, a_q <- s a_p + c a_q on A and on V.  Pairs come in
// the reference's circle order, so the final column order (and with it the
// unsorted sigma) is the reference's.
#include "jsvd_device.hpp"
#include "jsvd_kernels.hpp"

#include <algorithm>

namespace jsvd
{
    namespace
    {
        __device__ __forceinline__ int pow2_floor(int x)
        {
            return x <= 1 ? 1 : (1 << (31 - __clz(x)));
        }
        __device__ __forceinline__ int pow2_ceil(int x)
        {
            return x <= 1 ? 1 : (1 << (32 - __clz(x - 1)));
        }

        // One column pair held in LDS: G lanes share the rows.  Returns 1 when
        // the pair counts as rotated.
        template <int RMAX>
        __device__ __forceinline__ int pair_lds_f64(double *__restrict__ ap, double *__restrict__ aq,
                                                    double *__restrict__ vp, double *__restrict__ vq, int m, int n,
                                                    int G, int lane, double eps)
        {
            double pp = 0.0, qq = 0.0, pq = 0.0;
            const int rows = (m + G - 1) / G; // uniform over the group
            if (rows <= RMAX)
            {
                double xp[RMAX], xq[RMAX];
#pragma unroll
                for (int j = 0; j < RMAX; ++j)
                {
                    if (j < rows)
                    {
                        const int i = lane + j * G;
                        const bool in = i < m;
                        xp[j] = in ? ap[i] : 0.0;
                        xq[j] = in ? aq[i] : 0.0;
                        pp += xp[j] * xp[j];
                        qq += xq[j] * xq[j];
                        pq += xp[j] * xq[j];
                    }
                }
                pp = dev::group_sum(pp, G);
                qq = dev::group_sum(qq, G);
                pq = dev::group_sum(pq, G);
                double c, s;
                const bool rot = dev::rotation_refined(pp, qq, pq, eps, c, s);
                if (s != 0.0)
                {
#pragma unroll
                    for (int j = 0; j < RMAX; ++j)
                    {
                        if (j < rows)
                        {
                            const int i = lane + j * G;
                            if (i < m)
                            {
                                ap[i] = c * xp[j] - s * xq[j];
                                aq[i] = s * xp[j] + c * xq[j];
                            }
                        }
                    }
                    for (int i = lane; i < n; i += G)
                    {
                        const double x = vp[i], y = vq[i];
                        vp[i] = c * x - s * y;
                        vq[i] = s * x + c * y;
                    }
                }
                return rot ? 1 : 0;
            }
            for (int i = lane; i < m; i += G)
            {
                const double x = ap[i], y = aq[i];
                pp += x * x;
                qq += y * y;
                pq += x * y;
            }
            pp = dev::group_sum(pp, G);
            qq = dev::group_sum(qq, G);
            pq = dev::group_sum(pq, G);
            double c, s;
            const bool rot = dev::rotation_refined(pp, qq, pq, eps, c, s);
            if (s != 0.0)
            {
                for (int i = lane; i < m; i += G)
                {
                    const double x = ap[i], y = aq[i];
                    ap[i] = c * x - s * y;
                    aq[i] = s * x + c * y;
                }
                for (int i = lane; i < n; i += G)
                {
                    const double x = vp[i], y = vq[i];
                    vp[i] = c * x - s * y;
                    vq[i] = s * x + c * y;
                }
            }
            return rot ? 1 : 0;
        }

        template <int BS, int RMAX>
        __device__ void solve_lds_f64(const WorkItem &w, double *lds, const double *__restrict__ a,
                                      double *__restrict__ u, double *__restrict__ sigma, double *__restrict__ v,
                                      int32_t *__restrict__ sweeps, int32_t *__restrict__ stop, double eps,
                                      int max_sweeps)
        {
            const int m = w.m, n = w.n, tid = threadIdx.x;
            double *A = lds;                         // column-major, leading dimension m
            double *V = lds + (int64_t)m * n;        // column-major n x n
            double *S = V + (int64_t)n * n;          // sigma scratch
            const double *ag = a + w.a_off;
            const int mn = m * n, nn = n * n;
            for (int idx = tid; idx < mn; idx += BS)
            {
                const int i = idx / n, j = idx - (idx / n) * n;
                A[j * m + i] = ag[idx];
            }
            for (int idx = tid; idx < nn; idx += BS)
                V[idx] = (idx % (n + 1) == 0) ? 1.0 : 0.0;
            __syncthreads();

            const int even = n + (n & 1);
            const int half = even >> 1;
            const int rounds = n < 2 ? 0 : even - 1;
            int G = pow2_floor(BS / (half > 0 ? half : 1));
            G = min(G, 64);
            G = min(G, pow2_ceil(m));
            const int groups = BS / G;
            const int gid = tid / G, lane = tid & (G - 1);

            int executed = 0, clean = 0;
            for (int sweep = 0; sweep < max_sweeps; ++sweep)
            {
                int any = 0;
                for (int r = 0; r < rounds; ++r)
                {
                    for (int k = gid; k < half; k += groups)
                    {
                        const int p = dev::column_at(k, r, n, even);
                        const int q = dev::column_at(even - 1 - k, r, n, even);
                        if (p < 0 || q < 0)
                            continue;
                        any |= pair_lds_f64<RMAX>(A + p * m, A + q * m, V + p * n, V + q * n, m, n, G, lane, eps);
                    }
                    __syncthreads();
                }
                executed = sweep + 1;
                if (!__syncthreads_or(any))
                {
                    clean = 1;
                    break;
                }
            }

            // U / Sigma epilogue (reference src/domain/jacobi_rotation_kernels.cuh:370-413):
            // sigma_j = ||a_j||, U_ij = a_ij / sigma_j if sigma_j > eps else 0.
            for (int j = gid; j < n; j += groups)
            {
                double acc = 0.0;
                for (int i = lane; i < m; i += G)
                {
                    const double x = A[j * m + i];
                    acc += x * x;
                }
                acc = dev::group_sum(acc, G);
                if (lane == 0)
                {
                    const double sj = sqrt(fmax(acc, 0.0));
                    S[j] = sj;
                    sigma[w.s_off + j] = sj;
                }
            }
            __syncthreads();
            double *ug = u + w.a_off;
            for (int idx = tid; idx < mn; idx += BS)
            {
                const int i = idx / n, j = idx - (idx / n) * n;
                const double sj = S[j];
                ug[idx] = (sj > eps) ? (A[j * m + i] / sj) : 0.0;
            }
            double *vg = v + w.v_off;
            for (int idx = tid; idx < nn; idx += BS)
            {
                const int i = idx / n, j = idx - (idx / n) * n;
                vg[idx] = V[j * n + i];
            }
            if (tid == 0 && sweeps)
                sweeps[w.index] = executed;
            if (tid == 0 && stop)
                stop[w.index] = clean ? 0 : 1;
        }

        // counters[0]: next work-list entry; counters[1]: workgroups finished.
        // The last workgroup to finish resets both, so the next launch on the
        // stream starts from a clean work-list without a memset.
        template <int BS, int RMAX>
        __global__ __launch_bounds__(BS) void lds_batch_f64(const WorkItem *__restrict__ items, int count,
                                                            const double *__restrict__ a, double *__restrict__ u,
                                                            double *__restrict__ sigma, double *__restrict__ v,
                                                            int32_t *__restrict__ sweeps, int32_t *__restrict__ stop,
                                                            double eps, int max_sweeps,
                                                            unsigned int *__restrict__ counters)
        {
            extern __shared__ double lds[];
            __shared__ int s_item;
            for (;;)
            {
                if (threadIdx.x == 0)
                    s_item = (int)atomicAdd(&counters[0], 1u);
                __syncthreads();
                const int item = s_item;
                if (item >= count)
                    break;
                const WorkItem w = items[item];
                solve_lds_f64<BS, RMAX>(w, lds, a, u, sigma, v, sweeps, stop, eps, max_sweeps);
                __syncthreads();
            }
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

        // Persistent sweeps over one column-major matrix too large for LDS.
        // flags[s] is set when sweep s rotated any pair; it is read by every
        // workgroup after the sweep's last grid barrier, so all workgroups
        // leave the sweep loop together.  A barrier that hits its spin bound
        // ends every workgroup's loops; the stop reason is then 2.
        template <int BS, int RMAX>
        __global__ __launch_bounds__(BS) void persistent_sweeps_f64(double *__restrict__ A, double *__restrict__ V,
                                                                    int64_t m, int n, double eps, int max_sweeps,
                                                                    int32_t *__restrict__ flags,
                                                                    unsigned int *__restrict__ bar,
                                                                    int32_t *__restrict__ sweeps_out,
                                                                    int32_t *__restrict__ stop_out,
                                                                    dev::SpinBound bound)
        {
            constexpr int WAVES = BS / 64;
            __shared__ double red[3][WAVES];
            const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
            const int even = n + (n & 1);
            const int half = even >> 1;
            const int rounds = n < 2 ? 0 : even - 1;
            const bool in_regs = m <= (int64_t)BS * RMAX;
            int executed = 0, clean = 0, failed = 0;
            for (int sweep = 0; sweep < max_sweeps && !failed; ++sweep)
            {
                int any = 0;
                for (int r = 0; r < rounds; ++r)
                {
                    for (int k = blockIdx.x; k < half; k += gridDim.x)
                    {
                        const int p = dev::column_at(k, r, n, even);
                        const int q = dev::column_at(even - 1 - k, r, n, even);
                        if (p < 0 || q < 0)
                            continue;
                        double *ap = A + (int64_t)p * m;
                        double *aq = A + (int64_t)q * m;
                        double xp[RMAX], xq[RMAX];
                        double pp = 0.0, qq = 0.0, pq = 0.0;
                        if (in_regs)
                        {
#pragma unroll
                            for (int j = 0; j < RMAX; ++j)
                            {
                                const int64_t i = tid + (int64_t)j * BS;
                                const bool inb = i < m;
                                xp[j] = inb ? ap[i] : 0.0;
                                xq[j] = inb ? aq[i] : 0.0;
                                pp += xp[j] * xp[j];
                                qq += xq[j] * xq[j];
                                pq += xp[j] * xq[j];
                            }
                        }
                        else
                        {
                            for (int64_t i = tid; i < m; i += BS)
                            {
                                const double x = ap[i], y = aq[i];
                                pp += x * x;
                                qq += y * y;
                                pq += x * y;
                            }
                        }
                        pp = dev::wave_sum(pp);
                        qq = dev::wave_sum(qq);
                        pq = dev::wave_sum(pq);
                        if (lane == 0)
                        {
                            red[0][wave] = pp;
                            red[1][wave] = qq;
                            red[2][wave] = pq;
                        }
                        __syncthreads();
                        pp = red[0][0];
                        qq = red[1][0];
                        pq = red[2][0];
#pragma unroll
                        for (int w2 = 1; w2 < WAVES; ++w2)
                        {
                            pp += red[0][w2];
                            qq += red[1][w2];
                            pq += red[2][w2];
                        }
                        double c = 1.0, s = 0.0;
                        if (dev::rotation_needed(pp, qq, pq, eps))
                        {
                            any = 1;
                            dev::rotation_from_tangent(dev::rotation_tangent(pp, qq, pq), c, s);
                        }
                        if (s != 0.0)
                        {
                            if (in_regs)
                            {
#pragma unroll
                                for (int j = 0; j < RMAX; ++j)
                                {
                                    const int64_t i = tid + (int64_t)j * BS;
                                    if (i < m)
                                    {
                                        ap[i] = c * xp[j] - s * xq[j];
                                        aq[i] = s * xp[j] + c * xq[j];
                                    }
                                }
                            }
                            else
                            {
                                for (int64_t i = tid; i < m; i += BS)
                                {
                                    const double x = ap[i], y = aq[i];
                                    ap[i] = c * x - s * y;
                                    aq[i] = s * x + c * y;
                                }
                            }
                            double *vp = V + (int64_t)p * n;
                            double *vq = V + (int64_t)q * n;
                            for (int i = tid; i < n; i += BS)
                            {
                                const double x = vp[i], y = vq[i];
                                vp[i] = c * x - s * y;
                                vq[i] = s * x + c * y;
                            }
                        }
                        __syncthreads();
                    }
                    if (r == rounds - 1 && any && tid == 0)
                        __hip_atomic_fetch_or(&flags[sweep], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (!dev::grid_barrier(bar, gridDim.x, bound))
                    {
                        failed = 1;
                        break;
                    }
                }
                if (failed)
                    break;
                executed = sweep + 1;
                const int f = __hip_atomic_load(&flags[sweep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (f == 0)
                {
                    clean = 1;
                    break;
                }
            }
            if (blockIdx.x == 0 && tid == 0 && sweeps_out)
                *sweeps_out = executed;
            if (blockIdx.x == 0 && tid == 0 && stop_out)
            {
                const bool faulted =
                    failed || __hip_atomic_load(bound.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
                *stop_out = faulted ? 2 : (clean ? 0 : 1);
            }
        }

        // One column pair of a column-major m x n matrix (m <= 64 RW, n <= m)
        // by one wave: lane l loads rows l + 64 j of both A columns and of
        // both V columns at once (V's loads do not wait for the rotation
        // parameters), reduces the three dots across the wave, and applies
        // the rotation when the pair is not yet orthogonal.  Returns 1 when
        // it rotated.
        template <int RW>
        __device__ __forceinline__ int wave_pair_f64(double *__restrict__ A, double *__restrict__ V, int64_t m,
                                                     int n, int p, int q, int lane, double eps)
        {
            double *ap = A + (int64_t)p * m;
            double *aq = A + (int64_t)q * m;
            double *vp = V + (int64_t)p * n;
            double *vq = V + (int64_t)q * n;
            double xp[RW], xq[RW], yp[RW], yq[RW];
#pragma unroll
            for (int j = 0; j < RW; ++j)
            {
                const int64_t i = lane + 64 * j;
                xp[j] = i < m ? ap[i] : 0.0;
                xq[j] = i < m ? aq[i] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < RW; ++j)
            {
                const int i = lane + 64 * j;
                yp[j] = i < n ? vp[i] : 0.0;
                yq[j] = i < n ? vq[i] : 0.0;
            }
            double pp = 0.0, qq = 0.0, pq = 0.0;
#pragma unroll
            for (int j = 0; j < RW; ++j)
            {
                pp = fma(xp[j], xp[j], pp);
                qq = fma(xq[j], xq[j], qq);
                pq = fma(xp[j], xq[j], pq);
            }
            dev::group_sum3_dpp<64>(pp, qq, pq);
            if (!dev::rotation_needed(pp, qq, pq, eps))
                return 0;
            double c, s;
            dev::rotation_from_tangent(dev::rotation_tangent(pp, qq, pq), c, s);
#pragma unroll
            for (int j = 0; j < RW; ++j)
            {
                const int64_t i = lane + 64 * j;
                if (i < m)
                {
                    ap[i] = dev::rot_first(c, s, xp[j], xq[j]);
                    aq[i] = dev::rot_second(c, s, xp[j], xq[j]);
                }
            }
#pragma unroll
            for (int j = 0; j < RW; ++j)
            {
                const int i = lane + 64 * j;
                if (i < n)
                {
                    vp[i] = dev::rot_first(c, s, yp[j], yq[j]);
                    vq[i] = dev::rot_second(c, s, yp[j], yq[j]);
                }
            }
            return 1;
        }

        // One column pair of any height by one wave, rows streamed in chunks
        // of 256 (4 per lane): the dots accumulate rows lane + 64 j in
        // increasing j, the same FMA chain as wave_pair_f64 (equal dots, so
        // equal rotation decisions); the update re-reads the columns and
        // rounds like wave_pair_f64's (dev::rot_first / rot_second).  The
        // first HELD chunks stay in registers from the dots to the update,
        // so only the rows beyond them are read twice (a lone wide matrix
        // runs one wave per SIMD, whose registers are otherwise idle); the
        // arithmetic and its order do not depend on HELD.
        template <int HELD>
        __device__ __forceinline__ int wave_pair_stream_f64(double *__restrict__ A, double *__restrict__ V,
                                                            int64_t m, int n, int p, int q, int lane, double eps)
        {
            constexpr int CH = 4;
            constexpr int64_t HELD_ROWS = int64_t(HELD) * 64 * CH;
            double *ap = A + (int64_t)p * m;
            double *aq = A + (int64_t)q * m;
            double pp = 0.0, qq = 0.0, pq = 0.0;
            double hp[HELD > 0 ? HELD : 1][CH], hq[HELD > 0 ? HELD : 1][CH];
#pragma unroll
            for (int h = 0; h < HELD; ++h)
#pragma unroll
                for (int j = 0; j < CH; ++j)
                {
                    const int64_t i = int64_t(h) * 64 * CH + lane + 64 * j;
                    hp[h][j] = i < m ? ap[i] : 0.0;
                    hq[h][j] = i < m ? aq[i] : 0.0;
                }
#pragma unroll
            for (int h = 0; h < HELD; ++h)
#pragma unroll
                for (int j = 0; j < CH; ++j)
                {
                    pp = fma(hp[h][j], hp[h][j], pp);
                    qq = fma(hq[h][j], hq[h][j], qq);
                    pq = fma(hp[h][j], hq[h][j], pq);
                }
            for (int64_t r0 = HELD_ROWS; r0 < m; r0 += 64 * CH)
            {
                double xp[CH], xq[CH];
#pragma unroll
                for (int j = 0; j < CH; ++j)
                {
                    const int64_t i = r0 + lane + 64 * j;
                    xp[j] = i < m ? ap[i] : 0.0;
                    xq[j] = i < m ? aq[i] : 0.0;
                }
#pragma unroll
                for (int j = 0; j < CH; ++j)
                {
                    pp = fma(xp[j], xp[j], pp);
                    qq = fma(xq[j], xq[j], qq);
                    pq = fma(xp[j], xq[j], pq);
                }
            }
            dev::group_sum3_dpp<64>(pp, qq, pq);
            if (!dev::rotation_needed(pp, qq, pq, eps))
                return 0;
            double c, s;
            dev::rotation_from_tangent(dev::rotation_tangent(pp, qq, pq), c, s);
#pragma unroll
            for (int h = 0; h < HELD; ++h)
#pragma unroll
                for (int j = 0; j < CH; ++j)
                {
                    const int64_t i = int64_t(h) * 64 * CH + lane + 64 * j;
                    if (i < m)
                    {
                        ap[i] = dev::rot_first(c, s, hp[h][j], hq[h][j]);
                        aq[i] = dev::rot_second(c, s, hp[h][j], hq[h][j]);
                    }
                }
            // The streamed rows in reverse: the chunks the dots read last
            // are the likeliest to be still in L2.
            const int64_t last =
                m > HELD_ROWS ? HELD_ROWS + ((m - HELD_ROWS - 1) / (64 * CH)) * (64 * CH) : HELD_ROWS - 64 * CH;
            for (int64_t r0 = last; r0 >= HELD_ROWS; r0 -= 64 * CH)
            {
                double xp[CH], xq[CH];
#pragma unroll
                for (int j = 0; j < CH; ++j)
                {
                    const int64_t i = r0 + lane + 64 * j;
                    xp[j] = i < m ? ap[i] : 0.0;
                    xq[j] = i < m ? aq[i] : 0.0;
                }
#pragma unroll
                for (int j = 0; j < CH; ++j)
                {
                    const int64_t i = r0 + lane + 64 * j;
                    if (i < m)
                    {
                        ap[i] = dev::rot_first(c, s, xp[j], xq[j]);
                        aq[i] = dev::rot_second(c, s, xp[j], xq[j]);
                    }
                }
            }
            double *vp = V + (int64_t)p * n;
            double *vq = V + (int64_t)q * n;
            for (int i = lane; i < n; i += 64)
            {
                const double x = vp[i], y = vq[i];
                vp[i] = dev::rot_first(c, s, x, y);
                vq[i] = dev::rot_second(c, s, x, y);
            }
            return 1;
        }

        // Chunks of 256 rows a streamed pair keeps in registers: twelve (192
        // VGPRs) with one wave per SIMD (four-wave groups), none otherwise.
        template <int WAVES>
        constexpr int kStreamHeld = WAVES == 4 ? 20 : 0;

        // Matrices too large for LDS with at most 64 * RW rows (RW = 0: any
        // height, rows streamed by wave_pair_stream_f64), solved by
        // groups of `members` workgroups: group g takes matrices g, g +
        // groups, ... of the list (sorted longest first by the host), and its
        // workgroups meet at a barrier of their own after every round, so
        // the groups of one launch run independent solves side by side.  One
        // wave per column pair: lane l holds rows l + 64 j (j < RW) of both
        // columns in registers, the three dots are reduced across the wave
        // (DPP and lane shuffles), so a round needs no workgroup barrier.
        // WAVES waves per workgroup (chosen by the host): 8 for one or two
        // matrices (a lone solve is barrier-bound and eight-wave groups
        // measured fastest), 4 for one or two matrices of more than 512 rows
        // and at most 512 columns, or more than 256 rows and at most 256
        // columns (one wave per SIMD shortens the row loop) or when an
        // eight-wave group would exceed 64 workgroups (a lone wide matrix
        // spreads over more CUs); 16 for more matrices (half
        // the workgroups per group: fewer barrier arrivals and L2 write-backs
        // where many groups share the GPU), 8 for more matrices of more than
        // 512 rows while their eight-wave groups fit one workgroup per CU.
        // A wave takes pairs
        // k, k + WAVES members, ... of each round, so the arithmetic of a
        // pair, and with it the result, does not depend on the group shape.
        // Same pairs and order as the reference circle.
        template <int RW, int WAVES>
        __global__ __launch_bounds__(64 * WAVES) void wave_group_sweeps_f64(WaveGroupArgs g)
        {
            const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
            const int group = blockIdx.x / g.members, member = blockIdx.x - group * g.members;
            const int groups = gridDim.x / g.members;
            unsigned int *bar = g.bars + group * g.members;
            unsigned int epoch = 0;
            const int stride = g.members * WAVES;
            const dev::SpinBound bound{g.fault, g.spin_limit};
            int failed = 0;
            for (int idx = group; idx < g.count; idx += groups)
            {
                const WaveMatrix w = g.mats[idx];
                double *__restrict__ A = w.a;
                double *__restrict__ V = w.v;
                const int64_t m = w.m;
                const int n = w.n;
                const int even = n + (n & 1), half = even >> 1;
                const int rounds = n < 2 ? 0 : even - 1;
                int executed = 0, clean = 0;
                for (int sweep = 0; sweep < g.max_sweeps && !failed; ++sweep)
                {
                    int any = 0;
                    for (int r = 0; r < rounds; ++r)
                    {
                        for (int k = member * WAVES + wave; k < half; k += stride)
                        {
                            const int p = dev::column_at(k, r, n, even);
                            const int q = dev::column_at(even - 1 - k, r, n, even);
                            if (p < 0 || q < 0)
                                continue;
                            if constexpr (RW == 0)
                                any |= wave_pair_stream_f64<kStreamHeld<WAVES>>(A, V, m, n, p, q, lane, g.eps);
                            else
                                any |= wave_pair_f64<RW>(A, V, m, n, p, q, lane, g.eps);
                        }
                        if (r == rounds - 1 && any && lane == 0)
                            __hip_atomic_fetch_or(&w.flags[sweep], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (!dev::flag_barrier(bar, g.members, member, ++epoch, bound))
                        {
                            failed = 1;
                            break;
                        }
                    }
                    if (failed)
                        break;
                    executed = sweep + 1;
                    const int f = __hip_atomic_load(&w.flags[sweep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (f == 0)
                    {
                        clean = 1;
                        break;
                    }
                }
                if (member == 0 && threadIdx.x == 0 && g.sweeps)
                    g.sweeps[w.index] = executed;
                if (member == 0 && threadIdx.x == 0 && g.stop)
                {
                    const bool faulted =
                        failed || __hip_atomic_load(g.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
                    g.stop[w.index] = faulted ? 2 : (clean ? 0 : 1);
                }
                if (failed)
                {
                    // The group stops here: its remaining matrices are
                    // reported as not solved.
                    for (int rest = idx + groups; rest < g.count; rest += groups)
                        if (member == 0 && threadIdx.x == 0)
                        {
                            if (g.sweeps)
                                g.sweeps[g.mats[rest].index] = 0;
                            if (g.stop)
                                g.stop[g.mats[rest].index] = 2;
                        }
                    break;
                }
            }
        }

        // The same solves with one workgroup of sixteen waves per matrix:
        // workgroup w takes matrices w, w + grid, ... of the list.  A and V
        // stay in global memory (L2 / Infinity Cache), but only this
        // workgroup touches them, so a round ends with a workgroup barrier
        // alone (no cross-CU write-back or invalidate).  Wave k takes pairs
        // k, k + 16, ... of a round; the pair arithmetic is the grouped
        // kernel's, so results equal it bitwise.  Bound by the CU's L2
        // bandwidth: a round of a 256 x 256 matrix moves 2 MB through one CU.
        template <int RW>
        __global__ __launch_bounds__(1024) void wave_block_sweeps_f64(WaveGroupArgs g)
        {
            const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
            for (int idx = blockIdx.x; idx < g.count; idx += gridDim.x)
            {
                const WaveMatrix w = g.mats[idx];
                double *__restrict__ A = w.a;
                double *__restrict__ V = w.v;
                const int64_t m = w.m;
                const int n = w.n;
                const int even = n + (n & 1), half = even >> 1;
                const int rounds = n < 2 ? 0 : even - 1;
                int executed = 0, clean = 0;
                for (int sweep = 0; sweep < g.max_sweeps; ++sweep)
                {
                    int any = 0;
                    for (int r = 0; r < rounds; ++r)
                    {
                        for (int k = wave; k < half; k += 16)
                        {
                            const int p = dev::column_at(k, r, n, even);
                            const int q = dev::column_at(even - 1 - k, r, n, even);
                            if (p < 0 || q < 0)
                                continue;
                            any |= wave_pair_f64<RW>(A, V, m, n, p, q, lane, g.eps);
                        }
                        __syncthreads();
                    }
                    executed = sweep + 1;
                    if (!__syncthreads_or(any))
                    {
                        clean = 1;
                        break;
                    }
                }
                if (threadIdx.x == 0 && g.sweeps)
                    g.sweeps[w.index] = executed;
                if (threadIdx.x == 0 && g.stop)
                    g.stop[w.index] = clean ? 0 : 1;
            }
        }

        // wave_block_sweeps_f64 for matrices of any height: one workgroup of
        // sixteen waves per matrix, rows streamed (wave_pair_stream_f64).
        __global__ __launch_bounds__(1024) void wave_block_stream_f64(WaveGroupArgs g)
        {
            const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
            for (int idx = blockIdx.x; idx < g.count; idx += gridDim.x)
            {
                const WaveMatrix w = g.mats[idx];
                const int64_t m = w.m;
                const int n = w.n;
                const int even = n + (n & 1), half = even >> 1;
                const int rounds = n < 2 ? 0 : even - 1;
                int executed = 0, clean = 0;
                for (int sweep = 0; sweep < g.max_sweeps; ++sweep)
                {
                    int any = 0;
                    for (int r = 0; r < rounds; ++r)
                    {
                        for (int k = wave; k < half; k += 16)
                        {
                            const int p = dev::column_at(k, r, n, even);
                            const int q = dev::column_at(even - 1 - k, r, n, even);
                            if (p < 0 || q < 0)
                                continue;
                            any |= wave_pair_stream_f64<0>(w.a, w.v, m, n, p, q, lane, g.eps);
                        }
                        __syncthreads();
                    }
                    executed = sweep + 1;
                    if (!__syncthreads_or(any))
                    {
                        clean = 1;
                        break;
                    }
                }
                if (threadIdx.x == 0 && g.sweeps)
                    g.sweeps[w.index] = executed;
                if (threadIdx.x == 0 && g.stop)
                    g.stop[w.index] = clean ? 0 : 1;
            }
        }

        // Tiled transpose through a padded LDS tile (32 x 33 doubles: the
        // extra column puts the 32 elements of a tile column in distinct bank
        // pairs).  Block 32 x 8.  Tiles are numbered row of tiles by row of
        // tiles on a one-dimensional grid that strides over them, so any
        // shape launches (a two-dimensional grid would cap the tile rows at
        // gfx950's grid-y limit of 65536, i.e. about 2.1M matrix rows).
        __global__ __launch_bounds__(256) void transpose_f64(const double *__restrict__ in, int64_t ld_in,
                                                             double *__restrict__ out, int64_t ld_out, int64_t rows,
                                                             int64_t cols, const double *__restrict__ row_scale,
                                                             double eps)
        {
            __shared__ double tile[32][33];
            const int64_t tiles_c = (cols + 31) / 32, tiles = tiles_c * ((rows + 31) / 32);
            const int tx = threadIdx.x, ty = threadIdx.y;
            for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x)
            {
                const int64_t r0 = (t / tiles_c) * 32, c0 = (t % tiles_c) * 32;
#pragma unroll
                for (int k = 0; k < 32; k += 8)
                {
                    const int64_t r = r0 + ty + k, c = c0 + tx;
                    if (r < rows && c < cols)
                    {
                        double x = in[r * ld_in + c];
                        if (row_scale)
                        {
                            const double sc = row_scale[r];
                            x = (sc > eps) ? (x / sc) : 0.0;
                        }
                        tile[ty + k][tx] = x;
                    }
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < 32; k += 8)
                {
                    const int64_t c = c0 + ty + k, r = r0 + tx;
                    if (c < cols && r < rows)
                        out[c * ld_out + r] = tile[tx][ty + k];
                }
                __syncthreads();
            }
        }

        // sigma_j = ||a_j|| of a column-major m x n matrix, one workgroup per
        // column: thread t sums rows t, t + BS, ... (FMA chain), each wave's
        // sums are added by the butterfly of dev::wave_sum, then the waves'
        // totals in wave order.  The ring form of the wave solves reproduces
        // this order for BS = 256 from its register-held columns.
        template <int BS>
        __global__ __launch_bounds__(BS) void column_norms_f64(const double *__restrict__ a, double *__restrict__ sigma,
                                                               int64_t m, int64_t lda)
        {
            constexpr int WAVES = BS / 64;
            __shared__ double red[WAVES];
            const double *col = a + (int64_t)blockIdx.x * lda;
            double acc = 0.0;
            for (int64_t i = threadIdx.x; i < m; i += BS)
            {
                const double x = col[i];
                acc = fma(x, x, acc);
            }
            acc = dev::wave_sum(acc);
            if ((threadIdx.x & 63) == 0)
                red[threadIdx.x >> 6] = acc;
            __syncthreads();
            if (threadIdx.x == 0)
            {
                double t = red[0];
                for (int w = 1; w < WAVES; ++w)
                    t += red[w];
                sigma[blockIdx.x] = sqrt(fmax(t, 0.0));
            }
        }

        __global__ void identity_f64(double *v, int n)
        {
            const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
            const int64_t total = (int64_t)n * n;
            if (idx < total)
                v[idx] = (idx % (n + 1) == 0) ? 1.0 : 0.0;
        }

        // Network (big-endian) to host byte order, in place.
        __global__ void bswap64_inplace(uint64_t *p, int64_t count)
        {
            const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
            if (idx < count)
                p[idx] = __builtin_bswap64(p[idx]);
        }

        __global__ void fill_i32(int32_t *p, int32_t value, int64_t count)
        {
            const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
            if (idx < count)
                p[idx] = value;
        }
    } // namespace

    constexpr int kLdsRmax = 16;
    constexpr int kPersistentRmax = 16;
    constexpr int kWaveRowsMax = 8; // wave-per-pair kernel: up to 512 rows

    const void *lds_batch_f64_symbol(int block)
    {
        switch (block)
        {
        case 64: return reinterpret_cast<const void *>(&lds_batch_f64<64, kLdsRmax>);
        case 128: return reinterpret_cast<const void *>(&lds_batch_f64<128, kLdsRmax>);
        case 256: return reinterpret_cast<const void *>(&lds_batch_f64<256, kLdsRmax>);
        case 512: return reinterpret_cast<const void *>(&lds_batch_f64<512, kLdsRmax>);
        default: return nullptr;
        }
    }

    hipError_t launch_lds_batch_f64(int block, int grid, int lds_bytes, const WorkItem *items, int count,
                                    const double *a, double *u, double *sigma, double *v, int32_t *sweeps,
                                    int32_t *stop, double eps, int max_sweeps, unsigned int *counters,
                                    hipStream_t stream)
    {
        switch (block)
        {
#define JSVD_LDS_BATCH_CASE(BS)                                                                                     \
    case BS:                                                                                                       \
        lds_batch_f64<BS, kLdsRmax><<<grid, BS, lds_bytes, stream>>>(items, count, a, u, sigma, v, sweeps, stop, eps, \
                                                                     max_sweeps, counters);                       \
        break;
            JSVD_LDS_BATCH_CASE(64)
            JSVD_LDS_BATCH_CASE(128)
            JSVD_LDS_BATCH_CASE(256)
            JSVD_LDS_BATCH_CASE(512)
#undef JSVD_LDS_BATCH_CASE
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }

    const void *persistent_sweeps_f64_symbol(int block)
    {
        switch (block)
        {
        case 256: return reinterpret_cast<const void *>(&persistent_sweeps_f64<256, kPersistentRmax>);
        case 512: return reinterpret_cast<const void *>(&persistent_sweeps_f64<512, kPersistentRmax>);
        case 1024: return reinterpret_cast<const void *>(&persistent_sweeps_f64<1024, kPersistentRmax>);
        default: return nullptr;
        }
    }

    // The grid barrier needs every workgroup resident at once.  The caller
    // sizes `grid` to the occupancy-guaranteed resident count and issues all
    // barrier kernels of a device on one stream, so two of them never share
    // the GPU; any other kernel that delays residency finishes on its own.
    int wave_sweeps_max_rows()
    {
        return 64 * kWaveRowsMax;
    }

    template <int WAVES>
    const void *wave_group_symbol(int rw)
    {
        switch (rw)
        {
        case 0: return reinterpret_cast<const void *>(&wave_group_sweeps_f64<0, WAVES>);
        case 1: return reinterpret_cast<const void *>(&wave_group_sweeps_f64<1, WAVES>);
        case 2: return reinterpret_cast<const void *>(&wave_group_sweeps_f64<2, WAVES>);
        case 3: return reinterpret_cast<const void *>(&wave_group_sweeps_f64<3, WAVES>);
        case 4: return reinterpret_cast<const void *>(&wave_group_sweeps_f64<4, WAVES>);
        case 5: return reinterpret_cast<const void *>(&wave_group_sweeps_f64<5, WAVES>);
        case 6: return reinterpret_cast<const void *>(&wave_group_sweeps_f64<6, WAVES>);
        case 7: return reinterpret_cast<const void *>(&wave_group_sweeps_f64<7, WAVES>);
        case 8: return reinterpret_cast<const void *>(&wave_group_sweeps_f64<8, WAVES>);
        case 16:
            if constexpr (WAVES != 16)
                return reinterpret_cast<const void *>(&wave_group_sweeps_f64<16, WAVES>);
            else
                return nullptr;
        default: return nullptr;
        }
    }

    const void *wave_group_sweeps_f64_symbol(int64_t m, int waves)
    {
        // Up to 512 rows: the exact row-register count.  513 .. 1024 rows: 16
        // row registers per lane for the eight- and four-wave groups of one
        // or two matrices (one load round per pair instead of streamed
        // chunks: 1024 x 129 10.8 -> 7.9 ms, 1000 x 500 68.8 -> 48.8 ms); the
        // sixteen-wave groups of batches stream (the registers would halve
        // their occupancy: 64 x 1024x64 11.8 -> 16.9 ms).  Taller: streamed
        // (RW = 0).
        const int rw = m <= 64 * kWaveRowsMax ? static_cast<int>((m + 63) / 64)
                       : (m <= 1024 && waves != 16) ? 16
                                                      : 0;
        return waves == 16  ? wave_group_symbol<16>(rw)
               : waves == 8 ? wave_group_symbol<8>(rw)
               : waves == 4 ? wave_group_symbol<4>(rw)
                            : nullptr;
    }

    hipError_t launch_wave_block_sweeps_f64(int64_t m, int grid, const WaveGroupArgs &args, hipStream_t stream)
    {
        if (grid < 1 || args.count < 1)
            return hipErrorInvalidValue;
        if (m > 64 * kWaveRowsMax)
        {
            wave_block_stream_f64<<<grid, 1024, 0, stream>>>(args);
            return hipGetLastError();
        }
        const int rw = static_cast<int>((m + 63) / 64);
#define JSVD_WAVE_CASE(RW)                                                                                         \
    case RW:                                                                                                       \
        wave_block_sweeps_f64<RW><<<grid, 1024, 0, stream>>>(args);                                                \
        break;
        switch (rw)
        {
            JSVD_WAVE_CASE(1)
            JSVD_WAVE_CASE(2)
            JSVD_WAVE_CASE(3)
            JSVD_WAVE_CASE(4)
            JSVD_WAVE_CASE(5)
            JSVD_WAVE_CASE(6)
            JSVD_WAVE_CASE(7)
            JSVD_WAVE_CASE(8)
        default: return hipErrorInvalidValue;
        }
#undef JSVD_WAVE_CASE
        return hipGetLastError();
    }

    hipError_t launch_wave_group_sweeps_f64(int64_t m, int waves, int groups, const WaveGroupArgs &args,
                                            hipStream_t stream)
    {
        if (groups < 1 || args.members < 1 || args.count < 1)
            return hipErrorInvalidValue;
        const void *fn = wave_group_sweeps_f64_symbol(m, waves);
        if (!fn)
            return hipErrorInvalidValue;
        const unsigned int grid = static_cast<unsigned int>(groups) * static_cast<unsigned int>(args.members);
        void *params[] = {const_cast<WaveGroupArgs *>(&args)};
        return hipLaunchKernel(fn, dim3(grid), dim3(64 * waves), params, 0, stream);
    }

    hipError_t launch_persistent_sweeps_f64(int block, int grid, double *a_cm, double *v_cm, int64_t m, int n,
                                            double eps, int max_sweeps, int32_t *flags, unsigned int *barrier,
                                            int32_t *sweeps_out, int32_t *stop_out, unsigned int *fault,
                                            uint64_t spin_limit, hipStream_t stream)
    {
        const dev::SpinBound bound{fault, spin_limit};
        switch (block)
        {
        case 256:
            persistent_sweeps_f64<256, kPersistentRmax><<<grid, 256, 0, stream>>>(
                a_cm, v_cm, m, n, eps, max_sweeps, flags, barrier, sweeps_out, stop_out, bound);
            break;
        case 512:
            persistent_sweeps_f64<512, kPersistentRmax><<<grid, 512, 0, stream>>>(
                a_cm, v_cm, m, n, eps, max_sweeps, flags, barrier, sweeps_out, stop_out, bound);
            break;
        case 1024:
            persistent_sweeps_f64<1024, kPersistentRmax><<<grid, 1024, 0, stream>>>(
                a_cm, v_cm, m, n, eps, max_sweeps, flags, barrier, sweeps_out, stop_out, bound);
            break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }

    hipError_t launch_transpose_ld_f64(const double *in, int64_t ld_in, double *out, int64_t ld_out, int64_t rows,
                                       int64_t cols, const double *row_scale, double eps, hipStream_t stream)
    {
        if (rows == 0 || cols == 0)
            return hipSuccess;
        const int64_t tiles = ((cols + 31) / 32) * ((rows + 31) / 32);
        const unsigned grid = static_cast<unsigned>(std::min<int64_t>(tiles, int64_t(1) << 20));
        transpose_f64<<<grid, dim3(32, 8), 0, stream>>>(in, ld_in, out, ld_out, rows, cols, row_scale, eps);
        return hipGetLastError();
    }

    hipError_t launch_transpose_f64(const double *in, double *out, int64_t rows, int64_t cols,
                                    const double *row_scale, double eps, hipStream_t stream)
    {
        return launch_transpose_ld_f64(in, cols, out, rows, rows, cols, row_scale, eps, stream);
    }

    hipError_t launch_column_norms_f64(const double *a_cm, int64_t lda, double *sigma, int64_t m, int n,
                                       hipStream_t stream)
    {
        if (n == 0)
            return hipSuccess;
        column_norms_f64<256><<<n, 256, 0, stream>>>(a_cm, sigma, m, lda);
        return hipGetLastError();
    }

    hipError_t launch_identity_f64(double *v, int n, hipStream_t stream)
    {
        const int64_t total = (int64_t)n * n;
        if (total == 0)
            return hipSuccess;
        identity_f64<<<(unsigned)((total + 255) / 256), 256, 0, stream>>>(v, n);
        return hipGetLastError();
    }

    hipError_t launch_bswap64(void *p, int64_t count, hipStream_t stream)
    {
        if (count == 0)
            return hipSuccess;
        bswap64_inplace<<<(unsigned)((count + 255) / 256), 256, 0, stream>>>(static_cast<uint64_t *>(p), count);
        return hipGetLastError();
    }

    hipError_t launch_fill_i32(int32_t *p, int32_t value, int64_t count, hipStream_t stream)
    {
        if (count == 0)
            return hipSuccess;
        fill_i32<<<(unsigned)((count + 255) / 256), 256, 0, stream>>>(p, value, count);
        return hipGetLastError();
    }
} // namespace jsvd
