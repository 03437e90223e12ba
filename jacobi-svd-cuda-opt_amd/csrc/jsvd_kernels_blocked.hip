// jsvd_kernels_blocked.hip -- blocked one-sided Jacobi for matrices that do
// not fit in LDS (no reference counterpart: the reference rotates single
// column pairs only, src/infrastructure/jacobi_rotation_kernels.cuh:241-359).
//
// The n columns, padded with zero columns to an even number NB of blocks of
// width B, form the blocks; the m rows are padded with zero rows to lda (a
// multiple of 64).  Block pairs follow the circle order over blocks; one
// round is three launches over the NB/2 block pairs of that round:
//  1. gram_partial: G_c = P_c^T P_c for every block pair P = [A_I A_J]
//     (lda x 2B) and row chunk c on FP64 MFMA (v_mfma_f64_16x16x4f64).  Each
//     wave streams 16-row slices straight from HBM into MFMA operand
//     registers (the reduction over rows does not care which k-slot a row
//     lands in, so every lane reads 4 consecutive rows of one column) and
//     accumulates the upper triangle of 16 x 16 tiles; the four waves of a
//     workgroup are summed through LDS at the end.
//  2. gram_solve: G = sum_c G_c in a fixed order (deterministic), then one
//     cyclic pass of two-sided Jacobi on G in LDS.  Round 0 of a sweep visits
//     every column pair of the 2B columns (circle order over them, 2B - 1
//     inner rounds); later rounds visit only the B^2 pairs across the two
//     blocks (B inner rounds, column k of block I against column
//     (k + t) mod B of block J), so each sweep rotates every column pair of A
//     exactly once, as the column-pair kernels do.  Each inner round uses the
//     Rutishauser rotation and threshold of the column-pair path, updates
//     G <- J^T G J in one pass over 2 x 2 blocks and accumulates Q <- Q J.
//  3. block_update: [A_I A_J] <- [A_I A_J] Q and [V_I V_J] <- [V_I V_J] Q on
//     MFMA, formed as Q^T P^T with Q^T held in registers for the whole
//     workgroup and P streamed in 16-row slices: operand loads and result
//     stores are 128 contiguous bytes of a column per 16 lanes, no LDS.
// The sweep flag is raised by any applied rotation; a sweep without one has
// tested every column pair on an unchanged A, which is the column-pair
// path's convergence criterion, so sweep counts mean the same thing.
#include "jsvd_device.hpp"
#include "jsvd_kernels.hpp"

#include <atomic>

namespace jsvd
{
    namespace
    {
        typedef double f64x4 __attribute__((ext_vector_type(4)));

        // Blocks of pair slot k in round r (circle order over nb blocks).
        __device__ __forceinline__ void pair_blocks(int k, int r, int nb, int &I, int &J)
        {
            const int even = nb + (nb & 1);
            I = dev::column_at(k, r, nb, even);
            J = dev::column_at(even - 1 - k, r, nb, even);
        }

        // First matrix column of 16-column tile t of the panel [A_I A_J].
        template <int B>
        __device__ __forceinline__ int64_t tile_column(int t, int I, int J)
        {
            return t * 16 < B ? int64_t(I) * B + t * 16 : int64_t(J) * B + (t * 16 - B);
        }

        template <int B>
        __global__ __launch_bounds__(256) void gram_partial(const double *__restrict__ A, int64_t lda, int nb,
                                                           int round, int64_t chunk_rows, int chunks,
                                                           double *__restrict__ gram, const int32_t *__restrict__ done)
        {
            if (*done)
                return; // the solve has converged: the remaining launches are no-ops
            constexpr int W = 2 * B;
            constexpr int T16 = W / 16;
            constexpr int UT = T16 * (T16 + 1) / 2; // upper-triangle tiles (a <= b), row by row
            __shared__ double red[3 * UT * 4 * 64];
            const int pair = blockIdx.x / chunks, chunk = blockIdx.x - pair * chunks;
            int I, J;
            pair_blocks(pair, round, nb, I, J);
            if (I < 0 || J < 0)
                return;
            const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, i = l & 15, g = l >> 4;
            const int64_t r0 = int64_t(chunk) * chunk_rows;
            const int64_t r1 = r0 + chunk_rows < lda ? r0 + chunk_rows : lda;
            // Lane (i, g) reads rows base + 4g .. base + 4g + 3 of column i of
            // every tile; at MFMA step s both operands take row base + 4g + s.
            const double *col[T16];
#pragma unroll
            for (int t = 0; t < T16; ++t)
                col[t] = A + (tile_column<B>(t, I, J) + i) * lda + 4 * g;
            f64x4 acc[UT];
#pragma unroll
            for (int u = 0; u < UT; ++u)
                acc[u] = f64x4{0.0, 0.0, 0.0, 0.0};
            double2 x[T16][2], y[T16][2];
            int64_t base = r0 + 16 * wave;
            if (base < r1)
            {
#pragma unroll
                for (int t = 0; t < T16; ++t)
                {
                    x[t][0] = *reinterpret_cast<const double2 *>(col[t] + base);
                    x[t][1] = *reinterpret_cast<const double2 *>(col[t] + base + 2);
                }
            }
            for (; base < r1; base += 64)
            {
                if (base + 64 < r1)
                {
#pragma unroll
                    for (int t = 0; t < T16; ++t)
                    {
                        y[t][0] = *reinterpret_cast<const double2 *>(col[t] + base + 64);
                        y[t][1] = *reinterpret_cast<const double2 *>(col[t] + base + 66);
                    }
                }
#pragma unroll
                for (int s = 0; s < 4; ++s)
                {
                    int u = 0;
#pragma unroll
                    for (int a = 0; a < T16; ++a)
                    {
                        const double xa = (s & 1) ? x[a][s >> 1].y : x[a][s >> 1].x;
#pragma unroll
                        for (int b = a; b < T16; ++b)
                        {
                            const double xb = (s & 1) ? x[b][s >> 1].y : x[b][s >> 1].x;
                            acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa, xb, acc[u], 0, 0, 0);
                            ++u;
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < T16; ++t)
                {
                    x[t][0] = y[t][0];
                    x[t][1] = y[t][1];
                }
            }
            // Sum the four waves' tiles; wave 0 writes the partial Gram.
            if (wave > 0)
            {
                double *w = red + (wave - 1) * (UT * 4 * 64);
#pragma unroll
                for (int u = 0; u < UT; ++u)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        w[(u * 4 + r) * 64 + l] = acc[u][r];
            }
            __syncthreads();
            if (wave == 0)
            {
                double *out = gram + (int64_t(pair) * chunks + chunk) * (W * W);
                int u = 0;
#pragma unroll
                for (int a = 0; a < T16; ++a)
                {
#pragma unroll
                    for (int b = a; b < T16; ++b)
                    {
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                        {
                            double v = acc[u][r];
#pragma unroll
                            for (int w = 0; w < 3; ++w)
                                v += red[w * (UT * 4 * 64) + (u * 4 + r) * 64 + l];
                            // Item r of lane l: tile row g + 4r, tile column i.
                            out[(a * 16 + g + 4 * r) * W + b * 16 + i] = v;
                        }
                        ++u;
                    }
                }
            }
        }

        template <int B, int NT>
        __global__ __launch_bounds__(NT) void gram_solve(const double *__restrict__ gram, int chunks, int nb, int round,
                                                         double eps, double *__restrict__ qout,
                                                         int32_t *__restrict__ pair_rotated,
                                                         int32_t *__restrict__ sweep_flag, const int32_t *__restrict__ done)
        {
            if (*done)
                return;
            constexpr int W = 2 * B;
            constexpr int LDG = W + 1;
            constexpr int NBLK = B * (B + 1) / 2; // 2 x 2 blocks (i <= j) of the inner pairs
            constexpr int BPT = (NBLK + NT - 1) / NT;
            constexpr int QPT = (B * W + NT - 1) / NT;
            __shared__ double G[W * LDG];
            __shared__ double Q[W * LDG];
            __shared__ double rc[B], rs[B];
            __shared__ int rp[B], rq[B];
            const int pair = blockIdx.x, tid = threadIdx.x;
            int I, J;
            pair_blocks(pair, round, nb, I, J);
            if (I < 0 || J < 0)
                return;
            const double *gp = gram + int64_t(pair) * chunks * (W * W);
            for (int idx = tid; idx < W * W; idx += NT)
            {
                const int i = idx / W, j = idx - i * W;
                if (i <= j)
                {
                    double s = 0.0;
                    for (int c = 0; c < chunks; ++c)
                        s += gp[int64_t(c) * (W * W) + idx];
                    G[i * LDG + j] = s;
                    G[j * LDG + i] = s;
                }
                Q[i * LDG + j] = i == j ? 1.0 : 0.0;
            }
            // This thread's 2 x 2 blocks (pairs i <= j) and Q entries; the
            // assignment is the same in every inner round.
            int bi[BPT], bj[BPT];
#pragma unroll
            for (int t = 0; t < BPT; ++t)
            {
                int rem = tid + NT * t, i = 0;
                if (rem < NBLK)
                {
                    while (rem >= B - i)
                    {
                        rem -= B - i;
                        ++i;
                    }
                    bi[t] = i;
                    bj[t] = i + rem;
                }
                else
                {
                    bi[t] = -1;
                    bj[t] = -1;
                }
            }
            __syncthreads();
            const bool full = round == 0;
            const int inner = full ? W - 1 : B;
            int rotated = 0; // this thread saw a rotation (threads < B)
            for (int rr = 0; rr < inner; ++rr)
            {
                if (tid < B)
                {
                    int p, q;
                    if (full)
                    {
                        p = dev::column_at(tid, rr, W, W);
                        q = dev::column_at(W - 1 - tid, rr, W, W);
                    }
                    else
                    {
                        p = tid;
                        q = B + (tid + rr) % B;
                    }
                    // Rotation parameters from the refined hardware
                    // estimates (one dependent chain of about half the FP64
                    // operations of IEEE division and square root).
                    const double gpp = G[p * LDG + p], gqq = G[q * LDG + q], gpq = G[p * LDG + q];
                    double c = 1.0, s = 0.0;
                    if (dev::rotation_needed(gpp, gqq, gpq, eps))
                    {
                        dev::rotation_from_tangent(dev::rotation_tangent(gpp, gqq, gpq), c, s);
                        rotated = 1;
                    }
                    rc[tid] = c;
                    rs[tid] = s;
                    rp[tid] = p;
                    rq[tid] = q;
                }
                // One barrier per inner round: rounds without a rotation
                // still run the (skipping) update loops.
                __syncthreads();
                // G' = J^T G J; J restricted to pair k is [[c, s], [-s, c]] on
                // columns (p, q).
#pragma unroll
                for (int t = 0; t < BPT; ++t)
                {
                    const int i = bi[t], j = bj[t];
                    if (i < 0)
                        continue;
                    const double ci = rc[i], si = rs[i], cj = rc[j], sj = rs[j];
                    if (si == 0.0 && sj == 0.0)
                        continue;
                    const int pi = rp[i], qi = rq[i], pj = rp[j], qj = rq[j];
                    const double g00 = G[pi * LDG + pj], g01 = G[pi * LDG + qj];
                    const double g10 = G[qi * LDG + pj], g11 = G[qi * LDG + qj];
                    const double x00 = cj * g00 - sj * g01, x01 = sj * g00 + cj * g01;
                    const double x10 = cj * g10 - sj * g11, x11 = sj * g10 + cj * g11;
                    const double n00 = ci * x00 - si * x10, n01 = ci * x01 - si * x11;
                    const double n10 = si * x00 + ci * x10, n11 = si * x01 + ci * x11;
                    if (i == j)
                    {
                        G[pi * LDG + pi] = n00;
                        G[qi * LDG + qi] = n11;
                        // The rotation annihilates the pair's own entry.
                        G[pi * LDG + qi] = 0.0;
                        G[qi * LDG + pi] = 0.0;
                    }
                    else
                    {
                        G[pi * LDG + pj] = n00;
                        G[pj * LDG + pi] = n00;
                        G[pi * LDG + qj] = n01;
                        G[qj * LDG + pi] = n01;
                        G[qi * LDG + pj] = n10;
                        G[pj * LDG + qi] = n10;
                        G[qi * LDG + qj] = n11;
                        G[qj * LDG + qi] = n11;
                    }
                }
#pragma unroll
                for (int t = 0; t < QPT; ++t)
                {
                    const int e = tid + NT * t;
                    if (e >= B * W)
                        continue;
                    const int k = e / W, row = e - k * W;
                    const double c = rc[k], s = rs[k];
                    if (s == 0.0)
                        continue;
                    const int p = rp[k], q = rq[k];
                    const double x = Q[row * LDG + p], y = Q[row * LDG + q];
                    Q[row * LDG + p] = c * x - s * y;
                    Q[row * LDG + q] = s * x + c * y;
                }
                __syncthreads();
            }
            rotated = __syncthreads_or(rotated);
            if (rotated)
            {
                double *q = qout + int64_t(pair) * (W * W);
                for (int idx = tid; idx < W * W; idx += NT)
                {
                    const int i = idx / W, j = idx - i * W;
                    q[idx] = Q[i * LDG + j];
                }
            }
            if (tid == 0)
            {
                pair_rotated[pair] = rotated;
                if (rotated)
                    __hip_atomic_fetch_or(sweep_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }

        // [X_I X_J] <- [X_I X_J] Q.  Work units are 128-row pieces of A (ua
        // per pair) and V (uv per pair) in pair order; workgroup w takes the
        // contiguous units [w U / grid, (w + 1) U / grid), so a grid of
        // resident workgroups finishes together and Q^T is reloaded only
        // when the pair changes.
        constexpr int kUnitRows = 128;

        template <int B>
        __global__ __launch_bounds__(256) void block_update(double *__restrict__ A, double *__restrict__ V, int64_t lda,
                                                           int nv, int nb, int round, int ua, int uv, int64_t units,
                                                           const double *__restrict__ qin,
                                                           const int32_t *__restrict__ pair_rotated,
                                                           const int32_t *__restrict__ done)
        {
            if (*done)
                return;
            constexpr int W = 2 * B;
            constexpr int T16 = W / 16;
            constexpr int KS = W / 4; // MFMA k-steps
            const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, i = l & 15, g = l >> 4;
            const int per = ua + uv;
            const int64_t u0 = int64_t(blockIdx.x) * units / gridDim.x;
            const int64_t u1 = int64_t(blockIdx.x + 1) * units / gridDim.x;
            int cur = -1, I = -1, J = -1;
            bool active = false;
            double qa[T16][KS];
            for (int64_t u = u0; u < u1; ++u)
            {
                const int pair = static_cast<int>(u / per), part = static_cast<int>(u - int64_t(pair) * per);
                if (pair != cur)
                {
                    cur = pair;
                    pair_blocks(pair, round, nb, I, J);
                    active = I >= 0 && J >= 0 && pair_rotated[pair] != 0;
                    if (active)
                    {
                        // Operand A of result tile tq at k-step ks: element
                        // (i, g) of Q^T, i.e. Q[4 ks + g][16 tq + i].
                        const double *qg = qin + int64_t(pair) * (W * W);
#pragma unroll
                        for (int tq = 0; tq < T16; ++tq)
#pragma unroll
                            for (int ks = 0; ks < KS; ++ks)
                                qa[tq][ks] = qg[(4 * ks + g) * W + tq * 16 + i];
                    }
                }
                if (!active)
                    continue;
                const bool is_v = part >= ua;
                double *X = is_v ? V : A;
                const int64_t ld = is_v ? int64_t(nv) : lda;
                const int64_t r0 = int64_t(is_v ? part - ua : part) * kUnitRows;
                const int64_t r1 = r0 + kUnitRows < ld ? r0 + kUnitRows : ld;
                // Operand B at k-step ks: P[base + i][4 ks + g]; panel column
                // 4 ks + g lies in block I exactly when ks < B / 4.  Addresses
                // are a per-lane offset plus a wave-uniform one.
                const int64_t colI = int64_t(I) * B, colJ = int64_t(J) * B - B;
                const double *Xl = X + int64_t(g) * ld + i;
                double b[KS];
                int64_t base = r0 + 16 * wave;
                if (base < r1)
                {
#pragma unroll
                    for (int ks = 0; ks < KS; ++ks)
                        b[ks] = Xl[((ks < B / 4 ? colI : colJ) + 4 * ks) * ld + base];
                }
                for (; base < r1; base += 64)
                {
                    const bool next = base + 64 < r1;
                    f64x4 acc[T16];
#pragma unroll
                    for (int tq = 0; tq < T16; ++tq)
                        acc[tq] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int ks = 0; ks < KS; ++ks)
                    {
#pragma unroll
                        for (int tq = 0; tq < T16; ++tq)
                            acc[tq] = __builtin_amdgcn_mfma_f64_16x16x4f64(qa[tq][ks], b[ks], acc[tq], 0, 0, 0);
                        // Operand consumed: fetch the same slot of the next slice.
                        if (next)
                            b[ks] = Xl[((ks < B / 4 ? colI : colJ) + 4 * ks) * ld + base + 64];
                    }
                    // Item r of lane l: result column 16 tq + g + 4r, row base + i.
#pragma unroll
                    for (int tq = 0; tq < T16; ++tq)
                    {
                        double *dst = X + (tile_column<B>(tq, I, J) + g) * ld + base + i;
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            dst[4 * r * ld] = acc[tq][r];
                    }
                }
            }
        }

        // Workgroups of `fn` (256 threads, static LDS only) resident on the
        // current device at once, cached per device (threads that race on the
        // first query compute and store the same value).
        int resident_workgroups(const void *fn, int slot)
        {
            static std::atomic<int> cache[16][2];
            int dev = 0;
            if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16)
                return 256;
            int v = cache[dev][slot].load(std::memory_order_relaxed);
            if (v == 0)
            {
                int per_cu = 0, cus = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess || per_cu < 1)
                    per_cu = 1;
                if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
                    cus = 256;
                v = per_cu * cus;
                cache[dev][slot].store(v, std::memory_order_relaxed);
            }
            return v;
        }

        template <int B>
        hipError_t launch_round(const BlockedArgs &a, int round, hipStream_t stream)
        {
            const int pairs = (a.nb + (a.nb & 1)) / 2;
            const int32_t *done = a.state + 1;
            gram_partial<B><<<pairs * a.chunks, 256, 0, stream>>>(a.a, a.lda, a.nb, round, a.chunk_rows, a.chunks,
                                                                a.gram, done);
            constexpr int NT = B == 32 ? 1024 : 512;
            gram_solve<B, NT><<<pairs, NT, 0, stream>>>(a.gram, a.chunks, a.nb, round, a.eps, a.q, a.pair_rotated,
                                                    a.state, done);
            const int ua = static_cast<int>((a.lda + kUnitRows - 1) / kUnitRows);
            const int uv = (a.nv + kUnitRows - 1) / kUnitRows;
            const int64_t units = int64_t(pairs) * (ua + uv);
            const int slots = resident_workgroups(reinterpret_cast<const void *>(&block_update<B>), B == 32 ? 0 : 1);
            const int grid = static_cast<int>(units < slots ? units : slots);
            block_update<B><<<grid, 256, 0, stream>>>(a.a, a.v, a.lda, a.nv, a.nb, round, ua, uv, units, a.q,
                                                     a.pair_rotated, done);
            return hipGetLastError();
        }

        // Sweep bookkeeping on the device (one thread).  At the start of
        // sweep s > 0 it closes sweep s - 1: that sweep counts as executed,
        // and a sweep that rotated no pair ends the solve (done, converged),
        // after which every launch of the later sweeps returns at once.
        __global__ void blocked_sweep_step(int32_t *state, int sweep)
        {
            if (state[1])
                return;
            if (sweep > 0)
            {
                state[3] = sweep;
                if (state[0] == 0)
                {
                    state[1] = 1;
                    state[2] = 1;
                    return;
                }
            }
            state[0] = 0;
        }

        // After the last enqueued sweep: closes it if the solve is still
        // running, then reports sweeps and stop reason.
        __global__ void blocked_sweep_close(int32_t *state, int enqueued, int32_t *sweeps_out, int32_t *stop_out)
        {
            if (!state[1])
            {
                state[3] = enqueued;
                state[2] = state[0] == 0 ? 1 : 0;
                state[1] = 1;
            }
            if (sweeps_out)
                *sweeps_out = state[3];
            if (stop_out)
                *stop_out = state[2] ? 0 : 1;
        }
    } // namespace

    hipError_t launch_blocked_sweep_step(int32_t *state, int sweep, hipStream_t stream)
    {
        blocked_sweep_step<<<1, 1, 0, stream>>>(state, sweep);
        return hipGetLastError();
    }

    hipError_t launch_blocked_sweep_close(int32_t *state, int enqueued, int32_t *sweeps_out, int32_t *stop_out,
                                          hipStream_t stream)
    {
        blocked_sweep_close<<<1, 1, 0, stream>>>(state, enqueued, sweeps_out, stop_out);
        return hipGetLastError();
    }

    hipError_t launch_blocked_round(const BlockedArgs &args, int round, hipStream_t stream)
    {
        if (args.block == 16)
            return launch_round<16>(args, round, stream);
        if (args.block == 32)
            return launch_round<32>(args, round, stream);
        return hipErrorInvalidValue;
    }
} // namespace jsvd
