// jsvd_device.hpp -- device helpers shared by the gfx950 Jacobi kernels:
// circle-method schedule, Rutishauser rotation, wave64 reductions and the
// grid barrier of the persistent sweep kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace jsvd::dev
{
    // Column held by `position` of the circle in `round` (slot 0 fixed to
    // column 0, slots 1..E-1 rotate by one per round); -1 for the dummy column
    // of an odd n.  Same order as the reference schedule
    // (src/domain/jacobi_rotation_kernels.cuh:18-42, src/domain/jacobi_schedule.cuh:16-57).
    __device__ __forceinline__ int column_at(int position, int round, int n, int even_n)
    {
        if (position == 0)
            return 0;
        const int ring = even_n - 1;
        int r = (position - 1 - round) % ring;
        r += (r < 0) ? ring : 0;
        const int column = 1 + r;
        return column < n ? column : -1;
    }

    // Rutishauser rotation for the 2x2 Gram [[pp, pq], [pq, qq]] with the
    // reference's threshold |pq| > eps*sqrt(pp*qq) and |pq| > 1e-300
    // (src/domain/jacobi_rotation_kernels.cuh:305-326).  Returns true when the
    // pair counts as rotated.  Always evaluated in FP64.
    __device__ __forceinline__ bool rotation(double pp, double qq, double pq, double eps, double &c, double &s)
    {
        const double rhs = eps * sqrt(fmax(pp, 0.0) * fmax(qq, 0.0));
        const double apq = fabs(pq);
        if (apq > rhs && apq > 1.0e-300)
        {
            const double zeta = (qq - pp) / (2.0 * pq);
            const double root = sqrt(1.0 + zeta * zeta);
            const double t = (zeta >= 0.0) ? (1.0 / (zeta + root)) : (-1.0 / (-zeta + root));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
            return true;
        }
        c = 1.0;
        s = 0.0;
        return false;
    }

    // Butterfly all-reduce over aligned groups of G lanes (G a power of two,
    // G <= 64).  Every lane of the group ends with the group sum.
    template <typename T>
    __device__ __forceinline__ T group_sum(T x, int G)
    {
        for (int mask = G >> 1; mask > 0; mask >>= 1)
            x += __shfl_xor(x, mask, 64);
        return x;
    }

    template <typename T>
    __device__ __forceinline__ T wave_sum(T x)
    {
#pragma unroll
        for (int mask = 32; mask > 0; mask >>= 1)
            x += __shfl_xor(x, mask, 64);
        return x;
    }

    // Sense-free generation barrier across all workgroups of a cooperative
    // launch.  bar[0] counts arrivals, bar[1] is the generation.  The
    // agent-scope fences write back and invalidate the per-XCD L2 so that
    // columns written in one round are seen by every XCD in the next.
    __device__ __forceinline__ void grid_barrier(unsigned int *bar, unsigned int nblocks)
    {
        __syncthreads();
        if (threadIdx.x == 0)
        {
            const unsigned int gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            const unsigned int arrived =
                __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (arrived == nblocks - 1)
            {
                __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            else
            {
                while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen)
                    __builtin_amdgcn_s_sleep(1);
            }
            __threadfence();
        }
        __syncthreads();
    }
} // namespace jsvd::dev
