// jsvd_device.hpp -- device helpers shared by the gfx950 Jacobi kernels:
// circle-method schedule, Rutishauser rotation parameters, wave64
// reductions and the grid barrier of the persistent sweep kernels.
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

    // Incremental form of column_at for the two slots of pair position k:
    // at round 0 they hold columns k and E-1-k; each round every slot except
    // slot 0 takes the next lower column, wrapping 1 -> E-1.  A value >= n is
    // the dummy column of an odd n.  Replaces two integer modulos per round.
    struct CirclePair
    {
        int p, q;
        __device__ __forceinline__ void start(int k, int even)
        {
            p = k;
            q = even - 1 - k;
        }
        __device__ __forceinline__ void advance(int even)
        {
            if (p >= 1)
                p = (p == 1) ? even - 1 : p - 1;
            q = (q == 1) ? even - 1 : q - 1;
        }
    };

    // 1/sqrt(x) for x in [0.25, 2]: the hardware estimate (relative error
    // about 2^-22) refined by two Newton steps y <- y + y (1 - x y^2) / 2.
    __device__ __forceinline__ double rsqrt_refined(double x)
    {
        double y = __builtin_amdgcn_rsq(x);
#pragma unroll
        for (int i = 0; i < 2; ++i)
        {
            const double e = fma(-(x * y), y, 1.0);
            y = fma(0.5 * y, e, y);
        }
        return y;
    }

    // a / b for b in [0.5, 3]: the hardware reciprocal estimate, two Newton
    // steps y <- y + y (1 - b y), then one residual correction of the quotient.
    __device__ __forceinline__ double divide_refined(double a, double b)
    {
        double y = __builtin_amdgcn_rcp(b);
#pragma unroll
        for (int i = 0; i < 2; ++i)
            y = fma(y, fma(-b, y, 1.0), y);
        const double q = a * y;
        return fma(fma(-b, q, a), y, q);
    }

    // c = 1 / sqrt(1 + t^2), s = t c for |t| <= 1.  Exactly c = 1, s = t
    // for t = +-0.
    __device__ __forceinline__ void rotation_from_tangent(double t, double &c, double &s)
    {
        c = rsqrt_refined(fma(t, t, 1.0));
        s = t * c;
    }

    // The rotated pair x' = c x - s y, y' = s x + c y (the reference's update,
    // src/domain/jacobi_rotation_kernels.cuh:334-352) with its roundings
    // fixed: one product, then one fused multiply-add.  Written out so that
    // every kernel that shares these helpers rounds alike, whatever the
    // compiler would contract in each loop.
    __device__ __forceinline__ double rot_first(double c, double s, double x, double y)
    {
        return fma(c, x, -(s * y));
    }
    __device__ __forceinline__ double rot_second(double c, double s, double x, double y)
    {
        return fma(s, x, c * y);
    }

    // Rotation of the pair after the threshold test: Rutishauser's tangent
    // t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), zeta = (qq - pp) / (2 pq)
    // (src/domain/jacobi_rotation_kernels.cuh:313-322), evaluated as
    // t = sign(zeta) |e| / (|d| + sqrt(d^2 + e^2)) with d = qq - pp and e = 2 pq
    // both scaled by the power of two that brings max(|d|, |e|) into
    // [0.5, 1): no overflow of zeta^2 and every operand of the refined
    // hardware estimates in range.  Max error over 4e6 random Gram entries
    // spanning 2^+-100: t 3.1, c 1.6, s 4.0 ulp (the IEEE-exact evaluation
    // of the zeta form: 3.5, 1.9, 4.7 ulp).  zeta = +-0 (d = 0) gives t = 1, and
    // a ratio |d / e| beyond the double range gives t = +-0 where zeta^2
    // overflows in the zeta form.
    __device__ __forceinline__ double rotation_tangent(double pp, double qq, double pq)
    {
        const double d = qq - pp, e = 2.0 * pq;
        const int k = __builtin_amdgcn_frexp_exp(fmax(fabs(d), fabs(e)));
        const double ds = __builtin_amdgcn_ldexp(d, -k), es = __builtin_amdgcn_ldexp(e, -k);
        const double x = fma(ds, ds, es * es);
        const double root = x * rsqrt_refined(x);
        const double tmag = divide_refined(fabs(es), fabs(ds) + root);
        // zeta < 0 exactly when d and pq are nonzero with opposite signs
        // (bitwise operators: no short-circuit branches).
        const bool negative = ((d < 0.0) & (pq > 0.0)) | ((d > 0.0) & (pq < 0.0));
        return negative ? -tmag : tmag;
    }

    // The reference's threshold |pq| > eps * sqrt(pp * qq) and |pq| > 1e-300
    // (src/domain/jacobi_rotation_kernels.cuh:305-312) with the hardware
    // square root (relative error about 2^-23, far below the rounding noise
    // of pq itself near the threshold).
    __device__ __forceinline__ bool rotation_needed(double pp, double qq, double pq, double eps)
    {
        const double rhs = eps * __builtin_amdgcn_sqrt(fmax(pp, 0.0) * fmax(qq, 0.0));
        const double apq = fabs(pq);
        return apq > rhs && apq > 1.0e-300;
    }

    // rotation_needed, then c and s from rotation_tangent; c = 1, s = 0
    // when the pair does not rotate.  Returns true when the pair counts as
    // rotated.
    __device__ __forceinline__ bool rotation_refined(double pp, double qq, double pq, double eps, double &c,
                                                     double &s)
    {
        if (!rotation_needed(pp, qq, pq, eps))
        {
            c = 1.0;
            s = 0.0;
            return false;
        }
        rotation_from_tangent(rotation_tangent(pp, qq, pq), c, s);
        return true;
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

    // Cross-lane move of a double through DPP (two 32-bit halves).  CTRL is a
    // GFX9 DPP control: 0xB1 = quad_perm(1,0,3,2), 0x4E = quad_perm(2,3,0,1),
    // 0x141 = row_half_mirror (lane i <- 7-i in each group of 8),
    // 0x140 = row_mirror (lane i <- 15-i in each row of 16).
    template <int CTRL>
    __device__ __forceinline__ double dpp_f64(double x)
    {
        const long long bits = __double_as_longlong(x);
        const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(bits & 0xffffffffll), CTRL, 0xF, 0xF, true);
        const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(bits >> 32), CTRL, 0xF, 0xF, true);
        return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                    (static_cast<long long>(static_cast<unsigned int>(lo))));
    }

    // x + x[lane ^ 32] and x + x[lane ^ 16] for doubles without an LDS
    // permute: gfx950's v_permlane32_swap (v_permlane16_swap) exchanges the
    // upper half (odd rows of 16) of its first operand with the lower half
    // (even rows) of its second, so with both operands a copy of x the two
    // results hold x and its partner in every lane, and their sum is the
    // butterfly sum (the same value as x + __shfl_xor(x, 32 or 16): addition
    // commutes).
    __device__ __forceinline__ double xor32_sum(double x)
    {
        const long long b = __double_as_longlong(x);
        const auto lo = __builtin_amdgcn_permlane32_swap(static_cast<unsigned int>(b), static_cast<unsigned int>(b),
                                                         false, false);
        const auto hi = __builtin_amdgcn_permlane32_swap(static_cast<unsigned int>(b >> 32),
                                                         static_cast<unsigned int>(b >> 32), false, false);
        const double a = __longlong_as_double(static_cast<long long>(static_cast<unsigned int>(lo[0])) |
                                              (static_cast<long long>(hi[0]) << 32));
        const double c = __longlong_as_double(static_cast<long long>(static_cast<unsigned int>(lo[1])) |
                                              (static_cast<long long>(hi[1]) << 32));
        return a + c;
    }
    __device__ __forceinline__ double xor16_sum(double x)
    {
        const long long b = __double_as_longlong(x);
        const auto lo = __builtin_amdgcn_permlane16_swap(static_cast<unsigned int>(b), static_cast<unsigned int>(b),
                                                         false, false);
        const auto hi = __builtin_amdgcn_permlane16_swap(static_cast<unsigned int>(b >> 32),
                                                         static_cast<unsigned int>(b >> 32), false, false);
        const double a = __longlong_as_double(static_cast<long long>(static_cast<unsigned int>(lo[0])) |
                                              (static_cast<long long>(hi[0]) << 32));
        const double c = __longlong_as_double(static_cast<long long>(static_cast<unsigned int>(lo[1])) |
                                              (static_cast<long long>(hi[1]) << 32));
        return a + c;
    }

    // All-reduce over aligned groups of G lanes (compile-time G <= 64).  The
    // first four levels are DPP moves (no LDS round trip): after the xor-1 and
    // xor-2 levels every lane of a quad holds the quad sum, so the mirror
    // patterns pair each lane with one in the other half of its 8- and
    // 16-lane group.  Wider groups fall back to lane shuffles.
    template <int G>
    __device__ __forceinline__ double group_sum_dpp(double x)
    {
        if constexpr (G >= 2)
            x += dpp_f64<0xB1>(x);
        if constexpr (G >= 4)
            x += dpp_f64<0x4E>(x);
        if constexpr (G >= 8)
            x += dpp_f64<0x141>(x);
        if constexpr (G >= 16)
            x += dpp_f64<0x140>(x);
        if constexpr (G >= 32)
            x = xor16_sum(x);
        if constexpr (G >= 64)
            x = xor32_sum(x);
        return x;
    }

    // Two independent group all-reduces with their DPP levels interleaved.
    template <int G>
    __device__ __forceinline__ void group_sum2_dpp(double &a, double &b)
    {
        if constexpr (G >= 2)
        {
            const double ta = dpp_f64<0xB1>(a), tb = dpp_f64<0xB1>(b);
            a += ta;
            b += tb;
        }
        if constexpr (G >= 4)
        {
            const double ta = dpp_f64<0x4E>(a), tb = dpp_f64<0x4E>(b);
            a += ta;
            b += tb;
        }
        if constexpr (G >= 8)
        {
            const double ta = dpp_f64<0x141>(a), tb = dpp_f64<0x141>(b);
            a += ta;
            b += tb;
        }
        static_assert(G <= 8, "two-value DPP reduction covers groups of up to 8 lanes");
    }

    // Three independent group all-reduces with their DPP levels interleaved
    // (each level's moves read registers the previous level's adds wrote, so
    // three chains fill the wait states one chain would idle through).  Same
    // results as three group_sum_dpp calls.
    template <int G>
    __device__ __forceinline__ void group_sum3_dpp(double &a, double &b, double &c)
    {
        if constexpr (G >= 2)
        {
            const double ta = dpp_f64<0xB1>(a), tb = dpp_f64<0xB1>(b), tc = dpp_f64<0xB1>(c);
            a += ta;
            b += tb;
            c += tc;
        }
        if constexpr (G >= 4)
        {
            const double ta = dpp_f64<0x4E>(a), tb = dpp_f64<0x4E>(b), tc = dpp_f64<0x4E>(c);
            a += ta;
            b += tb;
            c += tc;
        }
        if constexpr (G >= 8)
        {
            const double ta = dpp_f64<0x141>(a), tb = dpp_f64<0x141>(b), tc = dpp_f64<0x141>(c);
            a += ta;
            b += tb;
            c += tc;
        }
        if constexpr (G >= 16)
        {
            const double ta = dpp_f64<0x140>(a), tb = dpp_f64<0x140>(b), tc = dpp_f64<0x140>(c);
            a += ta;
            b += tb;
            c += tc;
        }
        if constexpr (G >= 32)
        {
            a = xor16_sum(a);
            b = xor16_sum(b);
            c = xor16_sum(c);
        }
        if constexpr (G >= 64)
        {
            a = xor32_sum(a);
            b = xor32_sum(b);
            c = xor32_sum(c);
        }
    }

    template <typename T>
    __device__ __forceinline__ T wave_sum(T x)
    {
#pragma unroll
        for (int mask = 32; mask > 0; mask >>= 1)
            x += __shfl_xor(x, mask, 64);
        return x;
    }

    // wave_sum for doubles with the two cross-row levels on permlane swaps.
    template <>
    __device__ __forceinline__ double wave_sum<double>(double x)
    {
        x = xor32_sum(x);
        x = xor16_sum(x);
#pragma unroll
        for (int mask = 8; mask > 0; mask >>= 1)
            x += __shfl_xor(x, mask, 64);
        return x;
    }

    // Bound on the software barriers' waits.  A wait that lasts more than
    // `limit` ticks of the constant-rate wall clock (wall_clock64, rate from
    // hipDeviceAttributeWallClockRate) stores 1 to *fault and gives up; any
    // wait that finds *fault set gives up at once, so every workgroup of the
    // launch leaves its loops within one bound and the grid drains.  Such a
    // wait means the workgroups were not all resident (another process's
    // kernels holding CUs), which the occupancy-sized grids assume.
    struct SpinBound
    {
        unsigned int *fault; // zero at launch
        uint64_t limit;      // wall-clock ticks
    };

    __device__ __forceinline__ bool spin_expired(const SpinBound &b, uint64_t t0)
    {
        if (__hip_atomic_load(b.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
            return true;
        if (static_cast<uint64_t>(wall_clock64()) - t0 > b.limit)
        {
            __hip_atomic_store(b.fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return true;
        }
        return false;
    }

    // Sense-free generation barrier across `nblocks` co-resident workgroups
    // (a launch's grid, or one group of it; grid <= resident workgroups, one
    // such launch on the GPU at a time).  bar[0] counts arrivals, bar[1] is
    // the generation.  Each workgroup's arrival is a release RMW (its
    // columns, written before the workgroup barrier, reach the point of
    // agent coherence first); the last arrival acquires every other
    // arrival's writes through the RMW chain, then publishes the new
    // generation with a release store; the others poll with relaxed loads
    // (an acquire load would invalidate this XCD's L2 on every iteration)
    // and acquire once, after the generation has changed.  So each
    // workgroup pays one L2 write-back and one invalidate per barrier.
    // Returns false (in every thread) when the wait hit the spin bound.
    __device__ __forceinline__ bool grid_barrier(unsigned int *bar, unsigned int nblocks, const SpinBound &bound)
    {
        __syncthreads();
        int failed = 0;
        if (threadIdx.x == 0)
        {
            const unsigned int gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned int arrived =
                __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            if (arrived == nblocks - 1)
            {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            else
            {
                const uint64_t t0 = static_cast<uint64_t>(wall_clock64());
                while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen)
                {
                    if (spin_expired(bound, t0))
                    {
                        failed = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
        }
        return !__syncthreads_or(failed);
    }
    // Barrier among `members` (at most 256) co-resident workgroups through one
    // flag word each: workgroup `member` publishes its barrier count `epoch`
    // with a release fence and a relaxed store, then wave 0 polls every
    // member's word at once (lane l reads words l, l + 64, ...; relaxed
    // loads) until all have reached `epoch`, and acquires.  Two trips to the point of agent
    // coherence per barrier (the store, the loads that see the last store)
    // instead of the arrival RMW, the generation store and the poll of the
    // counter barrier.  Flags start at zero and only grow within a launch.
    // Returns false (in every thread) when the wait hit the spin bound.
    __device__ __forceinline__ bool flag_barrier(unsigned int *flags, int members, int member, unsigned int epoch,
                                                 const SpinBound &bound)
    {
        __syncthreads();
        int failed = 0;
        if (threadIdx.x < 64)
        {
            if (threadIdx.x == 0)
            {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_store(&flags[member], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            const int l = threadIdx.x;
            const uint64_t t0 = static_cast<uint64_t>(wall_clock64());
            for (;;)
            {
                bool ready = true;
                for (int w = l; w < members; w += 64)
                    ready &= __hip_atomic_load(&flags[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= epoch;
                if (__all(ready))
                    break;
                // Lane 0 decides for the wave, so the loop is left together.
                const bool expired = l == 0 && spin_expired(bound, t0);
                if (__any(expired))
                {
                    failed = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        return !__syncthreads_or(failed);
    }
} // namespace jsvd::dev
