// jsvd_kernels_ring.hip -- ring form of the grouped plain-order sweeps:
// mid-size matrices (at most 1024 rows and 512 columns) solved by groups of
// workgroups that never meet at a group barrier inside a sweep.
//
// Every column pair position k (0 <= k < half) belongs to one wave for the
// whole solve: wave k % W of workgroup k / W holds the pair's two columns of
// A and of V in registers (lane l: rows l + 64 j; RA row registers for A, RV
// for V; eight-wave workgroups up to 512 rows, four-wave ones above).  In the reference's circle
// (src/domain/jacobi_schedule.cuh:16-57, src/domain/jacobi_rotation_kernels.cuh:18-42)
// every column except column 0 moves to the next circle slot each round, so
// after a round's rotations
//   top(k) -> top(k + 1)                 for 1 <= k <= half - 2,
//   top(half - 1) -> bottom(half - 1),
//   bottom(k) -> bottom(k - 1)           for 1 <= k <= half - 1,
//   bottom(0) -> top(1),
// and top(0) keeps column 0.  Inside a workgroup the moved columns go
// through LDS (two workgroup barriers per round); between workgroups only
// two columns cross each boundary per round: the last wave's top forward and
// the first wave's bottom back.  They travel through mailboxes of 64-bit
// words, each carrying half of a double in its low 32 bits and the round
// number in its high 32 bits, written and read with relaxed agent-scope
// atomics: a reader takes a word once its tag is the current round, so no
// fence, flag or barrier orders the data (two mailboxes per direction,
// alternating by round parity; neighbours are never more than one round
// apart, because each needs the other's previous round).  A sweep ends with
// the group's flag barrier and the shared sweep flag, as in the grouped
// kernel.
//
// The pair arithmetic is wave_pair_f64's (same dot chains over the same row
// order, same reduction, rotation parameters and update roundings) and the
// pairs come in the same order, so the results equal the grouped kernel's
// bitwise.  After every whole sweep the columns are back at their starting
// positions (top(k) = column k, bottom(k) = column E - 1 - k).  The kernel
// reads A straight from the caller's packed row-major buffer, starts V as
// the identity in registers, and at the end writes sigma, U = A / sigma
// (zero where sigma <= eps, the reference's rule) and V row-major, with
// column norms summed in column_norms_f64's order: no workspace, no helper
// launches.
#include "jsvd_device.hpp"
#include "jsvd_kernels.hpp"

namespace jsvd
{
    namespace
    {
        constexpr int kRingRowsMax = 8;   // row registers per lane of the eight-wave form: m <= 512
        constexpr int kRingTallRows = 16; // the four-wave form: m <= 1024 (and n <= 512)

        // Words of one mailbox message: the A column (RA words of 64 lanes x
        // two halves), then the V column (RV words).
        template <int RA, int RV>
        constexpr int kRingMsg = 2 * 64 * (RA + RV);

        __device__ __forceinline__ unsigned long long tagged(unsigned int tag, unsigned int bits)
        {
            return (static_cast<unsigned long long>(tag) << 32) | bits;
        }

        // Posts the two columns.  With `local` (every member of the group on
        // one XCD) the words are stored at workgroup scope: they stop in the
        // XCD's shared L2, where the readers' agent-scope loads find them
        // (one-way latency about 1.2 us against 1.6 us for agent-scope
        // stores, which travel to the point of device coherence).
        template <int RA, int RV>
        __device__ __forceinline__ void ring_send(unsigned long long *msg, unsigned int tag, const double (&a)[RA],
                                                  const double (&v)[RV], int lane, bool local)
        {
#pragma unroll
            for (int j = 0; j < RA + RV; ++j)
                {
                    const unsigned long long bits =
                        static_cast<unsigned long long>(__double_as_longlong(j < RA ? a[j] : v[j - RA]));
                    unsigned long long *w = msg + j * 128 + lane * 2;
                    const unsigned long long lo = tagged(tag, static_cast<unsigned int>(bits));
                    const unsigned long long hi = tagged(tag, static_cast<unsigned int>(bits >> 32));
                    if (local)
                    {
                        __hip_atomic_store(w, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        __hip_atomic_store(w + 1, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                    else
                    {
                        __hip_atomic_store(w, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_store(w + 1, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
        }

        // Waits (bounded) until every word of the message carries `tag`, then
        // assembles the two columns.  Returns false when the wait expired.
        // The whole message is read at once (one round trip when it has
        // arrived); while it has not, each lane polls only the last word it
        // reads (the last one its sender lane writes), so a waiting wave
        // moves 512 bytes per poll instead of the whole message.
        template <int RA, int RV>
        __device__ __forceinline__ bool ring_receive(const unsigned long long *msg, unsigned int tag, double (&a)[RA],
                                                     double (&v)[RV], int lane, const dev::SpinBound &bound)
        {
            const uint64_t t0 = static_cast<uint64_t>(wall_clock64());
            const unsigned long long *last = msg + (RA + RV - 1) * 128 + lane * 2 + 1;
            for (;;)
            {
                bool ready = true;
#pragma unroll
                for (int j = 0; j < RA + RV; ++j)
                    {
                        const unsigned long long *w = msg + j * 128 + lane * 2;
                        const unsigned long long lo =
                            __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        const unsigned long long hi =
                            __hip_atomic_load(w + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ready &= static_cast<unsigned int>(lo >> 32) == tag && static_cast<unsigned int>(hi >> 32) == tag;
                        const double x = __longlong_as_double(static_cast<long long>((hi << 32) | (lo & 0xffffffffull)));
                        if (j < RA)
                            a[j] = x;
                        else
                            v[j - RA] = x;
                    }
                if (__all(ready))
                    return true;
                for (;;)
                {
                    const bool expired = lane == 0 && dev::spin_expired(bound, t0);
                    if (__any(expired))
                        return false;
                    const unsigned long long x = __hip_atomic_load(last, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (__all(static_cast<unsigned int>(x >> 32) == tag))
                        break;
                }
            }
        }

        // wave_pair_f64 on register-held columns: p is the top column, q the
        // bottom one.  Rows beyond m (and V rows beyond n) hold zeros, which
        // leave the dot chains and the updates unchanged.
        template <int RA, int RV>
        __device__ __forceinline__ int ring_pair(double (&xp)[RA], double (&xq)[RA], double (&yp)[RV],
                                                 double (&yq)[RV], double eps)
        {
            double pp = 0.0, qq = 0.0, pq = 0.0;
#pragma unroll
            for (int j = 0; j < RA; ++j)
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
            for (int j = 0; j < RA; ++j)
            {
                const double a = xp[j], b = xq[j];
                xp[j] = dev::rot_first(c, s, a, b);
                xq[j] = dev::rot_second(c, s, a, b);
            }
#pragma unroll
            for (int j = 0; j < RV; ++j)
            {
                const double a = yp[j], b = yq[j];
                yp[j] = dev::rot_first(c, s, a, b);
                yq[j] = dev::rot_second(c, s, a, b);
            }
            return 1;
        }

        // ||x|| of a register-held column in column_norms_f64<256>'s order:
        // its thread t = lane + 64 w sums rows t, t + 256, ..., i.e. this
        // lane's row registers j = w, w + 4, ... (FMA chain); each wave's sums
        // are reduced by dev::wave_sum, then the four totals added in order.
        // Rows beyond m hold zeros, which leave the chains unchanged.
        template <int RA>
        __device__ __forceinline__ double ring_norm(const double (&x)[RA])
        {
            double red[4];
#pragma unroll
            for (int w = 0; w < 4; ++w)
            {
                double acc = 0.0;
#pragma unroll
                for (int j = w; j < RA; j += 4)
                    acc = fma(x[j], x[j], acc);
                red[w] = dev::wave_sum(acc);
            }
            const double t = ((red[0] + red[1]) + red[2]) + red[3];
            return sqrt(fmax(t, 0.0));
        }

        template <int RA, int RV>
        __device__ __forceinline__ void lds_put(double *slot, const double (&a)[RA], const double (&v)[RV], int lane)
        {
#pragma unroll
            for (int j = 0; j < RA; ++j)
                slot[j * 64 + lane] = a[j];
#pragma unroll
            for (int j = 0; j < RV; ++j)
                slot[(RA + j) * 64 + lane] = v[j];
        }

        template <int RA, int RV>
        __device__ __forceinline__ void lds_get(const double *slot, double (&a)[RA], double (&v)[RV], int lane)
        {
#pragma unroll
            for (int j = 0; j < RA; ++j)
                a[j] = slot[j * 64 + lane];
#pragma unroll
            for (int j = 0; j < RV; ++j)
                v[j] = slot[(RA + j) * 64 + lane];
        }

        // LDS: per wave a top slot and a bottom slot of (RA + RV) x 64
        // doubles (A column, then V column).  Mailboxes of a group: for each
        // boundary b (between workgroups b and b + 1), direction d (0: top
        // forward, 1: bottom back) and round parity, one message.
        template <int RA, int RV, int W>
        __global__ __launch_bounds__(64 * W) void wave_ring_sweeps_f64(WaveGroupArgs g)
        {
            constexpr int MSG = kRingMsg<RA, RV>, SLOT = (RA + RV) * 64;
            extern __shared__ double ring_lds[];
            const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
            int group, member;
            if (g.xcd_spread > 0)
            {
                const int x = blockIdx.x % g.xcd_spread, slot = blockIdx.x / g.xcd_spread;
                member = slot % g.members;
                group = (slot / g.members) * g.xcd_spread + x;
            }
            else
            {
                group = blockIdx.x / g.members;
                member = blockIdx.x - group * g.members;
            }
            if (group >= g.ring_groups)
                return; // a slot of the XCD-spread grid without a group
            const int groups = g.ring_groups;
            unsigned int *bar = g.bars + group * g.members;
            unsigned long long *mb = g.mbox + group * g.mbox_group;
            const dev::SpinBound bound{g.fault, g.spin_limit};
            // With two slot sets (g.ring_db) round r uses set r % 2: a set is
            // written two rounds after its last reads, and the barrier of the
            // round between orders them, so the barrier before the writes
            // goes.
            const int set_stride = g.ring_db ? 2 * W * SLOT : 0;
            const int k = member * W + wave;
            unsigned int epoch = 0, tag = 0;
            int failed = 0;
            // Whether the group's members share one XCD (and so one L2):
            // each publishes its XCC id, and after the group's barrier
            // every member compares them.  Only then are the mailbox words
            // stored at workgroup scope.
            bool local = false;
            if (g.xcd_spread > 0)
            {
                unsigned int *ids = g.xcc + group * g.members;
                if (threadIdx.x == 0)
                {
                    unsigned int id;
                    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(id));
                    __hip_atomic_store(&ids[member], (id & 0xffu) + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                if (!dev::flag_barrier(bar, g.members, member, ++epoch, bound))
                    failed = 1;
                int same = 1;
                const unsigned int first = __hip_atomic_load(&ids[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                for (int q = threadIdx.x; q < g.members; q += blockDim.x)
                    same &= __hip_atomic_load(&ids[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == first;
                local = !__syncthreads_or(!same) && !failed;
            }
            for (int idx = group; idx < g.count; idx += groups)
            {
                const WaveMatrix w = g.mats[idx];
                const int64_t m = w.m;
                const int n = w.n;
                const int even = n + (n & 1), half = even >> 1;
                const int rounds = n < 2 ? 0 : even - 1;
                const bool active = k < half;
                const int ct = k, cb = even - 1 - k;
                // Columns ct and cb of the row-major A (each load touches 64
                // rows; once per solve), and of V = I.
                const double *a_in = g.in_a + w.a_off;
                double tA[RA], tV[RV], bA[RA], bV[RV];
#pragma unroll
                for (int j = 0; j < RA; ++j)
                {
                    const int64_t i = lane + 64 * j;
                    tA[j] = (active && ct < n && i < m) ? a_in[i * n + ct] : 0.0;
                    bA[j] = (active && cb < n && i < m) ? a_in[i * n + cb] : 0.0;
                }
#pragma unroll
                for (int j = 0; j < RV; ++j)
                {
                    const int i = lane + 64 * j;
                    tV[j] = (active && ct < n && i == ct) ? 1.0 : 0.0;
                    bV[j] = (active && cb < n && i == cb) ? 1.0 : 0.0;
                }
                // Message slots of this workgroup's boundaries.
                unsigned long long *fwd_out = mb + int64_t(member * 2 + 0) * 2 * MSG;       // top to member + 1
                unsigned long long *back_in = mb + int64_t(member * 2 + 1) * 2 * MSG;       // bottom from member + 1
                const int left = member > 0 ? member - 1 : 0; // used only when member > 0
                unsigned long long *fwd_in = mb + int64_t(left * 2 + 0) * 2 * MSG;  // top from member - 1
                unsigned long long *back_out = mb + int64_t(left * 2 + 1) * 2 * MSG; // bottom to member - 1
                // The last wave sends its top forward (and, as position
                // k + 1 < half exists, receives that position's bottom);
                // the first wave of a later workgroup sends its bottom back
                // (and receives the previous position's top).
                const bool send_fwd = active && wave == W - 1 && k + 1 < half;
                const bool send_back = active && wave == 0 && member > 0;
                int executed = 0, clean = 0;
                for (int sweep = 0; sweep < g.max_sweeps && !failed; ++sweep)
                {
                    int any = 0;
                    for (int r = 0; r < rounds; ++r)
                    {
                        ++tag;
                        const int par = static_cast<int>(tag & 1u) * MSG;
                        double *my_top = ring_lds + static_cast<int>(tag & 1u) * set_stride + (2 * wave) * SLOT;
                        double *my_bot = my_top + SLOT;
                        if (active)
                            any |= ring_pair<RA, RV>(tA, bA, tV, bV, g.eps);
                        // The boundary columns leave at once, so they travel
                        // while the workgroup exchanges the others.  The
                        // receiver has read this parity's previous message:
                        // it needed it to finish the round before the one
                        // this workgroup just completed.
                        if (send_fwd)
                            ring_send<RA, RV>(fwd_out + par, tag, tA, tV, lane, local);
                        if (send_back)
                            ring_send<RA, RV>(back_out + par, tag, bA, bV, lane, local);
                        // One slot set: every read of the previous round's
                        // slots is done (and every wave agrees on a failed
                        // wait) before they are written.
                        if (!g.ring_db && __syncthreads_or(failed))
                        {
                            failed = 1;
                            break;
                        }
                        if (active)
                        {
                            lds_put<RA, RV>(my_top, tA, tV, lane);
                            lds_put<RA, RV>(my_bot, bA, bV, lane);
                        }
                        // Two sets: a failed wait of the previous round is
                        // agreed on here, one round later.
                        if (g.ring_db)
                        {
                            if (__syncthreads_or(failed))
                            {
                                failed = 1;
                                break;
                            }
                        }
                        else
                            __syncthreads();
                        if (active)
                        {
                            // New bottom first: at k = half - 1 it is this
                            // wave's own top, which the new top replaces
                            // (with one pair, half = 1, both columns stay).
                            if (k == half - 1)
                            {
                                if (half > 1)
                                {
#pragma unroll
                                    for (int j = 0; j < RA; ++j)
                                        bA[j] = tA[j];
#pragma unroll
                                    for (int j = 0; j < RV; ++j)
                                        bV[j] = tV[j];
                                }
                            }
                            else if (wave < W - 1)
                                lds_get<RA, RV>(my_bot + 2 * SLOT, bA, bV, lane); // wave + 1's bottom
                            else if (!ring_receive<RA, RV>(back_in + par, tag, bA, bV, lane, bound))
                                failed = 1;
                            if (k == 1)
                                lds_get<RA, RV>(my_bot - 2 * SLOT, tA, tV, lane); // wave 0's bottom
                            else if (k >= 2 && wave > 0)
                                lds_get<RA, RV>(my_top - 2 * SLOT, tA, tV, lane); // wave - 1's top
                            else if (k >= 2 && !ring_receive<RA, RV>(fwd_in + par, tag, tA, tV, lane, bound))
                                failed = 1;
                        }
                    }
                    failed = __syncthreads_or(failed);
                    if (failed)
                        break;
                    executed = sweep + 1;
                    if (any && lane == 0)
                        __hip_atomic_fetch_or(&w.flags[sweep], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (!dev::flag_barrier(bar, g.members, member, ++epoch, bound))
                    {
                        failed = 1;
                        break;
                    }
                    const int f = __hip_atomic_load(&w.flags[sweep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (f == 0)
                    {
                        clean = 1;
                        break;
                    }
                }
                if (!failed && active)
                {
                    // sigma, then U = A / sigma (zero where sigma <= eps,
                    // src/domain/jacobi_rotation_kernels.cuh:370-413) and V,
                    // row-major like the transposes of the barrier forms.
                    const double st = ring_norm<RA>(tA), sb = ring_norm<RA>(bA);
                    double *u = g.out_u + w.a_off;
                    double *vv = g.out_v + w.v_off;
                    if (lane == 0)
                    {
                        g.out_s[w.s_off + ct] = st;
                        if (cb < n)
                            g.out_s[w.s_off + cb] = sb;
                    }
#pragma unroll
                    for (int j = 0; j < RA; ++j)
                    {
                        const int64_t i = lane + 64 * j;
                        if (i < m)
                        {
                            u[i * n + ct] = (st > g.eps) ? (tA[j] / st) : 0.0;
                            if (cb < n)
                                u[i * n + cb] = (sb > g.eps) ? (bA[j] / sb) : 0.0;
                        }
                    }
#pragma unroll
                    for (int j = 0; j < RV; ++j)
                    {
                        const int i = lane + 64 * j;
                        if (i < n)
                        {
                            vv[int64_t(i) * n + ct] = tV[j];
                            if (cb < n)
                                vv[int64_t(i) * n + cb] = bV[j];
                        }
                    }
                }
                if (member == 0 && threadIdx.x == 0 && g.sweeps)
                    g.sweeps[w.index] = failed ? 0 : executed;
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

        // Instances: eight-wave workgroups with RA = RV = ceil(m / 64) for m
        // <= 512; four-wave workgroups with 16 A rows and ceil(n / 64) V rows
        // per lane for 512 < m <= 1024 (n <= 512).
        template <int RA, int RV, int W>
        const void *ring_symbol_t()
        {
            return reinterpret_cast<const void *>(&wave_ring_sweeps_f64<RA, RV, W>);
        }
        const void *(*const kRingSymbol[])() = {ring_symbol_t<1, 1, 8>, ring_symbol_t<2, 2, 8>, ring_symbol_t<3, 3, 8>,
                                                ring_symbol_t<4, 4, 8>, ring_symbol_t<5, 5, 8>, ring_symbol_t<6, 6, 8>,
                                                ring_symbol_t<7, 7, 8>, ring_symbol_t<8, 8, 8>};
        static_assert(sizeof(kRingSymbol) / sizeof(kRingSymbol[0]) == kRingRowsMax, "one instance per row count");
        const void *(*const kRingTallSymbol[])() = {
            ring_symbol_t<16, 1, 4>, ring_symbol_t<16, 2, 4>, ring_symbol_t<16, 3, 4>, ring_symbol_t<16, 4, 4>,
            ring_symbol_t<16, 5, 4>, ring_symbol_t<16, 6, 4>, ring_symbol_t<16, 7, 4>, ring_symbol_t<16, 8, 4>};
        static_assert(sizeof(kRingTallSymbol) / sizeof(kRingTallSymbol[0]) == kRingRowsMax,
                      "one tall instance per V row count");

        int rows_of(int64_t x)
        {
            return static_cast<int>((x + 63) / 64);
        }

        // (A rows, V rows, waves) per lane of the instance for a launch whose
        // largest matrix has m rows and n columns; RA = 0 when none fits.
        void ring_shape(int64_t m, int64_t n, int &ra, int &rv, int &waves)
        {
            ra = rv = 0;
            waves = 8;
            if (m < 1 || n < 1 || n > m)
                return;
            if (m <= 64 * kRingRowsMax)
            {
                ra = rv = rows_of(m);
                return;
            }
            if (m <= 64 * kRingTallRows && n <= 64 * kRingRowsMax)
            {
                ra = kRingTallRows;
                rv = rows_of(n);
                waves = 4; // eight waves of this many registers would not be resident together
            }
        }
    } // namespace

    bool ring_fits(int64_t m, int64_t n)
    {
        int ra, rv, w;
        ring_shape(m, n, ra, rv, w);
        return ra > 0;
    }

    int ring_waves(int64_t m, int64_t n)
    {
        int ra, rv, w;
        ring_shape(m, n, ra, rv, w);
        return w;
    }

    int ring_lds_bytes(int64_t m, int64_t n)
    {
        // Two slots per wave of (RA + RV) x 64 doubles.
        int ra, rv, w;
        ring_shape(m, n, ra, rv, w);
        return w * 2 * (ra + rv) * 64 * 8;
    }

    int64_t ring_mailbox_words(int64_t m, int64_t n, int members)
    {
        // Per workgroup (as the left end of a boundary): two directions, two
        // parities, one message of 2 x 64 x (RA + RV) words each.
        int ra, rv, w;
        ring_shape(m, n, ra, rv, w);
        return int64_t(members) * 2 * 2 * (2 * 64 * (ra + rv));
    }

    const void *wave_ring_sweeps_f64_symbol(int64_t m, int64_t n)
    {
        int ra, rv, w;
        ring_shape(m, n, ra, rv, w);
        if (ra == 0)
            return nullptr;
        return w == 8 ? kRingSymbol[ra - 1]() : kRingTallSymbol[rv - 1]();
    }

    hipError_t launch_wave_ring_sweeps_f64(int64_t m, int64_t n, int groups, const WaveGroupArgs &args,
                                           hipStream_t stream)
    {
        const void *fn = wave_ring_sweeps_f64_symbol(m, n);
        if (!fn || groups < 1 || args.members < 1 || args.count < 1 || !args.mbox || !args.in_a || !args.out_u ||
            !args.out_s || !args.out_v || args.ring_groups != groups || (args.xcd_spread > 0 && !args.xcc))
            return hipErrorInvalidValue;
        const unsigned int spread = args.xcd_spread > 0 ? static_cast<unsigned int>(args.xcd_spread) : 1u;
        const unsigned int grid =
            ((static_cast<unsigned int>(groups) + spread - 1) / spread) * spread * static_cast<unsigned int>(args.members);
        void *params[] = {const_cast<WaveGroupArgs *>(&args)};
        return hipLaunchKernel(fn, dim3(grid), dim3(64 * ring_waves(m, n)), params,
                               ring_lds_bytes(m, n) * (args.ring_db ? 2 : 1), stream);
    }
} // namespace jsvd
