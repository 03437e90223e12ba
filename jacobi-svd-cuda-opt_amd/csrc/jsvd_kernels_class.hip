// jsvd_kernels_class.hip -- LDS-resident FP64 work-list kernels with
// compile-time shape classes (the batched hot path).
//
// One workgroup solves one matrix at a time, claimed from a device work-list:
// A (and V) are loaded into LDS once, every sweep runs with workgroup
// barriers only, the convergence decision is taken on the device and U,
// sigma, V are written once.
//
// Layout: columns are stored with a leading dimension padded with zero rows
// to a multiple of 2G, so a lane moves two consecutive rows with one 16-byte
// LDS access and the G lanes of a pair read 16*G contiguous bytes (no bank
// conflicts).  Lane l of a group owns row pairs l, l+G, ... of both columns,
// kept in registers between the dots and the update: every element crosses
// LDS once in and once out per round.  Group sums use DPP.
//
// Two ways to carry V:
//  * in LDS next to A (classes with vlog = false);
//  * as a rotation log (vlog = true): for matrices where A fits in LDS but
//    A and V together do not.  Each pair writes its tangent t (0 = no
//    rotation) to a per-workgroup log in HBM; after U and sigma are written,
//    V is rebuilt in LDS by replaying the log in the same order.  c and s are
//    recomputed from t with the same expressions, so the replay applies
//    bit-identical rotations.
//
// Arithmetic per pair follows the reference kernel
// (src/domain/jacobi_rotation_kernels.cuh:276-353): dots, the
// |pq| > eps*sqrt(pp*qq) test, Rutishauser's rotation, then
// a_p <- c a_p - s a_q, a_q <- s a_p + c a_q on A and V, in the reference's
// circle order.
#include "jsvd_device.hpp"
#include "jsvd_kernels.hpp"

#include <algorithm>

namespace jsvd
{
    namespace
    {
        struct ClassSpec
        {
            int block, group, rmax2, vmax2;
            bool vlog;
            int passes; // pair positions per lane group and round
            bool split; // half the waves rotate A, the other half V one round later
        };

        // Shape classes.  The default for a shape is the first fitting class
        // in select_class_f64's preference order; the autotuner may choose
        // another fitting class per column-count bucket.  rmax2 / vmax2 are register row pairs per lane for A and V.
        constexpr ClassSpec kClasses[] = {
            {64, 16, 1, 1, false, 1},  // 0: n <= 8, m <= 32: one wave, 16 lanes per pair
            {64, 8, 2, 1, false, 1},   // 1: n <= 16, m <= 32
            {64, 4, 8, 4, false, 1},   // 2: n <= 32, m <= 64
            {512, 16, 4, 2, false, 1}, // 3: n <= 64, m <= 128: eight waves
            {256, 8, 8, 4, false, 1},  // 4: n <= 64, m <= 128: four waves
            {512, 16, 8, 4, false, 2}, // 5: n <= 128, m <= 256
            {1024, 16, 4, 4, true, 1}, // 6: n <= 128, m <= 128, V by rotation log
            {512, 16, 8, 8, true, 4},  // 7: n <= 256, m <= 256, V by rotation log
            {512, 16, 4, 2, true, 1},  // 8: n <= 64, m <= 128, V by rotation log (two workgroups per CU)
            {512, 8, 8, 4, false, 1, true}, // 9: n <= 64, m <= 128: four waves on A, four on V
        };
        constexpr int kClassCount = sizeof(kClasses) / sizeof(kClasses[0]);

        __host__ __device__ __forceinline__ int padded_rows(int rows, int G)
        {
            return ((rows + 2 * G - 1) / (2 * G)) * (2 * G);
        }

        // A lane's share of one column in LDS: blocks j = 0 .. N-1 of G row
        // pairs (block j holds row pairs jG + lane).  With swap set, the
        // blocks of each aligned pair (0,1), (2,3), ... trade places (the
        // last block of an odd N stays).  A 16-byte LDS access serves 16
        // lanes (256 bytes, all 64 banks) per cycle: two 8-lane groups.  When
        // the column stride is a multiple of 256 bytes, the neighbouring
        // columns of two such groups start on the same banks; swapping in
        // every odd group puts their blocks 128 bytes apart instead.  Only the
        // order in which a lane visits its blocks changes.
        template <int G, int N>
        struct ColumnShare
        {
            double2 *even, *odd, *last;
            __device__ __forceinline__ ColumnShare(double2 *col, int lane, int swap)
                : even(col + lane + swap * G), odd(col + lane + (1 - swap) * G), last(col + lane)
            {
            }
            // j must be a compile-time constant after unrolling.
            __device__ __forceinline__ double2 &operator[](int j) const
            {
                if ((N & 1) && j == N - 1)
                    return last[j * G];
                return (j & 1) ? odd[(j - 1) * G] : even[j * G];
            }
        };

        template <int G, int N>
        __device__ __forceinline__ void load_share(double2 (&x)[N], double2 *col, int lane, int swap)
        {
            const ColumnShare<G, N> c(col, lane, swap);
#pragma unroll
            for (int j = 0; j < N; ++j)
                x[j] = c[j];
        }

        template <int G, int N>
        __device__ __forceinline__ void store_share(const double2 (&x)[N], double2 *col, int lane, int swap)
        {
            const ColumnShare<G, N> c(col, lane, swap);
#pragma unroll
            for (int j = 0; j < N; ++j)
                c[j] = x[j];
        }

        // Rotation of one pair held by G lanes.  The top column (p) is held
        // in registers by the lane group (xp for A, yp for V), the bottom
        // column (q) is read from and written back to LDS.  Each lane owns R
        // row pairs of A and VR of V (the matrix's exact counts, template
        // arguments, so every row loop is unrolled without bounds tests and
        // all loads are issued before any use).  Returns 1 when the pair
        // counts as rotated (reference flag semantics).  With VLOG the
        // group's lane 0 records t (0 when nothing is applied) at *tlog and
        // V is not touched.
        template <int G, int R, int VR, bool VLOG>
        __device__ __forceinline__ int pair_class_f64(double2 (&xp)[R], double2 (&yp)[VLOG ? 1 : VR],
                                                      double2 *__restrict__ acq, double2 *__restrict__ vcq,
                                                      int lane, int swap_a, int swap_v, double eps,
                                                      double *__restrict__ tlog)
        {
            const ColumnShare<G, R> aq(acq, lane, swap_a);
            double2 xq[R];
#pragma unroll
            for (int j = 0; j < R; ++j)
                xq[j] = aq[j];
            // Two FMA chains per dot (even and odd rows), joined at the end.
            double pp0 = 0.0, pp1 = 0.0, qq0 = 0.0, qq1 = 0.0, pq0 = 0.0, pq1 = 0.0;
#pragma unroll
            for (int j = 0; j < R; ++j)
            {
                pp0 = fma(xp[j].x, xp[j].x, pp0);
                pp1 = fma(xp[j].y, xp[j].y, pp1);
                qq0 = fma(xq[j].x, xq[j].x, qq0);
                qq1 = fma(xq[j].y, xq[j].y, qq1);
                pq0 = fma(xp[j].x, xq[j].x, pq0);
                pq1 = fma(xp[j].y, xq[j].y, pq1);
            }
            double pp = pp0 + pp1, qq = qq0 + qq1, pq = pq0 + pq1;
            dev::group_sum3_dpp<G>(pp, qq, pq);
            // Threshold test first (reference jacobi_rotation_kernels.cuh:310-312),
            // then the V loads, issued before the rotation parameters so
            // that their latency overlaps that computation.
            if (!dev::rotation_needed(pp, qq, pq, eps))
            {
                if (VLOG && lane == 0)
                    *tlog = 0.0;
                return 0;
            }
            double2 yq[VLOG ? 1 : VR];
            const ColumnShare<G, VR> vq(vcq, lane, swap_v);
            if constexpr (!VLOG)
            {
#pragma unroll
                for (int j = 0; j < VR; ++j)
                    yq[j] = vq[j];
                __builtin_amdgcn_sched_barrier(0);
            }
            const double t = dev::rotation_tangent(pp, qq, pq);
            double c, s;
            dev::rotation_from_tangent(t, c, s);
            if (VLOG && lane == 0)
                *tlog = t;
            // t = +-0 (a tangent below the double range) gives c = 1, s = 0:
            // the updates below leave both columns unchanged.
            // The new bottom row pair first, into the registers of the old
            // one (stored at once), then the new top in place: no register
            // copies to carry the held top column into the next round.
#pragma unroll
            for (int j = 0; j < R; ++j)
            {
                const double2 a = xp[j], b = xq[j];
                xq[j].x = s * a.x + c * b.x;
                xq[j].y = s * a.y + c * b.y;
                aq[j] = xq[j];
                xp[j].x = c * a.x - s * b.x;
                xp[j].y = c * a.y - s * b.y;
            }
            if constexpr (!VLOG)
            {
#pragma unroll
                for (int j = 0; j < VR; ++j)
                {
                    const double2 a = yp[j], b = yq[j];
                    yq[j].x = s * a.x + c * b.x;
                    yq[j].y = s * a.y + c * b.y;
                    vq[j] = yq[j];
                    yp[j].x = c * a.x - s * b.x;
                    yp[j].y = c * a.y - s * b.y;
                }
            }
            return 1;
        }

        // Replays one logged rotation on the V columns of a pair.
        template <int G, int VR>
        __device__ __forceinline__ void replay_pair_f64(double2 *__restrict__ vcp, double2 *__restrict__ vcq, int lane,
                                                        int swap_v, double t)
        {
            if (t == 0.0)
                return;
            double c, s;
            dev::rotation_from_tangent(t, c, s);
            const ColumnShare<G, VR> vp(vcp, lane, swap_v), vq(vcq, lane, swap_v);
            double2 yp[VR], yq[VR];
#pragma unroll
            for (int j = 0; j < VR; ++j)
            {
                yp[j] = vp[j];
                yq[j] = vq[j];
            }
#pragma unroll
            for (int j = 0; j < VR; ++j)
            {
                double2 np, nq;
                np.x = c * yp[j].x - s * yq[j].x;
                np.y = c * yp[j].y - s * yq[j].y;
                nq.x = s * yp[j].x + c * yq[j].x;
                nq.y = s * yp[j].y + c * yq[j].y;
                vp[j] = np;
                vq[j] = nq;
            }
        }

        // Sweeps with the waves split in two roles.  The first BS/2 threads
        // (A waves) run the circle on A exactly as the one-role loop does,
        // with V left out: each pair's tangent t (0 when not rotated) goes to
        // a two-round ring T in LDS.  The other BS/2 threads (V waves) follow
        // one round behind with the same positions and held top columns, and
        // apply c, s recomputed from t to V.  V's loads, stores and
        // arithmetic thereby leave the A waves' critical path (they overlap
        // the next round's dots and rotation parameters), and each role
        // holds only its own columns in registers.  c and s come from t by
        // the same expressions in both roles, so V receives the rotations
        // applied to A bit for bit.  Returns the number of sweeps.
        template <int BS, int G, int R, int VR>
        __device__ __forceinline__ int sweeps_split_f64(int n, int lda, int ldv, double *A, double *V, double *T,
                                                        const ClassArgs &args, int &clean_out)
        {
            constexpr int HALF_BS = BS / 2;
            const int tid = threadIdx.x;
            const int role = tid >= HALF_BS ? 1 : 0; // 0: A waves, 1: V waves
            const int t2 = tid - role * HALF_BS;
            const int k = t2 / G, lane = t2 & (G - 1);
            const int even = n + (n & 1), half = even >> 1, ring = even - 1;
            const int rounds = n < 2 ? 0 : even - 1;
            double2 *A2 = reinterpret_cast<double2 *>(A);
            double2 *V2 = reinterpret_cast<double2 *>(V);
            const int lda2 = lda / 2, ldv2 = ldv / 2;
            const int swap_a = (G == 8 && (k & 1) && lda % 32 == 0) ? 1 : 0;
            const int swap_v = (G == 8 && (k & 1) && ldv % 32 == 0) ? 1 : 0;
            // Circle positions and held top columns as in the one-role loop;
            // each role advances its own copy after handling a round.
            int pos = k, ct = k, cb = even - 1 - k, fresh = 1, dirty = 0, z = even - 1;
            // The held top column: R row pairs of A in the A waves, the first
            // VR (VR <= R) of V in the V waves (one array, so the two roles
            // share registers).
            double2 held[R];
            double2 unused[1];
            auto load_v = [&](double2 *col) {
                const ColumnShare<G, VR> c(col, lane, swap_v);
#pragma unroll
                for (int j = 0; j < VR; ++j)
                    held[j] = c[j];
            };
            auto store_v = [&](double2 *col) {
                const ColumnShare<G, VR> c(col, lane, swap_v);
#pragma unroll
                for (int j = 0; j < VR; ++j)
                    c[j] = held[j];
            };
            auto advance = [&](int p, int q) {
                const int z1 = z - 1 < 1 ? ring : z - 1;
                const int z2 = z - 2 < 1 ? z - 2 + ring : z - 2;
                const int wrap = (k > 0) & (pos == half - 1);
                if (wrap & (p < n) & dirty)
                {
                    if (role == 0)
                        store_share<G, R>(held, A2 + __umul24(p, lda2), lane, swap_a);
                    else
                        store_v(V2 + __umul24(p, ldv2));
                }
                const int q1 = q - 1 < 1 ? ring : q - 1;
                const int q2 = q - 2 < 1 ? q - 2 + ring : q - 2;
                cb = k == 0 ? q1 : (wrap ? z2 : q2);
                ct = wrap ? z : p;
                pos = wrap ? 1 : (k > 0 ? pos + 1 : 0);
                fresh = wrap;
                z = z1;
            };
            // V waves: apply the rotations logged for round `gr` at slot gr & 1.
            auto v_round = [&](int gr) {
                const int p = ct, q = cb;
                if (fresh & (p < n))
                    load_v(V2 + __umul24(p, ldv2));
                int rot = 0;
                if ((p < n) & (q < n))
                {
                    const double t = T[(gr & 1) * half + pos];
                    if (t != 0.0)
                    {
                        const ColumnShare<G, VR> vq(V2 + __umul24(q, ldv2), lane, swap_v);
                        double2 yq[VR];
#pragma unroll
                        for (int j = 0; j < VR; ++j)
                            yq[j] = vq[j];
                        double c, s;
                        dev::rotation_from_tangent(t, c, s);
#pragma unroll
                        for (int j = 0; j < VR; ++j)
                        {
                            const double2 a = held[j], b = yq[j];
                            yq[j].x = s * a.x + c * b.x;
                            yq[j].y = s * a.y + c * b.y;
                            vq[j] = yq[j];
                            held[j].x = c * a.x - s * b.x;
                            held[j].y = c * a.y - s * b.y;
                        }
                        rot = 1;
                    }
                }
                dirty = (dirty & (fresh ^ 1)) | rot;
                advance(p, q);
            };
            int executed = 0, gr = 0, clean = 0;
            for (int sweep = 0; sweep < args.max_sweeps; ++sweep)
            {
                int any = 0;
                for (int r = 0; r < rounds; ++r, ++gr)
                {
                    if (k < half)
                    {
                        if (role == 0)
                        {
                            const int p = ct, q = cb;
                            if (fresh & (p < n))
                                load_share<G, R>(held, A2 + __umul24(p, lda2), lane, swap_a);
                            int rot = 0;
                            if ((p < n) & (q < n))
                                rot = pair_class_f64<G, R, VR, true>(held, unused, A2 + __umul24(q, lda2), nullptr,
                                                                     lane, swap_a, 0, args.eps,
                                                                     T + (gr & 1) * half + pos);
                            any |= rot;
                            dirty = (dirty & (fresh ^ 1)) | rot;
                            advance(p, q);
                        }
                        else if (gr > 0)
                            v_round(gr - 1);
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
            // The V waves' last round, then every held top column to LDS.
            if (role == 1 && k < half && gr > 0)
                v_round(gr - 1);
            __syncthreads();
            if (k < half && !fresh && dirty && ct < n)
            {
                if (role == 0)
                    store_share<G, R>(held, A2 + __umul24(ct, lda2), lane, swap_a);
                else
                    store_v(V2 + __umul24(ct, ldv2));
            }
            __syncthreads();
            clean_out = clean;
            return executed;
        }

        // Compact layout for the one-role classes that keep V in LDS.  Each
        // pair group holds its top column (A and V) in registers, so LDS only
        // stores the bottom row of the circle: half + 2 column slots of A
        // and of V instead of n.  The bottom column at position P in global
        // round r lives in slot (r + P) mod (half + 2) (it stays in its slot
        // while it moves down the bottom row); the group that leaves
        // position half-1 stores its top column in slot (r - 2) mod
        // (half + 2), freed a round earlier, and at position 1 loads the new
        // top from slot (r - 1) mod (half + 2), bottom[0]'s slot of the
        // previous round.  With one pair (n <= 2) the bottom stays in slot 0.
        // Same pairs in the same order as the reference circle; about half
        // the LDS of the full layout, so more workgroups share a CU.
        template <int BS, int G, int R, int VR, int PASSES>
        __device__ __forceinline__ void solve_compact_f64(const WorkItem &w, double *lds, const ClassArgs &args)
        {
            constexpr int GROUPS = BS / G;
            const int m = w.m, n = w.n, tid = threadIdx.x;
            const int lda = padded_rows(m, G), ldv = padded_rows(n, G);
            const int even = n + (n & 1), half = even >> 1, ring = even - 1;
            const int rounds = n < 2 ? 0 : even - 1;
            const int ns = half + 2;
            double *As = lds;                                     // ns slots of lda rows
            double *Vs = As + ns * lda;                           // ns slots of ldv rows
            double *S = Vs + ns * ldv;                            // sigma scratch (n)
            int *colmap = reinterpret_cast<int *>(S + n);         // column held by each slot
            const int gid = tid / G, lane = tid & (G - 1);
            double2 *A2 = reinterpret_cast<double2 *>(As);
            double2 *V2 = reinterpret_cast<double2 *>(Vs);
            const int lda2 = lda / 2, ldv2 = ldv / 2;
            const int swap_a = (G == 8 && (gid & 1) && lda % 32 == 0) ? 1 : 0;
            const int swap_v = (G == 8 && (gid & 1) && ldv % 32 == 0) ? 1 : 0;

            // Slots 0 .. cnt-1 <- columns col(0 .. cnt-1) of the start matrix
            // and of the initial V (zeros for the dummy column and padding).
            auto fill = [&](int cnt, auto col) {
                for (int idx = tid; idx < cnt * lda; idx += BS)
                {
                    const int sl = idx / lda, i = idx - sl * lda, j = col(sl);
                    double x = 0.0;
                    if (i < m && j < n)
                        x = args.b_cm ? args.b_cm[w.a_off + (int64_t)j * m + i] : args.a[w.a_off + (int64_t)i * n + j];
                    As[idx] = x;
                }
                for (int idx = tid; idx < cnt * ldv; idx += BS)
                {
                    const int sl = idx / ldv, i = idx - sl * ldv, j = col(sl);
                    double x = 0.0;
                    if (i < n && j < n)
                        x = args.q_init ? args.q_init[w.v_off + (int64_t)j * n + i] : (i == j ? 1.0 : 0.0);
                    Vs[idx] = x;
                }
            };

            int pos[PASSES], ct[PASSES], cb[PASSES], fresh[PASSES];
            double2 ta[PASSES][R];
            double2 tv[PASSES][VR];
            // The top row (position k holds column k) is staged through slots
            // 0 .. half-1 into registers, then the bottom row (position P
            // holds column even-1-P) fills slots 0 .. half-1.
            fill(half, [](int sl) { return sl; });
            __syncthreads();
#pragma unroll
            for (int i = 0; i < PASSES; ++i)
            {
                const int k = gid + i * GROUPS;
                pos[i] = k;
                ct[i] = k;
                cb[i] = even - 1 - k;
                fresh[i] = 0;
                if (k < half && k < n)
                {
                    load_share<G, R>(ta[i], A2 + __umul24(k, lda2), lane, swap_a);
                    load_share<G, VR>(tv[i], V2 + __umul24(k, ldv2), lane, swap_v);
                }
            }
            __syncthreads();
            fill(half, [even](int sl) { return even - 1 - sl; });
            __syncthreads();

            int z = even - 1; // bottom[0]'s column this round
            int gs = 0;       // global round mod ns
            double dummy_t;
            int executed = 0, clean = 0;
            for (int sweep = 0; sweep < args.max_sweeps; ++sweep)
            {
                int any = 0;
                for (int r = 0; r < rounds; ++r)
                {
                    const int z1 = z - 1 < 1 ? ring : z - 1;
                    const int z2 = z - 2 < 1 ? z - 2 + ring : z - 2;
                    const int ls = gs == 0 ? ns - 1 : gs - 1;       // fresh top loads
                    const int ws = gs < 2 ? gs + ns - 2 : gs - 2;   // wrap stores
#pragma unroll
                    for (int i = 0; i < PASSES; ++i)
                    {
                        const int k = gid + i * GROUPS;
                        if (k < half)
                        {
                            const int p = ct[i], q = cb[i];
                            if (fresh[i] & (p < n))
                            {
                                load_share<G, R>(ta[i], A2 + __umul24(ls, lda2), lane, swap_a);
                                load_share<G, VR>(tv[i], V2 + __umul24(ls, ldv2), lane, swap_v);
                            }
                            int bsl = gs + pos[i];
                            bsl = half == 1 ? 0 : (bsl >= ns ? bsl - ns : bsl);
                            int rot = 0;
                            if ((p < n) & (q < n))
                                rot = pair_class_f64<G, R, VR, false>(ta[i], tv[i], A2 + __umul24(bsl, lda2),
                                                                      V2 + __umul24(bsl, ldv2), lane, swap_a, swap_v,
                                                                      args.eps, &dummy_t);
                            any |= rot;
                            // Leaving position half-1: the top column becomes
                            // bottom[half-1] in slot ws; the group restarts at
                            // position 1 with the previous bottom[0].
                            const int wrap = (k > 0) & (pos[i] == half - 1);
                            if (wrap & (p < n))
                            {
                                store_share<G, R>(ta[i], A2 + __umul24(ws, lda2), lane, swap_a);
                                store_share<G, VR>(tv[i], V2 + __umul24(ws, ldv2), lane, swap_v);
                            }
                            const int q1 = q - 1 < 1 ? ring : q - 1;
                            const int q2 = q - 2 < 1 ? q - 2 + ring : q - 2;
                            cb[i] = k == 0 ? q1 : (wrap ? z2 : q2);
                            ct[i] = wrap ? z : p;
                            pos[i] = wrap ? 1 : (k > 0 ? pos[i] + 1 : 0);
                            fresh[i] = wrap;
                        }
                    }
                    z = z1;
                    gs = gs + 1 == ns ? 0 : gs + 1;
                    __syncthreads();
                }
                executed = sweep + 1;
                if (!__syncthreads_or(any))
                {
                    clean = 1;
                    break;
                }
            }

            // Epilogue (reference src/domain/jacobi_rotation_kernels.cuh:370-413):
            // sigma_j = ||a_j||; U_ij = a_ij / sigma_j if sigma_j > eps else 0.
            // Pending top loads first, so that every top is in registers.
            const int ls = gs == 0 ? ns - 1 : gs - 1;
#pragma unroll
            for (int i = 0; i < PASSES; ++i)
            {
                const int k = gid + i * GROUPS;
                if (k < half && fresh[i] && ct[i] < n)
                {
                    load_share<G, R>(ta[i], A2 + __umul24(ls, lda2), lane, swap_a);
                    load_share<G, VR>(tv[i], V2 + __umul24(ls, ldv2), lane, swap_v);
                }
            }
            // Column norms: the held top from registers, the bottom from its slot.
#pragma unroll
            for (int i = 0; i < PASSES; ++i)
            {
                const int k = gid + i * GROUPS;
                if (k < half)
                {
                    double acc = 0.0;
#pragma unroll
                    for (int j = 0; j < R; ++j)
                        acc = fma(ta[i][j].y, ta[i][j].y, fma(ta[i][j].x, ta[i][j].x, acc));
                    acc = dev::group_sum_dpp<G>(acc);
                    int bsl = gs + pos[i];
                    bsl = half == 1 ? 0 : (bsl >= ns ? bsl - ns : bsl);
                    double accb = 0.0;
                    for (int row = lane; row < m; row += G)
                    {
                        const double x = As[bsl * lda + row];
                        accb = fma(x, x, accb);
                    }
                    accb = dev::group_sum_dpp<G>(accb);
                    if (lane == 0)
                    {
                        if (ct[i] < n)
                        {
                            const double sj = sqrt(fmax(acc, 0.0));
                            S[ct[i]] = sj;
                            args.sigma[w.s_off + ct[i]] = sj;
                        }
                        if (cb[i] < n)
                        {
                            const double sj = sqrt(fmax(accb, 0.0));
                            S[cb[i]] = sj;
                            args.sigma[w.s_off + cb[i]] = sj;
                        }
                        colmap[bsl] = cb[i];
                    }
                }
            }
            __syncthreads();
            double *ug = args.u + w.a_off;
            // U rule: the reference's absolute eps, or sigma > 0 when the
            // caller asks for every nonzero column normalised.
            const double u_floor = args.u_all_nonzero ? 0.0 : args.eps;
            double *vg = args.v + w.v_off;
            // U and V columns of the slots listed in colmap[0 .. cnt-1]
            // (other slots: colmap >= n), row by row.
            auto write_out = [&](int cnt, const int *cols) {
                for (int idx = tid; idx < cnt * m; idx += BS)
                {
                    const int row = idx / cnt, sl = idx - row * cnt, j = cols[sl];
                    if (j < n)
                    {
                        const double sj = S[j];
                        ug[row * n + j] = (sj > u_floor) ? (As[sl * lda + row] / sj) : 0.0;
                    }
                }
                for (int idx = tid; idx < cnt * n; idx += BS)
                {
                    const int row = idx / cnt, sl = idx - row * cnt, j = cols[sl];
                    if (j < n)
                        vg[row * n + j] = Vs[sl * ldv + row];
                }
            };
            // Slots not holding a bottom column this round hold none.
            if (tid < ns)
            {
                bool bottom = false;
                for (int P = 0; P < half; ++P)
                {
                    int bsl = gs + P;
                    bsl = half == 1 ? 0 : (bsl >= ns ? bsl - ns : bsl);
                    bottom = bottom || bsl == tid;
                }
                if (!bottom)
                    colmap[tid] = n;
            }
            __syncthreads();
            write_out(ns, colmap);
            __syncthreads();
            // Then the held tops: position k's column to slot k.
#pragma unroll
            for (int i = 0; i < PASSES; ++i)
            {
                const int k = gid + i * GROUPS;
                if (k < half)
                {
                    if (ct[i] < n)
                    {
                        store_share<G, R>(ta[i], A2 + __umul24(k, lda2), lane, swap_a);
                        store_share<G, VR>(tv[i], V2 + __umul24(k, ldv2), lane, swap_v);
                    }
                    if (lane == 0)
                        colmap[k] = ct[i];
                }
            }
            __syncthreads();
            write_out(half, colmap);
            if (tid == 0 && args.sweeps)
            {
                if (args.add_sweeps)
                    args.sweeps[w.index] += executed;
                else
                    args.sweeps[w.index] = executed;
            }
            if (tid == 0 && args.stop)
                args.stop[w.index] = clean ? 0 : 1;
        }

        // Inlined into the kernel: as a call, the kernel-argument struct it
        // takes by reference would be copied to private (scratch) memory.
        template <int BS, int G, int R, int VR, bool VLOG, int PASSES, bool SPLIT>
        __device__ __forceinline__ void solve_class_f64(const WorkItem &w, double *lds, const ClassArgs &args)
        {
            if constexpr (!VLOG && !SPLIT)
            {
                solve_compact_f64<BS, G, R, VR, PASSES>(w, lds, args);
                return;
            }
            constexpr int GROUPS = BS / G;
            const int m = w.m, n = w.n, tid = threadIdx.x;
            const int lda = padded_rows(m, G), ldv = padded_rows(n, G);
            double *A = lds;                              // column-major, lda rows (zero padded)
            double *V = VLOG ? lds : lds + lda * n;       // column-major, ldv rows (zero padded)
            double *S = (VLOG ? lds + max(lda, ldv) * n : V + ldv * n); // sigma scratch
            if (args.b_cm)
            {
                const double *bg = args.b_cm + w.a_off; // column-major m x n start matrix
                for (int idx = tid; idx < lda * n; idx += BS)
                {
                    const int j = idx / lda, i = idx - j * lda;
                    A[idx] = i < m ? bg[j * m + i] : 0.0;
                }
            }
            else
            {
                const double *ag = args.a + w.a_off; // row-major m x n input
                for (int idx = tid; idx < lda * n; idx += BS)
                {
                    const int j = idx / lda, i = idx - j * lda;
                    A[idx] = i < m ? ag[i * n + j] : 0.0;
                }
            }
            auto init_v = [&]() {
                for (int idx = tid; idx < ldv * n; idx += BS)
                {
                    const int j = idx / ldv, i = idx - j * ldv;
                    double x = (i == j) ? 1.0 : 0.0;
                    if (args.q_init)
                        x = i < n ? args.q_init[w.v_off + j * n + i] : 0.0;
                    V[idx] = x;
                }
            };
            if constexpr (!VLOG)
                init_v();
            __syncthreads();

            const int even = n + (n & 1);
            const int half = even >> 1;
            const int rounds = n < 2 ? 0 : even - 1;
            const int gid = tid / G, lane = tid & (G - 1);
            double2 *A2 = reinterpret_cast<double2 *>(A);
            double2 *V2 = reinterpret_cast<double2 *>(V);
            const int lda2 = lda / 2, ldv2 = ldv / 2;
            // Block order per lane group (see ColumnShare): swapped in odd
            // 8-lane groups when the column stride is a multiple of 256 bytes.
            const int swap_a = (G == 8 && (gid & 1) && lda % 32 == 0) ? 1 : 0;
            const int swap_v = (G == 8 && (gid & 1) && ldv % 32 == 0) ? 1 : 0;
            double *tlog = VLOG ? args.vlog + (int64_t)blockIdx.x * args.vlog_stride : nullptr;
            double dummy_t;
            // Pair positions of the circle (reference order): top[k] and
            // bottom[k], k = 0 .. half-1; position 0's top is column 0.  Each
            // round every other column moves one slot along the ring
            // top[1] -> top[2] -> ... -> top[half-1] -> bottom[half-1] -> ... ->
            // bottom[0] -> top[1].  Virtual group 0 (lane group 0, pass 0)
            // stays at position 0; virtual group v >= 1 starts at position v
            // and moves one position per round (half-1 -> 1), following its
            // top column, which it holds in registers: only bottom columns
            // travel through LDS.  A group that leaves position half-1 (its
            // top column turns into bottom[half-1]) stores that column and,
            // at position 1, loads the new top (the previous bottom[0]).
            // Column indices >= n are the dummy column of an odd n.
            int executed = 0, clean = 0;
            if constexpr (SPLIT)
            {
                static_assert(!VLOG && PASSES == 1, "split roles carry V in LDS, one pair per group");
                executed = sweeps_split_f64<BS, G, R, VR>(n, lda, ldv, A, V, S + n, args, clean);
            }
            else
            {
                const int ring = even - 1;
                // Per pass: position, top and bottom column, and two flags kept
                // as integers (0/1) rather than lane masks.
                int pos[PASSES], ct[PASSES], cb[PASSES], fresh[PASSES], dirty[PASSES];
                double2 ta[PASSES][R];
                double2 tv[PASSES][VLOG ? 1 : VR];
#pragma unroll
                for (int i = 0; i < PASSES; ++i)
                {
                    const int k = gid + i * GROUPS;
                    pos[i] = k;
                    ct[i] = k;
                    cb[i] = even - 1 - k;
                    fresh[i] = 1; // top column to be loaded
                    dirty[i] = 0;
                }
                int z = even - 1; // bottom[0]'s column this round
                for (int sweep = 0; sweep < args.max_sweeps; ++sweep)
                {
                    int any = 0;
                    for (int r = 0; r < rounds; ++r)
                    {
                        const int z1 = z - 1 < 1 ? ring : z - 1;
                        const int z2 = z - 2 < 1 ? z - 2 + ring : z - 2;
#pragma unroll
                        for (int i = 0; i < PASSES; ++i)
                        {
                            const int k = gid + i * GROUPS;
                            if (k < half)
                            {
                                const int p = ct[i], q = cb[i];
                                if (fresh[i] & (p < n))
                                {
                                    load_share<G, R>(ta[i], A2 + __umul24(p, lda2), lane, swap_a);
                                    if constexpr (!VLOG)
                                        load_share<G, VR>(tv[i], V2 + __umul24(p, ldv2), lane, swap_v);
                                }
                                int rot = 0;
                                if ((p < n) & (q < n))
                                {
                                    double *slot = VLOG ? tlog + ((int64_t)sweep * rounds + r) * half + pos[i] : &dummy_t;
                                    // 24-bit products (full-rate multiplies): column
                                    // indices and leading dimensions are far below 2^24.
                                    rot = pair_class_f64<G, R, VR, VLOG>(ta[i], tv[i], A2 + __umul24(q, lda2),
                                                                         V2 + __umul24(q, ldv2), lane, swap_a, swap_v,
                                                                         args.eps, slot);
                                }
                                any |= rot;
                                dirty[i] = (dirty[i] & (fresh[i] ^ 1)) | rot;
                                // Leaving position half-1: the top column becomes
                                // bottom[half-1] (stored if changed); the group
                                // restarts at position 1 with the previous bottom[0].
                                const int wrap = (k > 0) & (pos[i] == half - 1);
                                if (wrap & (p < n) & dirty[i])
                                {
                                    store_share<G, R>(ta[i], A2 + __umul24(p, lda2), lane, swap_a);
                                    if constexpr (!VLOG)
                                        store_share<G, VR>(tv[i], V2 + __umul24(p, ldv2), lane, swap_v);
                                }
                                const int q1 = q - 1 < 1 ? ring : q - 1;
                                const int q2 = q - 2 < 1 ? q - 2 + ring : q - 2;
                                cb[i] = k == 0 ? q1 : (wrap ? z2 : q2);
                                ct[i] = wrap ? z : p;
                                pos[i] = wrap ? 1 : (k > 0 ? pos[i] + 1 : 0);
                                fresh[i] = wrap;
                            }
                        }
                        z = z1;
                        __syncthreads();
                    }
                    executed = sweep + 1;
                    if (!__syncthreads_or(any))
                    {
                        clean = 1;
                        break;
                    }
                }
                // Held top columns back to LDS (those loaded and changed since).
#pragma unroll
                for (int i = 0; i < PASSES; ++i)
                {
                    const int k = gid + i * GROUPS;
                    if (k < half && !fresh[i] && dirty[i] && ct[i] < n)
                    {
                        store_share<G, R>(ta[i], A2 + __umul24(ct[i], lda2), lane, swap_a);
                        if constexpr (!VLOG)
                            store_share<G, VR>(tv[i], V2 + __umul24(ct[i], ldv2), lane, swap_v);
                    }
                }
                __syncthreads();
            }
            dev::CirclePair circ[PASSES]; // pair positions for the rotation-log replay

            // U / Sigma epilogue (reference src/domain/jacobi_rotation_kernels.cuh:370-413):
            // sigma_j = ||a_j||; U_ij = a_ij / sigma_j if sigma_j > eps else 0.
            for (int j = gid; j < n; j += GROUPS)
            {
                double acc = 0.0;
                for (int i = lane; i < m; i += G)
                {
                    const double x = A[j * lda + i];
                    acc += x * x;
                }
                acc = dev::group_sum_dpp<G>(acc);
                if (lane == 0)
                {
                    const double sj = sqrt(fmax(acc, 0.0));
                    S[j] = sj;
                    args.sigma[w.s_off + j] = sj;
                }
            }
            __syncthreads();
            double *ug = args.u + w.a_off;
            // U rule: the reference's absolute eps, or sigma > 0 when the
            // caller asks for every nonzero column normalised.
            const double u_floor = args.u_all_nonzero ? 0.0 : args.eps;
            const int mn = m * n, nn = n * n;
            for (int idx = tid; idx < mn; idx += BS)
            {
                const int i = idx / n, j = idx - i * n;
                const double sj = S[j];
                ug[idx] = (sj > u_floor) ? (A[j * lda + i] / sj) : 0.0;
            }
            if constexpr (VLOG)
            {
                // A is no longer needed: rebuild V in its place from the log.
                __syncthreads();
                init_v();
                __syncthreads();
                // The tangents are read G rounds at a time, spread across the
                // group's lanes (lane l holds round j*G + l of chunk j) and
                // broadcast per round; the next chunk is fetched while the
                // current one is replayed.
                constexpr int RT = (2 * PASSES * GROUPS - 1 + G - 1) / G; // chunks per sweep (upper bound)
                const int group_base = threadIdx.x & ~(G - 1);
                const int chunks = (rounds + G - 1) / G;
                auto fetch = [&](int sweep, int j, int i) -> double {
                    const int k = gid + i * GROUPS;
                    const int r = lane + j * G;
                    // Agent-scope loads bypass the CU's vector L1, which may
                    // still hold this slot's entries from the previous matrix.
                    return (k < half && r < rounds)
                               ? __longlong_as_double(static_cast<long long>(__hip_atomic_load(
                                     reinterpret_cast<unsigned long long *>(tlog + ((int64_t)sweep * rounds + r) * half + k),
                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
                               : 0.0;
                };
                static_assert(RT >= 1, "at least one chunk");
                for (int sweep = 0; sweep < executed; ++sweep)
                {
                    double cur[PASSES], nxt[PASSES];
#pragma unroll
                    for (int i = 0; i < PASSES; ++i)
                    {
                        circ[i].start(gid + i * GROUPS, even);
                        cur[i] = fetch(sweep, 0, i);
                    }
                    for (int j = 0; j < chunks; ++j)
                    {
                        if (j + 1 < chunks)
                        {
#pragma unroll
                            for (int i = 0; i < PASSES; ++i)
                                nxt[i] = fetch(sweep, j + 1, i);
                        }
                        for (int rr = 0; rr < G; ++rr)
                        {
                            const int r = j * G + rr;
                            if (r >= rounds)
                                break;
#pragma unroll
                            for (int i = 0; i < PASSES; ++i)
                            {
                                const int k = gid + i * GROUPS;
                                const double t = __shfl(cur[i], group_base + rr, 64);
                                if (k < half)
                                {
                                    const int p = circ[i].p, q = circ[i].q;
                                    if (p < n && q < n)
                                        replay_pair_f64<G, VR>(V2 + __umul24(p, ldv2), V2 + __umul24(q, ldv2),
                                                                              lane, swap_v, t);
                                    circ[i].advance(even);
                                }
                            }
                            __syncthreads();
                        }
#pragma unroll
                        for (int i = 0; i < PASSES; ++i)
                            cur[i] = nxt[i];
                    }
                }
            }
            double *vg = args.v + w.v_off;
            for (int idx = tid; idx < nn; idx += BS)
            {
                const int i = idx / n, j = idx - i * n;
                vg[idx] = V[j * ldv + i];
            }
            if (tid == 0 && args.sweeps)
            {
                if (args.add_sweeps)
                    args.sweeps[w.index] += executed;
                else
                    args.sweeps[w.index] = executed;
            }
            if (tid == 0 && args.stop)
                args.stop[w.index] = clean ? 0 : 1;
        }

        // Runs solve_class_f64 instantiated for the matrix's exact numbers of
        // row pairs per lane: R = 1 .. RMAX2 for A, VR = 1 .. min(R, VMAX2)
        // for V (n <= m, so V never has more row pairs than A).
        template <int BS, int G, int VMAX2, bool VLOG, int PASSES, bool SPLIT, int R, int VR>
        __device__ __forceinline__ void solve_vrows(int vrows2, const WorkItem &w, double *lds, const ClassArgs &args)
        {
            if constexpr (VR < VMAX2 && VR < R)
            {
                if (vrows2 != VR)
                {
                    solve_vrows<BS, G, VMAX2, VLOG, PASSES, SPLIT, R, VR + 1>(vrows2, w, lds, args);
                    return;
                }
            }
            solve_class_f64<BS, G, R, VR, VLOG, PASSES, SPLIT>(w, lds, args);
        }

        template <int BS, int G, int RMAX2, int VMAX2, bool VLOG, int PASSES, bool SPLIT, int R>
        __device__ __forceinline__ void solve_rows(int rows2, int vrows2, const WorkItem &w, double *lds,
                                                   const ClassArgs &args)
        {
            if constexpr (R < RMAX2)
            {
                if (rows2 != R)
                {
                    solve_rows<BS, G, RMAX2, VMAX2, VLOG, PASSES, SPLIT, R + 1>(rows2, vrows2, w, lds, args);
                    return;
                }
            }
            solve_vrows<BS, G, VMAX2, VLOG, PASSES, SPLIT, R, 1>(vrows2, w, lds, args);
        }

        // counters[0]: next work-list entry; counters[1]: workgroups finished.
        // The last workgroup to finish resets both, so the next launch on the
        // stream starts from a clean work-list without a memset.
        template <int BS, int G, int RMAX2, int VMAX2, bool VLOG, int PASSES, bool SPLIT>
        __global__ __launch_bounds__(BS) void lds_class_f64(ClassArgs args)
        {
            // Declared as double2 so the dynamic LDS block is 16-byte aligned
            // (the row-pair accesses are 16-byte ds_read_b128 / ds_write_b128).
            extern __shared__ double2 lds_pairs[];
            double *lds = reinterpret_cast<double *>(lds_pairs);
            __shared__ int s_item;
            // Raised issue priority over co-resident workgroups of other
            // buckets: set for the buckets whose matrices need a whole CU's
            // LDS, the longest single solves of a plan.
            if (args.priority)
                __builtin_amdgcn_s_setprio(3);
            for (;;)
            {
                if (threadIdx.x == 0)
                    s_item = (int)atomicAdd(&args.counters[0], 1u);
                __syncthreads();
                const int item = s_item;
                if (item >= args.count)
                    break;
                const WorkItem w = args.items[item];
                solve_rows<BS, G, RMAX2, VMAX2, VLOG, PASSES, SPLIT, 1>(padded_rows(w.m, G) / (2 * G),
                                                                 padded_rows(w.n, G) / (2 * G), w, lds, args);
                __syncthreads();
            }
            if (threadIdx.x == 0)
            {
                const unsigned int done = atomicAdd(&args.counters[1], 1u);
                if (done == gridDim.x - 1)
                {
                    __hip_atomic_store(&args.counters[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&args.counters[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }

        template <int I>
        const void *class_symbol_t()
        {
            constexpr ClassSpec c = kClasses[I];
            return reinterpret_cast<const void *>(
                &lds_class_f64<c.block, c.group, c.rmax2, c.vmax2, c.vlog, c.passes, c.split>);
        }

        template <int I>
        void class_launch_t(int grid, int lds_bytes, const ClassArgs &args, hipStream_t stream)
        {
            constexpr ClassSpec c = kClasses[I];
            lds_class_f64<c.block, c.group, c.rmax2, c.vmax2, c.vlog, c.passes, c.split>
                <<<grid, c.block, lds_bytes, stream>>>(args);
        }

        using LaunchFn = void (*)(int, int, const ClassArgs &, hipStream_t);
        const LaunchFn kClassLaunch[] = {class_launch_t<0>, class_launch_t<1>, class_launch_t<2>,
                                         class_launch_t<3>, class_launch_t<4>, class_launch_t<5>,
                                         class_launch_t<6>, class_launch_t<7>, class_launch_t<8>,
                                         class_launch_t<9>};
        const void *(*const kClassSymbol[])() = {class_symbol_t<0>, class_symbol_t<1>, class_symbol_t<2>,
                                                 class_symbol_t<3>, class_symbol_t<4>, class_symbol_t<5>,
                                                 class_symbol_t<6>, class_symbol_t<7>, class_symbol_t<8>,
                                                 class_symbol_t<9>};
        static_assert(sizeof(kClassLaunch) / sizeof(kClassLaunch[0]) == kClassCount);
        static_assert(sizeof(kClassSymbol) / sizeof(kClassSymbol[0]) == kClassCount);
    } // namespace

    int class_count()
    {
        return kClassCount;
    }

    bool class_uses_vlog(int cls)
    {
        return cls >= 0 && cls < kClassCount && kClasses[cls].vlog;
    }

    int64_t class_lds_bytes(int cls, int64_t m, int64_t n)
    {
        const int g = kClasses[cls].group;
        const int64_t lda = ((m + 2 * g - 1) / (2 * g)) * (2 * g);
        const int64_t ldv = ((n + 2 * g - 1) / (2 * g)) * (2 * g);
        if (kClasses[cls].vlog)
            return (std::max(lda, ldv) * n + n) * 8;
        if (kClasses[cls].split)
        {
            // A and V in full, sigma, a two-round tangent ring.
            return (lda * n + ldv * n + n + 2 * ((n + 1) / 2)) * 8;
        }
        // Compact layout: (n+1)/2 + 2 slots of A and of V, sigma, and one
        // int per slot.
        const int64_t ns = (n + 1) / 2 + 2;
        return (ns * (lda + ldv) + n + (ns + 1) / 2) * 8;
    }

    bool class_fits(int cls, int64_t m, int64_t n)
    {
        if (cls < 0 || cls >= kClassCount)
            return false;
        const ClassSpec &c = kClasses[cls];
        const int64_t half = (n + 1) / 2;
        const int64_t groups = (c.split ? c.block / 2 : c.block) / c.group;
        // One wave: one pair per group per round.  Multi-wave classes may
        // loop over pairs (up to two passes, four with the rotation log).
        const int64_t passes = c.passes;
        if (half > passes * groups)
            return false;
        const int64_t lda = ((m + 2 * c.group - 1) / (2 * c.group)) * (2 * c.group);
        const int64_t ldv = ((n + 2 * c.group - 1) / (2 * c.group)) * (2 * c.group);
        if (lda / (2 * c.group) > c.rmax2 || ldv / (2 * c.group) > c.vmax2)
            return false;
        return class_lds_bytes(cls, m, n) <= kMaxLdsBytes;
    }

    int select_class_f64(int64_t m, int64_t n)
    {
        // Preference order: the four-wave class 4 before the one-wave class 2
        // and the eight-wave class 3.  Against class 3 (same shapes) half the
        // waves repeat the rotation arithmetic and the group reductions are
        // one DPP level shorter; against class 2 (n <= 32), the 8-lane
        // groups keep the LDS accesses free of bank conflicts and four waves
        // share one workgroup's barriers (7% faster on the uneven
        // 4096-matrix batch than class 2 first).  The split-role class 9 is
        // 9% faster per round than class 4 for a 128 x 64 matrix alone on its
        // CU, but its eight waves displace other buckets' workgroups: on the
        // uneven batch it loses 2% as the default for whole-CU matrices, so
        // it comes last (the autotuner may still choose it per bucket).
        static constexpr int kOrder[] = {0, 1, 4, 2, 3, 5, 6, 7, 8, 9};
        static_assert(sizeof(kOrder) / sizeof(kOrder[0]) == kClassCount, "every class once");
        for (int i : kOrder)
            if (class_fits(i, m, n))
                return i;
        return -1;
    }

    int class_block(int cls)
    {
        return kClasses[cls].block;
    }

    const void *lds_class_f64_symbol(int cls)
    {
        return (cls >= 0 && cls < kClassCount) ? kClassSymbol[cls]() : nullptr;
    }

    hipError_t launch_lds_class_f64(int cls, int grid, int lds_bytes, const ClassArgs &args, hipStream_t stream)
    {
        if (cls < 0 || cls >= kClassCount)
            return hipErrorInvalidValue;
        kClassLaunch[cls](grid, lds_bytes, args, stream);
        return hipGetLastError();
    }
} // namespace jsvd
