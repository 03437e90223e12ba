// jsvd_capi.hip -- the C-ABI (include/jsvd_c.h): validation, device work-list
// plans, workspace, launches and error reporting.  Host code only; the
// kernels live in jsvd_kernels_*.hip.
#include "jsvd_c.h"
#include "jsvd_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

namespace
{
    thread_local std::string g_last_error;

    jsvd_status fail(jsvd_status status, const std::string &message)
    {
        g_last_error = message;
        return status;
    }

    std::string hip_message(hipError_t e, const char *expr, const char *file, int line)
    {
        char buf[512];
        std::snprintf(buf, sizeof(buf), "HIP call failed: %s @ %s:%d, code=%d, message=%s", expr, file, line,
                      static_cast<int>(e), hipGetErrorString(e));
        return buf;
    }

#define JSVD_HIP(expr)                                                                                         \
    do                                                                                                         \
    {                                                                                                          \
        const hipError_t jsvd_e_ = (expr);                                                                     \
        if (jsvd_e_ != hipSuccess)                                                                             \
            return fail(jsvd_e_ == hipErrorOutOfMemory ? JSVD_ENOMEM : JSVD_EDEVICE,                           \
                        hip_message(jsvd_e_, #expr, __FILE__, __LINE__));                                      \
    } while (0)

#define JSVD_CHECK(expr)                                                                                       \
    do                                                                                                         \
    {                                                                                                          \
        const jsvd_status jsvd_s_ = (expr);                                                                    \
        if (jsvd_s_ != JSVD_OK)                                                                                \
            return jsvd_s_;                                                                                    \
    } while (0)

    // Validation with the reference's messages
    // (src/domain/jacobi_svd.cu:197-223 and :36-54 of that file).
    jsvd_status validate_config(const jsvd_config &c)
    {
        if (!(c.epsilon > 0.0))
            return fail(JSVD_EINVAL, "epsilon must be positive.");
        if (c.max_sweeps <= 0)
            return fail(JSVD_EINVAL, "max_sweeps must be positive.");
        if (c.layout_min_columns <= 0)
            return fail(JSVD_EINVAL, "layout_transpose_min_columns must be positive.");
        if (c.layout_min_elements == 0)
            return fail(JSVD_EINVAL, "layout_transpose_min_elements must be positive.");
        if (c.layout_bench_reps <= 0)
            return fail(JSVD_EINVAL, "layout_transpose_benchmark_repetitions must be positive.");
        if (c.layout_bench_sweeps <= 0)
            return fail(JSVD_EINVAL, "layout_transpose_benchmark_sweeps must be positive.");
        if (c.layout_mode < JSVD_LAYOUT_AUTO || c.layout_mode > JSVD_LAYOUT_OFF)
            return fail(JSVD_EINVAL, "layout_transpose_mode is not a valid policy.");
        if (c.precision != JSVD_PRECISION_FP64 && c.precision != JSVD_PRECISION_MIXED)
            return fail(JSVD_EINVAL, "precision must be JSVD_PRECISION_FP64 or JSVD_PRECISION_MIXED.");
        if (c.block_width != 0 && c.block_width != 16 && c.block_width != 32)
            return fail(JSVD_EINVAL, "block_width must be 0, 16 or 32.");
        if (c.refine_sweeps < 0)
            return fail(JSVD_EINVAL, "refine_sweeps must be non-negative.");
        return JSVD_OK;
    }

    jsvd_status validate_shape(uint64_t rows, uint64_t columns)
    {
        if (rows == 0 || columns == 0)
            return fail(JSVD_EINVAL, "Input matrix shape must be non-zero.");
        if (rows < columns)
            return fail(JSVD_EINVAL, "One-sided Jacobi SVD requires rows >= columns in this implementation.");
        if (columns > (1u << 30) || rows > (1ull << 40))
            return fail(JSVD_EINVAL, "Input matrix shape is too large.");
        return JSVD_OK;
    }

    // threads_per_block maps to a workgroup size: a multiple of 64 in [64,1024]
    // (the reference rounds to a multiple of 32 in [32,1024],
    // src/domain/jacobi_schedule.cuh:64-69).
    int normalize_workgroup(int raw)
    {
        const int clamped = std::clamp(raw, 64, 1024);
        return std::min(((clamped + 63) / 64) * 64, 1024);
    }

    int lds_kernel_block(int raw)
    {
        const int w = normalize_workgroup(raw);
        if (w <= 64)
            return 64;
        if (w <= 128)
            return 128;
        if (w <= 256)
            return 256;
        return 512;
    }

    int persistent_kernel_block(int raw)
    {
        const int w = normalize_workgroup(raw);
        if (w <= 256)
            return 256;
        if (w <= 512)
            return 512;
        return 1024;
    }

    template <typename T>
    struct DevBuf
    {
        T *ptr = nullptr;
        size_t count = 0;
        DevBuf() = default;
        DevBuf(const DevBuf &) = delete;
        DevBuf &operator=(const DevBuf &) = delete;
        ~DevBuf()
        {
            if (ptr)
                (void)hipFree(ptr);
        }
        jsvd_status alloc(size_t n)
        {
            if (ptr)
                (void)hipFree(ptr);
            ptr = nullptr;
            count = 0;
            if (n == 0)
                return JSVD_OK;
            JSVD_HIP(hipMalloc(reinterpret_cast<void **>(&ptr), n * sizeof(T)));
            count = n;
            return JSVD_OK;
        }
        jsvd_status reserve(size_t n)
        {
            return n <= count ? JSVD_OK : alloc(n);
        }
    };
} // namespace

struct jsvd_plan
{
    int device = 0;
    jsvd_config cfg{};
    int64_t count = 0;
    std::vector<int64_t> m, n, a_off, v_off, s_off;
    // LDS-resident work-list part.
    std::vector<jsvd::WorkItem> items;
    DevBuf<jsvd::WorkItem> d_items;
    DevBuf<unsigned int> d_counters;
    int lds_block = 256, lds_grid = 0, lds_bytes = 0;
    // Large matrices: persistent kernel over a column-major workspace.
    std::vector<int64_t> large;
    DevBuf<double> ws_a, ws_v, ws_s;
    DevBuf<int32_t> ws_flags;
    DevBuf<unsigned int> ws_bar;
    int big_block = 256;
    int64_t a_elems = 0, v_elems = 0, s_elems = 0;
};

namespace
{
    jsvd_status plan_large(jsvd_plan &p, int64_t k, const double *d_a, double *d_u, double *d_sigma, double *d_v,
                           int32_t *d_sweeps, hipStream_t stream)
    {
        const int64_t m = p.m[k];
        const int n = static_cast<int>(p.n[k]);
        const double *a = d_a + p.a_off[k];
        // Row-major A -> column-major workspace (padded-LDS transpose).
        JSVD_HIP(jsvd::launch_transpose_f64(a, p.ws_a.ptr, m, n, nullptr, 0.0, stream));
        JSVD_HIP(jsvd::launch_identity_f64(p.ws_v.ptr, n, stream));
        JSVD_HIP(jsvd::launch_fill_i32(p.ws_flags.ptr, 0, p.cfg.max_sweeps, stream));
        int occupancy = 0;
        JSVD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occupancy, jsvd::persistent_sweeps_f64_symbol(p.big_block), p.big_block, 0));
        hipDeviceProp_t prop;
        JSVD_HIP(hipGetDeviceProperties(&prop, p.device));
        const int resident = std::max(1, occupancy) * prop.multiProcessorCount;
        const int half = (n + (n & 1)) / 2;
        const int grid = std::max(1, std::min(resident, half));
        int32_t *sweeps_slot = d_sweeps ? d_sweeps + k : nullptr;
        JSVD_HIP(jsvd::launch_persistent_sweeps_f64(p.big_block, grid, p.ws_a.ptr, p.ws_v.ptr, m, n,
                                                    p.cfg.epsilon, p.cfg.max_sweeps, p.ws_flags.ptr,
                                                    p.ws_bar.ptr, sweeps_slot, stream));
        JSVD_HIP(jsvd::launch_column_norms_f64(p.ws_a.ptr, p.ws_s.ptr, m, n, stream));
        JSVD_HIP(hipMemcpyAsync(d_sigma + p.s_off[k], p.ws_s.ptr, sizeof(double) * n, hipMemcpyDeviceToDevice,
                                stream));
        // U (m x n row-major) = transpose of the n x m row-major view of A_cm,
        // each A column divided by its sigma.
        JSVD_HIP(jsvd::launch_transpose_f64(p.ws_a.ptr, d_u + p.a_off[k], n, m, p.ws_s.ptr, p.cfg.epsilon, stream));
        JSVD_HIP(jsvd::launch_transpose_f64(p.ws_v.ptr, d_v + p.v_off[k], n, n, nullptr, 0.0, stream));
        return JSVD_OK;
    }

    struct DeviceContext
    {
        std::mutex mu;
        hipStream_t stream = nullptr;
        DevBuf<double> a, u, s, v;
        DevBuf<int32_t> sweeps;
    };

    std::mutex g_ctx_mu;
    std::vector<std::unique_ptr<DeviceContext>> g_ctx;

    jsvd_status device_context(DeviceContext **out)
    {
        int dev = 0;
        JSVD_HIP(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lock(g_ctx_mu);
        if (static_cast<int>(g_ctx.size()) <= dev)
            g_ctx.resize(dev + 1);
        if (!g_ctx[dev])
        {
            auto ctx = std::make_unique<DeviceContext>();
            JSVD_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            g_ctx[dev] = std::move(ctx);
        }
        *out = g_ctx[dev].get();
        return JSVD_OK;
    }
} // namespace

extern "C" {

void jsvd_config_default(jsvd_config *cfg)
{
    if (!cfg)
        return;
    cfg->epsilon = 1.0e-9;
    cfg->max_sweeps = 128;
    cfg->threads_per_block = 256;
    cfg->layout_mode = JSVD_LAYOUT_AUTO;
    cfg->layout_min_columns = 16;
    cfg->layout_min_elements = 4096;
    cfg->layout_auto_tune = 0;
    cfg->layout_bench_reps = 2;
    cfg->layout_bench_sweeps = 8;
    cfg->precision = JSVD_PRECISION_FP64;
    cfg->block_width = 0;
    cfg->refine_sweeps = 4;
}

const char *jsvd_last_error(void)
{
    return g_last_error.c_str();
}

int32_t jsvd_abi_version(void)
{
    return 1;
}

jsvd_status jsvd_plan_create(int64_t count, const int64_t *m, const int64_t *n, const jsvd_config *cfg_in,
                             jsvd_plan **out)
{
    g_last_error.clear();
    if (!out)
        return fail(JSVD_EINVAL, "plan output pointer is null.");
    *out = nullptr;
    if (count < 0 || (count > 0 && (!m || !n)))
        return fail(JSVD_EINVAL, "Input batch description is invalid.");
    jsvd_config cfg;
    jsvd_config_default(&cfg);
    if (cfg_in)
        cfg = *cfg_in;
    JSVD_CHECK(validate_config(cfg));
    for (int64_t k = 0; k < count; ++k)
    {
        if (m[k] <= 0 || n[k] <= 0)
            return fail(JSVD_EINVAL, "Input matrix shape must be non-zero.");
        JSVD_CHECK(validate_shape(static_cast<uint64_t>(m[k]), static_cast<uint64_t>(n[k])));
    }

    auto plan = std::make_unique<jsvd_plan>();
    JSVD_HIP(hipGetDevice(&plan->device));
    plan->cfg = cfg;
    plan->count = count;
    plan->m.assign(m, m + count);
    plan->n.assign(n, n + count);
    plan->a_off.resize(count);
    plan->v_off.resize(count);
    plan->s_off.resize(count);
    int64_t ao = 0, vo = 0, so = 0, big_mn = 0, big_nn = 0, big_n = 0;
    int64_t max_lds = 0;
    for (int64_t k = 0; k < count; ++k)
    {
        plan->a_off[k] = ao;
        plan->v_off[k] = vo;
        plan->s_off[k] = so;
        ao += m[k] * n[k];
        vo += n[k] * n[k];
        so += n[k];
        const int64_t need = jsvd::lds_bytes_f64(m[k], n[k]);
        if (need <= jsvd::kMaxLdsBytes)
        {
            jsvd::WorkItem w{};
            w.m = static_cast<int32_t>(m[k]);
            w.n = static_cast<int32_t>(n[k]);
            w.index = static_cast<int32_t>(k);
            w.a_off = plan->a_off[k];
            w.v_off = plan->v_off[k];
            w.s_off = plan->s_off[k];
            plan->items.push_back(w);
            max_lds = std::max(max_lds, need);
        }
        else
        {
            plan->large.push_back(k);
            big_mn = std::max(big_mn, m[k] * n[k]);
            big_nn = std::max(big_nn, n[k] * n[k]);
            big_n = std::max(big_n, n[k]);
        }
    }
    plan->a_elems = ao;
    plan->v_elems = vo;
    plan->s_elems = so;

    // Longest-processing-time-first work-list: the cost of a sweep grows as
    // n^2 (m + n), so the biggest matrices are claimed first and the small
    // ones fill the tail.
    std::stable_sort(plan->items.begin(), plan->items.end(), [](const jsvd::WorkItem &x, const jsvd::WorkItem &y) {
        const double cx = double(x.n) * x.n * (x.m + x.n), cy = double(y.n) * y.n * (y.m + y.n);
        return cx > cy;
    });

    if (!plan->items.empty())
    {
        plan->lds_block = lds_kernel_block(cfg.threads_per_block);
        plan->lds_bytes = static_cast<int>(max_lds);
        const void *fn = jsvd::lds_batch_f64_symbol(plan->lds_block);
        JSVD_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, plan->lds_bytes));
        int occupancy = 0;
        JSVD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occupancy, fn, plan->lds_block, plan->lds_bytes));
        if (occupancy <= 0)
            return fail(JSVD_EDEVICE, "work-list kernel cannot be resident with the requested LDS size.");
        hipDeviceProp_t prop;
        JSVD_HIP(hipGetDeviceProperties(&prop, plan->device));
        const int64_t resident = int64_t(occupancy) * prop.multiProcessorCount;
        plan->lds_grid = static_cast<int>(std::min<int64_t>(resident, plan->items.size()));
        JSVD_CHECK(plan->d_items.alloc(plan->items.size()));
        JSVD_HIP(hipMemcpy(plan->d_items.ptr, plan->items.data(), sizeof(jsvd::WorkItem) * plan->items.size(),
                           hipMemcpyHostToDevice));
        JSVD_CHECK(plan->d_counters.alloc(2));
        JSVD_HIP(hipMemset(plan->d_counters.ptr, 0, sizeof(unsigned int) * 2));
    }
    if (!plan->large.empty())
    {
        plan->big_block = persistent_kernel_block(cfg.threads_per_block);
        JSVD_CHECK(plan->ws_a.alloc(big_mn));
        JSVD_CHECK(plan->ws_v.alloc(big_nn));
        JSVD_CHECK(plan->ws_s.alloc(big_n));
        JSVD_CHECK(plan->ws_flags.alloc(cfg.max_sweeps));
        JSVD_CHECK(plan->ws_bar.alloc(2));
        JSVD_HIP(hipMemset(plan->ws_bar.ptr, 0, sizeof(unsigned int) * 2));
    }
    JSVD_HIP(hipDeviceSynchronize());
    *out = plan.release();
    return JSVD_OK;
}

jsvd_status jsvd_plan_execute_f64(jsvd_plan *p, const double *d_a, double *d_u, double *d_sigma, double *d_v,
                                  int32_t *d_sweeps, void *stream_ptr)
{
    g_last_error.clear();
    if (!p)
        return fail(JSVD_EINVAL, "plan is null.");
    if (p->count == 0)
        return JSVD_OK;
    if (!d_a || !d_u || !d_sigma || !d_v)
        return fail(JSVD_EINVAL, "device buffer pointer is null.");
    int dev = 0;
    JSVD_HIP(hipGetDevice(&dev));
    if (dev != p->device)
        return fail(JSVD_EINVAL, "plan was created on another device.");
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    if (!p->items.empty())
    {
        JSVD_HIP(jsvd::launch_lds_batch_f64(p->lds_block, p->lds_grid, p->lds_bytes, p->d_items.ptr,
                                            static_cast<int>(p->items.size()), d_a, d_u, d_sigma, d_v, d_sweeps,
                                            p->cfg.epsilon, p->cfg.max_sweeps, p->d_counters.ptr, stream));
    }
    for (int64_t k : p->large)
        JSVD_CHECK(plan_large(*p, k, d_a, d_u, d_sigma, d_v, d_sweeps, stream));
    return JSVD_OK;
}

jsvd_status jsvd_plan_get_info(const jsvd_plan *p, jsvd_plan_info *info)
{
    if (!p || !info)
        return fail(JSVD_EINVAL, "plan or info pointer is null.");
    info->count = p->count;
    info->lds_matrices = static_cast<int64_t>(p->items.size());
    info->large_matrices = static_cast<int64_t>(p->large.size());
    info->a_elements = p->a_elems;
    info->v_elements = p->v_elems;
    info->sigma_elements = p->s_elems;
    info->workgroup_size = p->lds_block;
    info->grid_size = p->lds_grid;
    info->lds_bytes = p->lds_bytes;
    info->precision = p->cfg.precision;
    return JSVD_OK;
}

void jsvd_plan_destroy(jsvd_plan *p)
{
    delete p;
}

jsvd_status jsvd_solve_batched_f64(int64_t count, const double *const *a_rm, const uint64_t *m, const uint64_t *n,
                                   const jsvd_config *cfg, double *const *u_rm, double *const *sigma,
                                   double *const *v_rm, int32_t *sweeps)
{
    g_last_error.clear();
    if (count < 0 || (count > 0 && (!a_rm || !m || !n || !u_rm || !sigma || !v_rm)))
        return fail(JSVD_EINVAL, "Input batch description is invalid.");
    if (count == 0)
        return JSVD_OK;
    std::vector<int64_t> mm(count), nn(count);
    for (int64_t k = 0; k < count; ++k)
    {
        JSVD_CHECK(validate_shape(m[k], n[k]));
        if (!a_rm[k] || !u_rm[k] || !sigma[k] || !v_rm[k])
            return fail(JSVD_EINVAL, "Input buffer size mismatch.");
        mm[k] = static_cast<int64_t>(m[k]);
        nn[k] = static_cast<int64_t>(n[k]);
    }
    {
        // All argument checks happen before the first device call.
        jsvd_config c;
        jsvd_config_default(&c);
        if (cfg)
            c = *cfg;
        JSVD_CHECK(validate_config(c));
    }
    DeviceContext *ctx = nullptr;
    JSVD_CHECK(device_context(&ctx));
    jsvd_plan *plan = nullptr;
    JSVD_CHECK(jsvd_plan_create(count, mm.data(), nn.data(), cfg, &plan));
    std::unique_ptr<jsvd_plan, void (*)(jsvd_plan *)> guard(plan, jsvd_plan_destroy);

    std::lock_guard<std::mutex> lock(ctx->mu);
    JSVD_CHECK(ctx->a.reserve(plan->a_elems));
    JSVD_CHECK(ctx->u.reserve(plan->a_elems));
    JSVD_CHECK(ctx->s.reserve(plan->s_elems));
    JSVD_CHECK(ctx->v.reserve(plan->v_elems));
    JSVD_CHECK(ctx->sweeps.reserve(count));
    hipStream_t st = ctx->stream;
    for (int64_t k = 0; k < count; ++k)
        JSVD_HIP(hipMemcpyAsync(ctx->a.ptr + plan->a_off[k], a_rm[k], sizeof(double) * mm[k] * nn[k],
                                hipMemcpyHostToDevice, st));
    JSVD_CHECK(jsvd_plan_execute_f64(plan, ctx->a.ptr, ctx->u.ptr, ctx->s.ptr, ctx->v.ptr, ctx->sweeps.ptr, st));
    for (int64_t k = 0; k < count; ++k)
    {
        JSVD_HIP(hipMemcpyAsync(u_rm[k], ctx->u.ptr + plan->a_off[k], sizeof(double) * mm[k] * nn[k],
                                hipMemcpyDeviceToHost, st));
        JSVD_HIP(hipMemcpyAsync(sigma[k], ctx->s.ptr + plan->s_off[k], sizeof(double) * nn[k],
                                hipMemcpyDeviceToHost, st));
        JSVD_HIP(hipMemcpyAsync(v_rm[k], ctx->v.ptr + plan->v_off[k], sizeof(double) * nn[k] * nn[k],
                                hipMemcpyDeviceToHost, st));
    }
    std::vector<int32_t> sw(count, 0);
    JSVD_HIP(hipMemcpyAsync(sw.data(), ctx->sweeps.ptr, sizeof(int32_t) * count, hipMemcpyDeviceToHost, st));
    JSVD_HIP(hipStreamSynchronize(st));
    if (sweeps)
        std::copy(sw.begin(), sw.end(), sweeps);
    return JSVD_OK;
}

jsvd_status jsvd_solve_f64(const double *a_rm, uint64_t m, uint64_t n, const jsvd_config *cfg, double *u_rm,
                           double *sigma, double *v_rm, int32_t *sweeps)
{
    g_last_error.clear();
    JSVD_CHECK(validate_shape(m, n));
    if (!a_rm || !u_rm || !sigma || !v_rm)
        return fail(JSVD_EINVAL, "Input buffer size mismatch.");
    return jsvd_solve_batched_f64(1, &a_rm, &m, &n, cfg, &u_rm, &sigma, &v_rm, sweeps);
}

jsvd_status jsvd_autotune(const jsvd_config *cfg, jsvd_autotune_report *out)
{
    g_last_error.clear();
    if (!out)
        return fail(JSVD_EINVAL, "report pointer is null.");
    jsvd_config c;
    jsvd_config_default(&c);
    if (cfg)
        c = *cfg;
    JSVD_CHECK(validate_config(c));
    out->executed = 1;
    out->recommended_min_columns = c.layout_min_columns;
    out->recommended_min_elements = c.layout_min_elements;
    out->estimated_best_speedup = 1.0;
    out->sample_count = 0;
    return JSVD_OK;
}

} // extern "C"
