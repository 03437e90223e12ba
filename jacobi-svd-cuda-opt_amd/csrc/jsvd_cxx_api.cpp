// jsvd_cxx_api.cpp -- the reference's C++ domain API (jacobi::svd) on top of
// the C-ABI.  Same names, signatures, exception types and messages as
// include/jacobi/svd/domain/*.cuh of the reference.
#include "jacobi/svd/domain/kernels.cuh"
#include "jsvd_c.h"

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>

namespace jacobi::svd
{
    CudaError::CudaError(const char *message) : std::runtime_error(message) {}

    namespace
    {
        void rethrow(jsvd_status status)
        {
            if (status == JSVD_OK)
                return;
            const std::string msg = jsvd_last_error();
            if (status == JSVD_EINVAL)
                throw std::invalid_argument(msg);
            throw CudaError(msg.c_str());
        }

        // Same message shape as the reference's JACOBI_CUDA_CHECK
        // (src/domain/cuda_check.cuh:18-29): "<runtime> call failed: <expr> @
        // <file>:<line>, code=<n>, message=<text>".
        void check_hip(hipError_t e, const char *expr, const char *file, int line)
        {
            if (e == hipSuccess)
                return;
            (void)hipGetLastError(); // consumed here, so later launches do not report it again
            char buf[768];
            std::snprintf(buf, sizeof(buf), "HIP call failed: %s @ %s:%d, code=%d, message=%s", expr, file, line,
                          static_cast<int>(e), hipGetErrorString(e));
            throw CudaError(buf);
        }
#define JSVD_CXX_HIP(expr) check_hip((expr), #expr, __FILE__, __LINE__)

        jsvd_config to_c(const JacobiSvdConfig &c)
        {
            jsvd_config out;
            jsvd_config_default(&out);
            out.epsilon = c.epsilon;
            out.max_sweeps = c.max_sweeps;
            out.threads_per_block = c.threads_per_block;
            out.layout_mode = c.layout_transpose_mode == LayoutTransposeMode::force_enable    ? JSVD_LAYOUT_ON
                              : c.layout_transpose_mode == LayoutTransposeMode::force_disable ? JSVD_LAYOUT_OFF
                                                                                               : JSVD_LAYOUT_AUTO;
            out.layout_min_columns = c.layout_transpose_min_columns;
            out.layout_min_elements = c.layout_transpose_min_elements;
            out.layout_auto_tune = c.layout_transpose_auto_tune ? 1 : 0;
            out.layout_bench_reps = c.layout_transpose_benchmark_repetitions;
            out.layout_bench_sweeps = c.layout_transpose_benchmark_sweeps;
            out.precision = c.precision == JacobiPrecision::mixed ? JSVD_PRECISION_MIXED : JSVD_PRECISION_FP64;
            out.block_width = c.block_width;
            out.refine_sweeps = c.refine_sweeps;
            return out;
        }
    } // namespace

    JacobiSvdResult one_sided_jacobi_svd(std::span<const double> host_input, std::size_t rows, std::size_t columns,
                                         const JacobiSvdConfig &config)
    {
        // Validation order and messages follow src/domain/jacobi_svd.cu:197-223.
        if (rows == 0 || columns == 0)
            throw std::invalid_argument("Input matrix shape must be non-zero.");
        if (rows < columns)
            throw std::invalid_argument("One-sided Jacobi SVD requires rows >= columns in this implementation.");
        if (host_input.size() != rows * columns)
            throw std::invalid_argument("Input buffer size mismatch.");
        JacobiSvdResult result;
        result.rows = rows;
        result.columns = columns;
        result.u.assign(rows * columns, 0.0);
        result.sigma.assign(columns, 0.0);
        result.v.assign(columns * columns, 0.0);
        const jsvd_config cfg = to_c(config);
        int32_t sweeps = 0, stop = JSVD_STOP_CONVERGED;
        const double *a = host_input.data();
        double *u = result.u.data(), *s = result.sigma.data(), *v = result.v.data();
        const uint64_t m64 = rows, n64 = columns;
        rethrow(jsvd_solve_batched_status_f64(1, &a, &m64, &n64, &cfg, &u, &s, &v, &sweeps, &stop));
        result.converged = stop == JSVD_STOP_CONVERGED;
        result.sweeps = sweeps;
        return result;
    }

    LayoutTransposeAutoTuneReport auto_tune_layout_transpose_threshold(const JacobiSvdConfig &config)
    {
        const jsvd_config cfg = to_c(config);
        jsvd_autotune_report rep{};
        rethrow(jsvd_autotune(&cfg, &rep));
        LayoutTransposeAutoTuneReport out;
        out.executed = rep.executed != 0;
        out.recommended_min_columns = rep.recommended_min_columns;
        out.recommended_min_elements = rep.recommended_min_elements;
        out.estimated_best_speedup = rep.estimated_best_speedup;
        out.sample_count = rep.sample_count;
        return out;
    }

    DeviceMatrix::DeviceMatrix(std::size_t rows, std::size_t columns)
    {
        reset(rows, columns);
    }

    DeviceMatrix::~DeviceMatrix()
    {
        if (data_)
            (void)hipFree(data_);
    }

    DeviceMatrix::DeviceMatrix(DeviceMatrix &&other) noexcept
        : rows_(std::exchange(other.rows_, 0)), columns_(std::exchange(other.columns_, 0)),
          data_(std::exchange(other.data_, nullptr))
    {
    }

    DeviceMatrix &DeviceMatrix::operator=(DeviceMatrix &&other) noexcept
    {
        if (this != &other)
        {
            if (data_)
                (void)hipFree(data_);
            rows_ = std::exchange(other.rows_, 0);
            columns_ = std::exchange(other.columns_, 0);
            data_ = std::exchange(other.data_, nullptr);
        }
        return *this;
    }

    // As the reference (src/domain/device_matrix.cu:48-63): zero dimensions
    // are rejected before the old allocation is released.
    void DeviceMatrix::reset(std::size_t rows, std::size_t columns)
    {
        if (rows == 0 || columns == 0)
            throw std::invalid_argument("DeviceMatrix dimensions must be positive.");
        if (columns > SIZE_MAX / sizeof(double) / rows)
            throw std::overflow_error("DeviceMatrix size exceeds size_t.");
        if (data_)
        {
            JSVD_CXX_HIP(hipFree(data_));
            data_ = nullptr;
        }
        rows_ = rows;
        columns_ = columns;
        JSVD_CXX_HIP(hipMalloc(reinterpret_cast<void **>(&data_), bytes()));
    }

    std::size_t DeviceMatrix::rows() const noexcept
    {
        return rows_;
    }
    std::size_t DeviceMatrix::columns() const noexcept
    {
        return columns_;
    }
    std::size_t DeviceMatrix::size() const noexcept
    {
        return rows_ * columns_;
    }
    std::size_t DeviceMatrix::bytes() const noexcept
    {
        return size() * sizeof(double);
    }
    double *DeviceMatrix::data() noexcept
    {
        return data_;
    }
    const double *DeviceMatrix::data() const noexcept
    {
        return data_;
    }

    void DeviceMatrix::copy_from_host(std::span<const double> host_values)
    {
        if (host_values.size() != size())
            throw std::invalid_argument("Host data size does not match device matrix shape.");
        if (size() > 0)
            JSVD_CXX_HIP(hipMemcpy(data_, host_values.data(), bytes(), hipMemcpyHostToDevice));
    }

    std::vector<double> DeviceMatrix::copy_to_host() const
    {
        std::vector<double> out(size(), 0.0);
        if (size() > 0)
            JSVD_CXX_HIP(hipMemcpy(out.data(), data_, bytes(), hipMemcpyDeviceToHost));
        return out;
    }
} // namespace jacobi::svd
