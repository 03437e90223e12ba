#pragma once
// Move-only owner of a row-major FP64 matrix in device memory (reference
// include/jacobi/svd/domain/device_matrix.cuh:13-127), backed by hipMalloc.
#include <cstddef>
#include <span>
#include <vector>

namespace jacobi::svd
{
    class DeviceMatrix final
    {
    public:
        DeviceMatrix() = default;
        DeviceMatrix(std::size_t rows, std::size_t columns);
        ~DeviceMatrix();
        DeviceMatrix(const DeviceMatrix &) = delete;
        DeviceMatrix &operator=(const DeviceMatrix &) = delete;
        DeviceMatrix(DeviceMatrix &&other) noexcept;
        DeviceMatrix &operator=(DeviceMatrix &&other) noexcept;

        void reset(std::size_t rows, std::size_t columns);
        [[nodiscard]] std::size_t rows() const noexcept;
        [[nodiscard]] std::size_t columns() const noexcept;
        [[nodiscard]] std::size_t size() const noexcept;
        [[nodiscard]] std::size_t bytes() const noexcept;
        [[nodiscard]] double *data() noexcept;
        [[nodiscard]] const double *data() const noexcept;
        void copy_from_host(std::span<const double> host_values);
        [[nodiscard]] std::vector<double> copy_to_host() const;

    private:
        std::size_t rows_ = 0;
        std::size_t columns_ = 0;
        double *data_ = nullptr;
    };
} // namespace jacobi::svd
