#pragma once
// Public solve API (reference include/jacobi/svd/domain/jacobi_svd.cuh:25-37).
// Implemented over the C-ABI in include/jsvd_c.h; validation failures throw
// std::invalid_argument with the reference's messages, device failures throw
// CudaError.
#include "jacobi/svd/domain/jacobi_svd_config.cuh"
#include "jacobi/svd/domain/jacobi_svd_result.cuh"
#include "jacobi/svd/domain/layout_transpose.cuh"

#include <cstddef>
#include <span>

namespace jacobi::svd
{
    [[nodiscard]] JacobiSvdResult one_sided_jacobi_svd(std::span<const double> host_input,
                                                       std::size_t rows,
                                                       std::size_t columns,
                                                       const JacobiSvdConfig &config = {});

    [[nodiscard]] LayoutTransposeAutoTuneReport auto_tune_layout_transpose_threshold(
        const JacobiSvdConfig &config = {});
} // namespace jacobi::svd
