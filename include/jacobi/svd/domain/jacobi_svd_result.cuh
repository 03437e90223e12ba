#pragma once
// Host result of one solve (reference include/jacobi/svd/domain/jacobi_svd_result.cuh:11-42):
// U m x n row-major, sigma n values in final column order (unsorted),
// V n x n row-major, and the sweep count including the verifying sweep.
#include <cstddef>
#include <vector>

namespace jacobi::svd
{
    struct JacobiSvdResult final
    {
        std::size_t rows = 0;
        std::size_t columns = 0;
        std::vector<double> u;
        std::vector<double> sigma;
        std::vector<double> v;
        int sweeps = 0;
        // Appended (the reference stops silently at max_sweeps): true when the
        // last sweep rotated no pair, false when max_sweeps ran out first.
        bool converged = true;
    };
} // namespace jacobi::svd
