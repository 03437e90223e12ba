#pragma once
// Dense row-major FP64 matrix as read from and written to matrix streams
// (reference include/jacobi/svd/io/matrix.cuh).
#include <cstddef>
#include <vector>

namespace jacobi::svd::io
{
    struct Matrix final
    {
        std::size_t rows = 0;
        std::size_t columns = 0;
        std::vector<double> values;
    };
} // namespace jacobi::svd::io
