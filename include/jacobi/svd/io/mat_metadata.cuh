#pragma once
// Header of one *.mat record: two 64-bit unsigned integers in network byte
// order (rows, columns), followed by rows*columns IEEE-754 doubles, also in
// network byte order, row-major (reference include/jacobi/svd/io/mat_metadata.cuh).
#include <cstdint>

namespace jacobi::svd::io
{
    struct MatMetaData final
    {
        std::uint64_t rows = 0;
        std::uint64_t columns = 0;
    };
} // namespace jacobi::svd::io
