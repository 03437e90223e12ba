#pragma once
// Solver configuration.  The first nine fields, their order and defaults
// follow the reference (include/jacobi/svd/domain/jacobi_svd_config.cuh:12-58)
// so designated-initializer call sites keep compiling; the fields after them
// have no reference counterpart.
#include "jacobi/svd/domain/layout_transpose.cuh"

#include <cstddef>

namespace jacobi::svd
{
    enum class JacobiPrecision
    {
        fp64,  // plain FP64 sweeps in the reference's pair order
        mixed  // FP32 sweeps, FP64 rotation parameters and FP64 refinement sweeps
    };

    struct JacobiSvdConfig final
    {
        double epsilon = 1.0e-9;
        int max_sweeps = 128;
        int threads_per_block = 256;
        LayoutTransposeMode layout_transpose_mode = LayoutTransposeMode::auto_select;
        int layout_transpose_min_columns = 16;
        std::size_t layout_transpose_min_elements = 4096;
        bool layout_transpose_auto_tune = false;
        int layout_transpose_benchmark_repetitions = 2;
        int layout_transpose_benchmark_sweeps = 8;
        JacobiPrecision precision = JacobiPrecision::fp64;
        int block_width = 0;
        int refine_sweeps = 0; // 0: FP64 refinement runs to convergence (at most max_sweeps)
    };
} // namespace jacobi::svd
