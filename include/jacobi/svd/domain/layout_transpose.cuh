#pragma once
// Layout policy and its autotune report (reference
// include/jacobi/svd/domain/layout_transpose.cuh:10-57).  The gfx950 solver
// always works on a column-major copy, so the policy is accepted and
// validated but does not change the arithmetic.
#include <cstddef>

namespace jacobi::svd
{
    enum class LayoutTransposeMode
    {
        auto_select,
        force_enable,
        force_disable
    };

    struct LayoutTransposeAutoTuneReport final
    {
        bool executed = false;
        int recommended_min_columns = 16;
        std::size_t recommended_min_elements = 4096;
        double estimated_best_speedup = 1.0;
        std::size_t sample_count = 0;
    };
} // namespace jacobi::svd
