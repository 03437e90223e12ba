#pragma once
// Device-runtime failure raised by the solver.  The type name is kept from
// the reference (include/jacobi/svd/domain/cuda_error.cuh:10-18) so callers
// that catch it keep compiling; here it wraps HIP runtime errors.
#include <stdexcept>

namespace jacobi::svd
{
    class CudaError final : public std::runtime_error
    {
    public:
        explicit CudaError(const char *message);
    };
} // namespace jacobi::svd
