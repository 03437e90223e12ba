#pragma once
// Umbrella header of the domain API (reference include/jacobi/svd/domain/kernels.cuh).
#include "jacobi/svd/domain/cuda_error.cuh"
#include "jacobi/svd/domain/device_matrix.cuh"
#include "jacobi/svd/domain/jacobi_svd.cuh"
#include "jacobi/svd/domain/jacobi_svd_config.cuh"
#include "jacobi/svd/domain/jacobi_svd_result.cuh"
#include "jacobi/svd/domain/layout_transpose.cuh"
