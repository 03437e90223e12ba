#pragma once
// Umbrella header of the IO API (reference include/jacobi/svd/io/io.cuh).
#include "jacobi/svd/io/files.cuh"
#include "jacobi/svd/io/mat_dispatch.cuh"
#include "jacobi/svd/io/mat_file.cuh"
#include "jacobi/svd/io/mat_metadata.cuh"
#include "jacobi/svd/io/matrix.cuh"
#include "jacobi/svd/io/matrix_stream.cuh"
#include "jacobi/svd/io/pinned_host_task_buffer.cuh"
#include "jacobi/svd/io/txt_file.cuh"
