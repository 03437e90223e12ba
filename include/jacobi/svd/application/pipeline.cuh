#pragma once
// Batched pipeline API (reference include/jacobi/svd/application/pipeline.cuh:8-142):
// reads a matrix stream, solves every matrix, and writes U, Sigma (1 x n) and
// V for each input matrix, in input order.
#include "jacobi/svd/domain/kernels.cuh"

#include <cstddef>
#include <filesystem>

namespace jacobi::svd::pipeline
{
    enum class MatrixFileFormat
    {
        auto_detect, // from the file extension (.mat / .txt)
        mat,
        txt
    };

    struct PipelineConfig final
    {
        std::filesystem::path input_path;
        std::filesystem::path output_path;
        MatrixFileFormat input_format = MatrixFileFormat::auto_detect;
        MatrixFileFormat output_format = MatrixFileFormat::auto_detect;
        // Bounded in-flight window: the number of device batches that may be
        // in flight before the oldest is written out.
        std::size_t max_queued_results = 4;
        JacobiSvdConfig kernel_config{};
    };

    struct PipelineReport final
    {
        std::size_t testcase_count = 0;
        std::size_t emitted_matrix_count = 0;
        std::size_t total_sweeps = 0;
        LayoutTransposeMode layout_transpose_mode = LayoutTransposeMode::auto_select;
        int layout_transpose_min_columns = 16;
        std::size_t layout_transpose_min_elements = 4096;
        bool layout_transpose_auto_tuned = false;
        double layout_transpose_estimated_best_speedup = 1.0;
    };

    class JacobiSvdPipeline final
    {
    public:
        explicit JacobiSvdPipeline(PipelineConfig config);
        [[nodiscard]] PipelineReport run() const;

    private:
        PipelineConfig config_;
    };

    [[nodiscard]] PipelineReport run_pipeline(const PipelineConfig &config);
} // namespace jacobi::svd::pipeline
