#pragma once
// Batched pipeline API (reference include/jacobi/svd/application/pipeline.cuh:8-142):
// reads a matrix stream, solves every matrix, and writes U, Sigma (1 x n) and
// V for each input matrix, in input order.
#include "jacobi/svd/domain/kernels.cuh"

#include <cstddef>
#include <filesystem>
#include <vector>

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
        // Appended (no reference counterpart): compute per-testcase error
        // metrics on the device (orthogonality of U and V, relative
        // residual) and, when reference_sigma_path names a matrix stream of
        // 1 x n rows (one per testcase, input order, same formats as the
        // input), the sorted-sigma error against it.
        bool report_errors = false;
        std::filesystem::path reference_sigma_path{};
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
        // Appended (no reference counterpart; the reference stops silently
        // at max_sweeps): per testcase, in input order.
        std::vector<int> matrix_sweeps;
        std::vector<bool> matrix_converged; // false: max_sweeps ran out with rotations still applied
        std::size_t unconverged_count = 0;
        // Filled when PipelineConfig::report_errors is set.
        struct MatrixErrors
        {
            double orth_u = 0.0;          // ||U^T U - D||_F, D = diag(sigma > epsilon)
            double orth_v = 0.0;          // ||V^T V - I||_F
            double residual = 0.0;        // ||A - U S V^T||_F / ||A||_F
            double sigma_err_norm = -1.0; // max |s - r| / max r, both sorted (-1: no reference)
            double sigma_err_comp = -1.0; // max |s - r| / r over r > 0
        };
        bool errors_reported = false;
        std::vector<MatrixErrors> matrix_errors;
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
