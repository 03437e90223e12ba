// cli.hpp -- command-line front end of the batched pipeline.  Same option
// table, defaults, normalisation, validation and reports as the reference
// CLI (src/interfaces/cli_parser.cu:14-43, cli_support.cu:271-464,
// cli_types.cuh:53-149, main.cu:17-81); argument errors exit with 2, other
// failures with 1.
#pragma once

#include "jacobi/svd/application/pipeline.cuh"

#include <cstddef>
#include <filesystem>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace jacobi::svd::cli
{
    class CliArgumentError final : public std::runtime_error
    {
    public:
        explicit CliArgumentError(const std::string &message) : std::runtime_error(message) {}
    };

    enum class ParseAction
    {
        run,
        help,
        version
    };

    struct CliOptions final
    {
        std::filesystem::path input_path;
        std::filesystem::path output_path;
        pipeline::MatrixFileFormat input_format = pipeline::MatrixFileFormat::auto_detect;
        pipeline::MatrixFileFormat output_format = pipeline::MatrixFileFormat::auto_detect;
        double epsilon = 1.0e-9;
        int max_sweeps = 128;
        int threads_per_block = 256;
        LayoutTransposeMode layout_transpose_mode = LayoutTransposeMode::auto_select;
        int layout_transpose_min_columns = 16;
        std::size_t layout_transpose_min_elements = 4096;
        bool layout_transpose_auto_tune = false;
        int layout_transpose_benchmark_repetitions = 2;
        int layout_transpose_benchmark_sweeps = 8;
        std::size_t queue_capacity = 4;
        bool force_overwrite = false;
        bool dry_run = false;
        bool print_config = false;
        bool json_report = false;
        bool quiet = false;
        // No reference counterpart: arithmetic of the solve (--precision fp64|mixed)
        // and the column block width of large matrices (--block-width 0|16|32).
        JacobiPrecision precision = JacobiPrecision::fp64;
        int block_width = 0;
        // No reference counterpart: per-testcase error metrics in the report
        // (--report-errors), against reference singular values when
        // --reference-sigma names a stream of 1 x n rows (implies
        // --report-errors).
        bool report_errors = false;
        std::filesystem::path reference_sigma_path;
    };

    struct ParseResult final
    {
        ParseAction action = ParseAction::run;
        CliOptions options;
    };

    enum class OptionId
    {
        input, output, input_format, output_format, format, epsilon, max_sweeps, threads_per_block,
        layout_transpose_mode, layout_transpose_min_columns, layout_transpose_min_elements,
        layout_transpose_auto_tune, layout_transpose_benchmark_repetitions, layout_transpose_benchmark_sweeps,
        queue_capacity, force, dry_run, print_config, json_report, quiet, help, version, precision, block_width,
        report_errors, reference_sigma
    };

    struct OptionDefinition final
    {
        OptionId id;
        std::string_view long_name;
        char short_name;
        bool requires_value;
        std::string_view value_hint;
        std::string_view description;
    };

    class ArgParser final
    {
    public:
        ArgParser();
        [[nodiscard]] ParseResult parse(int argc, char *const argv[]) const;
        [[nodiscard]] std::string help_message(std::string_view executable) const;

    private:
        void parse_long(std::string_view token, char *const argv[], int argc, std::size_t &index,
                        ParseResult &result) const;
        void parse_short(std::string_view token, char *const argv[], int argc, std::size_t &index,
                         ParseResult &result) const;
        void apply(const OptionDefinition &option, const std::optional<std::string_view> &value,
                   ParseResult &result) const;
        [[nodiscard]] const OptionDefinition *find_long(std::string_view name) const;
        [[nodiscard]] const OptionDefinition *find_short(char name) const;
        [[nodiscard]] std::string_view next_value(std::string_view name, char *const argv[], int argc,
                                                  std::size_t &index) const;
        std::vector<OptionDefinition> options_;
    };

    void normalize_run_options(CliOptions &options);
    void validate_run_options(const CliOptions &options);
    [[nodiscard]] pipeline::PipelineConfig make_pipeline_config(const CliOptions &options);
    void print_dry_run_config(const CliOptions &options, bool banner);
    void print_text_report(const pipeline::PipelineReport &report, double elapsed_ms);
    void print_json_report(const pipeline::PipelineReport &report, double elapsed_ms);
    [[nodiscard]] std::string_view format_name(pipeline::MatrixFileFormat format);
    [[nodiscard]] std::string_view layout_name(LayoutTransposeMode mode);

    // The whole program: returns the process exit code.
    int run_main(int argc, char *argv[]);
} // namespace jacobi::svd::cli
