// cli.cpp -- argument parsing, normalisation, validation and reports of the
// pipeline CLI (behaviour of the reference's src/interfaces/*).
#include "cli.hpp"

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <system_error>

namespace jacobi::svd::cli
{
    namespace
    {
        std::string lower(std::string_view text)
        {
            std::string s(text);
            std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
            return s;
        }

        pipeline::MatrixFileFormat parse_format(std::string_view raw, std::string_view name)
        {
            const std::string v = lower(raw);
            if (v == "auto")
                return pipeline::MatrixFileFormat::auto_detect;
            if (v == "mat")
                return pipeline::MatrixFileFormat::mat;
            if (v == "txt")
                return pipeline::MatrixFileFormat::txt;
            throw CliArgumentError("Invalid value for --" + std::string(name) + ": expected one of {auto, mat, txt}.");
        }

        LayoutTransposeMode parse_layout(std::string_view raw, std::string_view name)
        {
            const std::string v = lower(raw);
            if (v == "auto")
                return LayoutTransposeMode::auto_select;
            if (v == "on")
                return LayoutTransposeMode::force_enable;
            if (v == "off")
                return LayoutTransposeMode::force_disable;
            throw CliArgumentError("Invalid value for --" + std::string(name) + ": expected one of {auto, on, off}.");
        }

        JacobiPrecision parse_precision(std::string_view raw, std::string_view name)
        {
            const std::string v = lower(raw);
            if (v == "fp64")
                return JacobiPrecision::fp64;
            if (v == "mixed")
                return JacobiPrecision::mixed;
            throw CliArgumentError("Invalid value for --" + std::string(name) + ": expected one of {fp64, mixed}.");
        }

        int parse_block_width(std::string_view raw, std::string_view name)
        {
            if (raw == "0")
                return 0;
            if (raw == "16")
                return 16;
            if (raw == "32")
                return 32;
            throw CliArgumentError("Invalid value for --" + std::string(name) + ": expected one of {0, 16, 32}.");
        }

        std::optional<pipeline::MatrixFileFormat> format_from_extension(const std::filesystem::path &path)
        {
            const std::string ext = lower(path.extension().string());
            if (ext == ".mat")
                return pipeline::MatrixFileFormat::mat;
            if (ext == ".txt")
                return pipeline::MatrixFileFormat::txt;
            return std::nullopt;
        }

        std::string extension_of(pipeline::MatrixFileFormat f)
        {
            return f == pipeline::MatrixFileFormat::txt ? ".txt" : ".mat";
        }

        template <typename T, typename Parse>
        T parse_number(std::string_view raw, std::string_view name, const char *kind, Parse parse)
        {
            const std::string text(raw);
            std::size_t used = 0;
            T value{};
            try
            {
                value = parse(text, &used);
            }
            catch (const std::exception &)
            {
                throw CliArgumentError(std::string("Invalid ") + kind + " value for --" + std::string(name) + ".");
            }
            if (used != text.size())
                throw CliArgumentError(std::string("Invalid ") + kind + " value for --" + std::string(name) + ".");
            return value;
        }

        int positive_int(std::string_view raw, std::string_view name)
        {
            const int v = parse_number<int>(raw, name, "integer",
                                            [](const std::string &t, std::size_t *u) { return std::stoi(t, u, 10); });
            if (v <= 0)
                throw CliArgumentError("Option --" + std::string(name) + " must be positive.");
            return v;
        }

        std::size_t positive_size(std::string_view raw, std::string_view name)
        {
            const unsigned long long v = parse_number<unsigned long long>(
                raw, name, "integer", [](const std::string &t, std::size_t *u) { return std::stoull(t, u, 10); });
            if (v == 0)
                throw CliArgumentError("Option --" + std::string(name) + " must be positive.");
            return static_cast<std::size_t>(v);
        }

        double positive_double(std::string_view raw, std::string_view name)
        {
            const double v = parse_number<double>(raw, name, "floating-point",
                                                  [](const std::string &t, std::size_t *u) { return std::stod(t, u); });
            if (!std::isfinite(v) || v <= 0.0)
                throw CliArgumentError("Option --" + std::string(name) + " must be a positive finite number.");
            return v;
        }

        std::filesystem::path comparable(const std::filesystem::path &p)
        {
            std::error_code ec;
            const std::filesystem::path abs = std::filesystem::absolute(p, ec);
            return ec ? p.lexically_normal() : abs.lexically_normal();
        }
    } // namespace

    ArgParser::ArgParser()
        : options_{
              {OptionId::input, "input", 'i', true, "PATH", "Input matrix stream file path."},
              {OptionId::output, "output", 'o', true, "PATH", "Output matrix stream file path."},
              {OptionId::input_format, "input-format", '\0', true, "FMT", "Input format: auto|mat|txt."},
              {OptionId::output_format, "output-format", '\0', true, "FMT", "Output format: auto|mat|txt."},
              {OptionId::format, "format", 'f', true, "FMT", "Set both input/output format: auto|mat|txt."},
              {OptionId::epsilon, "epsilon", 'e', true, "NUM", "Convergence epsilon (>0)."},
              {OptionId::max_sweeps, "max-sweeps", 's', true, "N", "Maximum sweep count (>0)."},
              {OptionId::threads_per_block, "threads-per-block", 't', true, "N",
               "Threads per workgroup (>0; rounded to a multiple of 64)."},
              {OptionId::layout_transpose_mode, "layout-transpose-mode", '\0', true, "MODE",
               "Layout-transpose mode: auto|on|off."},
              {OptionId::layout_transpose_min_columns, "layout-transpose-min-cols", '\0', true, "N",
               "Auto-mode threshold: minimum columns (>0)."},
              {OptionId::layout_transpose_min_elements, "layout-transpose-min-elems", '\0', true, "N",
               "Auto-mode threshold: minimum elements (>0)."},
              {OptionId::layout_transpose_auto_tune, "layout-transpose-auto-tune", '\0', false, "",
               "Run the kernel-shape autotuner before the pipeline."},
              {OptionId::layout_transpose_benchmark_repetitions, "layout-transpose-bench-reps", '\0', true, "N",
               "Auto-tune repetitions per candidate (>0)."},
              {OptionId::layout_transpose_benchmark_sweeps, "layout-transpose-bench-sweeps", '\0', true, "N",
               "Auto-tune sweep cap per benchmark run (>0)."},
              {OptionId::queue_capacity, "queue-capacity", 'c', true, "N", "In-flight/reorder window size (>0)."},
              {OptionId::force, "force", 'y', false, "", "Overwrite existing output file."},
              {OptionId::dry_run, "dry-run", '\0', false, "", "Validate arguments and print config only."},
              {OptionId::print_config, "print-config", '\0', false, "", "Print effective config before execution."},
              {OptionId::json_report, "json-report", '\0', false, "", "Print execution report in JSON."},
              {OptionId::quiet, "quiet", 'q', false, "", "Suppress text report."},
              {OptionId::help, "help", 'h', false, "", "Show this help message."},
              {OptionId::version, "version", 'v', false, "", "Show version information."},
              {OptionId::precision, "precision", '\0', true, "P", "Arithmetic: fp64|mixed (default fp64)."},
              {OptionId::block_width, "block-width", '\0', true, "B",
               "Blocked Jacobi for large matrices: 0 (off, default), 16 or 32."},
              {OptionId::report_errors, "report-errors", '\0', false, "",
               "Report per-testcase orthogonality of U and V and the relative residual."},
              {OptionId::reference_sigma, "reference-sigma", '\0', true, "PATH",
               "Stream of reference singular values (1 x n per testcase); implies --report-errors."},
          }
    {
    }

    ParseResult ArgParser::parse(int argc, char *const argv[]) const
    {
        ParseResult result;
        std::vector<std::string_view> positional;
        std::size_t i = 1;
        while (i < static_cast<std::size_t>(argc))
        {
            const std::string_view tok(argv[i] ? argv[i] : "");
            if (tok.empty())
            {
                ++i;
                continue;
            }
            if (tok == "--")
            {
                for (++i; i < static_cast<std::size_t>(argc); ++i)
                    positional.emplace_back(argv[i] ? argv[i] : "");
                break;
            }
            if (tok.rfind("--", 0) == 0)
                parse_long(tok, argv, argc, i, result);
            else if (tok.size() > 1 && tok[0] == '-')
                parse_short(tok, argv, argc, i, result);
            else
                positional.push_back(tok);
            ++i;
        }
        std::size_t p = 0;
        if (result.options.input_path.empty() && p < positional.size())
            result.options.input_path = std::filesystem::path(std::string(positional[p++]));
        if (result.options.output_path.empty() && p < positional.size())
            result.options.output_path = std::filesystem::path(std::string(positional[p++]));
        if (p < positional.size())
            throw CliArgumentError("Unexpected positional argument: " + std::string(positional[p]));
        if (result.action == ParseAction::run)
        {
            normalize_run_options(result.options);
            validate_run_options(result.options);
        }
        return result;
    }

    std::string ArgParser::help_message(std::string_view exe) const
    {
        std::ostringstream s;
        s << "Jacobi SVD CLI (MI355X)\n\nUsage:\n  " << exe << " [OPTIONS] <input> [output]\n\nOptions:\n";
        for (const OptionDefinition &o : options_)
        {
            std::ostringstream names;
            if (o.short_name != '\0')
                names << '-' << o.short_name << ", ";
            else
                names << "    ";
            names << "--" << o.long_name;
            if (o.requires_value)
                names << ' ' << o.value_hint;
            const std::string n = names.str();
            s << "  " << n << std::string(n.size() < 32 ? 32 - n.size() : 1, ' ') << o.description << '\n';
        }
        s << "\nExamples:\n"
          << "  " << exe << " -i inputs/a.mat -o outputs/r.mat\n"
          << "  " << exe << " inputs/a.mat --print-config\n"
          << "  " << exe << " input.txt output.txt --format txt --epsilon 1e-10\n"
          << "  " << exe << " --input a.mat --output b.txt --output-format txt --json-report --force\n"
          << "\nNotes:\n"
          << "  - When [output] is omitted, default output is <input-stem>.svd.{mat|txt}.\n"
          << "  - Existing output file requires --force to overwrite.\n";
        return s.str();
    }

    void ArgParser::parse_long(std::string_view tok, char *const argv[], int argc, std::size_t &index,
                               ParseResult &result) const
    {
        const std::string_view body = tok.substr(2);
        if (body.empty())
            throw CliArgumentError("Invalid option token '--'.");
        const std::size_t eq = body.find('=');
        const std::string_view name = eq == std::string_view::npos ? body : body.substr(0, eq);
        const OptionDefinition *o = find_long(name);
        if (!o)
            throw CliArgumentError("Unknown option: --" + std::string(name));
        if (o->requires_value)
        {
            apply(*o, eq != std::string_view::npos ? body.substr(eq + 1) : next_value(o->long_name, argv, argc, index),
                  result);
            return;
        }
        if (eq != std::string_view::npos)
            throw CliArgumentError("Option --" + std::string(o->long_name) + " does not take a value.");
        apply(*o, std::nullopt, result);
    }

    void ArgParser::parse_short(std::string_view tok, char *const argv[], int argc, std::size_t &index,
                                ParseResult &result) const
    {
        for (std::size_t off = 1; off < tok.size(); ++off)
        {
            const OptionDefinition *o = find_short(tok[off]);
            if (!o)
                throw CliArgumentError(std::string("Unknown short option: -") + tok[off]);
            if (!o->requires_value)
            {
                apply(*o, std::nullopt, result);
                continue;
            }
            // A value-taking short option consumes the rest of the token, or the next argument.
            apply(*o, off + 1 < tok.size() ? tok.substr(off + 1) : next_value(o->long_name, argv, argc, index), result);
            return;
        }
    }

    void ArgParser::apply(const OptionDefinition &o, const std::optional<std::string_view> &v,
                          ParseResult &r) const
    {
        CliOptions &opt = r.options;
        switch (o.id)
        {
        case OptionId::input: opt.input_path = std::string(*v); return;
        case OptionId::output: opt.output_path = std::string(*v); return;
        case OptionId::input_format: opt.input_format = parse_format(*v, o.long_name); return;
        case OptionId::output_format: opt.output_format = parse_format(*v, o.long_name); return;
        case OptionId::format:
            opt.input_format = opt.output_format = parse_format(*v, o.long_name);
            return;
        case OptionId::epsilon: opt.epsilon = positive_double(*v, o.long_name); return;
        case OptionId::max_sweeps: opt.max_sweeps = positive_int(*v, o.long_name); return;
        case OptionId::threads_per_block: opt.threads_per_block = positive_int(*v, o.long_name); return;
        case OptionId::layout_transpose_mode: opt.layout_transpose_mode = parse_layout(*v, o.long_name); return;
        case OptionId::layout_transpose_min_columns:
            opt.layout_transpose_min_columns = positive_int(*v, o.long_name);
            return;
        case OptionId::layout_transpose_min_elements:
            opt.layout_transpose_min_elements = positive_size(*v, o.long_name);
            return;
        case OptionId::layout_transpose_auto_tune: opt.layout_transpose_auto_tune = true; return;
        case OptionId::layout_transpose_benchmark_repetitions:
            opt.layout_transpose_benchmark_repetitions = positive_int(*v, o.long_name);
            return;
        case OptionId::layout_transpose_benchmark_sweeps:
            opt.layout_transpose_benchmark_sweeps = positive_int(*v, o.long_name);
            return;
        case OptionId::queue_capacity: opt.queue_capacity = positive_size(*v, o.long_name); return;
        case OptionId::force: opt.force_overwrite = true; return;
        case OptionId::dry_run: opt.dry_run = true; return;
        case OptionId::print_config: opt.print_config = true; return;
        case OptionId::json_report: opt.json_report = true; return;
        case OptionId::quiet: opt.quiet = true; return;
        case OptionId::help: r.action = ParseAction::help; return;
        case OptionId::version: r.action = ParseAction::version; return;
        case OptionId::precision: opt.precision = parse_precision(*v, o.long_name); return;
        case OptionId::block_width: opt.block_width = parse_block_width(*v, o.long_name); return;
        case OptionId::report_errors: opt.report_errors = true; return;
        case OptionId::reference_sigma:
            opt.reference_sigma_path = std::filesystem::path(std::string(*v));
            opt.report_errors = true;
            return;
        }
    }

    const OptionDefinition *ArgParser::find_long(std::string_view name) const
    {
        for (const OptionDefinition &o : options_)
            if (o.long_name == name)
                return &o;
        return nullptr;
    }

    const OptionDefinition *ArgParser::find_short(char name) const
    {
        for (const OptionDefinition &o : options_)
            if (o.short_name != '\0' && o.short_name == name)
                return &o;
        return nullptr;
    }

    std::string_view ArgParser::next_value(std::string_view name, char *const argv[], int argc,
                                           std::size_t &index) const
    {
        if (index + 1 >= static_cast<std::size_t>(argc) || argv[index + 1] == nullptr)
            throw CliArgumentError("Option --" + std::string(name) + " requires a value.");
        ++index;
        return std::string_view(argv[index]);
    }

    void normalize_run_options(CliOptions &o)
    {
        if (o.input_path.empty())
            throw CliArgumentError("Missing input file. Use --input <path>.");
        const auto in_ext = format_from_extension(o.input_path);
        if (o.input_format == pipeline::MatrixFileFormat::auto_detect && !in_ext)
            throw CliArgumentError("Cannot infer input format from extension. Use --input-format {mat|txt}.");
        if (o.output_path.empty())
        {
            if (o.output_format == pipeline::MatrixFileFormat::auto_detect)
                o.output_format = o.input_format != pipeline::MatrixFileFormat::auto_detect ? o.input_format
                                  : in_ext                                                   ? *in_ext
                                                                                             : pipeline::MatrixFileFormat::mat;
            std::string stem = o.input_path.stem().string();
            if (stem.empty())
                stem = "result";
            o.output_path = o.input_path.parent_path() / (stem + ".svd" + extension_of(o.output_format));
        }
        const auto out_ext = format_from_extension(o.output_path);
        if (o.output_format == pipeline::MatrixFileFormat::auto_detect && !out_ext)
        {
            const pipeline::MatrixFileFormat fallback = o.input_format != pipeline::MatrixFileFormat::auto_detect
                                                            ? o.input_format
                                                        : in_ext ? *in_ext
                                                                 : pipeline::MatrixFileFormat::mat;
            if (o.output_path.extension().empty())
                o.output_path += extension_of(fallback);
            o.output_format = fallback;
        }
    }

    void validate_run_options(const CliOptions &o)
    {
        if (!std::filesystem::exists(o.input_path))
            throw CliArgumentError("Input file does not exist: " + o.input_path.string());
        if (std::filesystem::is_directory(o.input_path))
            throw CliArgumentError("Input path is a directory, expected a file: " + o.input_path.string());
        if (std::filesystem::is_directory(o.output_path))
            throw CliArgumentError("Output path points to a directory: " + o.output_path.string());
        if (comparable(o.input_path) == comparable(o.output_path))
            throw CliArgumentError("Input and output paths must be different files.");
        if (std::filesystem::exists(o.output_path) && !o.force_overwrite)
            throw CliArgumentError("Output file already exists. Use --force to overwrite: " + o.output_path.string());
        if (!o.reference_sigma_path.empty() && !std::filesystem::is_regular_file(o.reference_sigma_path))
            throw CliArgumentError("Reference sigma file does not exist: " + o.reference_sigma_path.string());
    }

    pipeline::PipelineConfig make_pipeline_config(const CliOptions &o)
    {
        pipeline::PipelineConfig c;
        c.input_path = o.input_path;
        c.output_path = o.output_path;
        c.input_format = o.input_format;
        c.output_format = o.output_format;
        c.max_queued_results = o.queue_capacity;
        c.report_errors = o.report_errors;
        c.reference_sigma_path = o.reference_sigma_path;
        c.kernel_config = JacobiSvdConfig{
            .epsilon = o.epsilon,
            .max_sweeps = o.max_sweeps,
            .threads_per_block = o.threads_per_block,
            .layout_transpose_mode = o.layout_transpose_mode,
            .layout_transpose_min_columns = o.layout_transpose_min_columns,
            .layout_transpose_min_elements = o.layout_transpose_min_elements,
            .layout_transpose_auto_tune = o.layout_transpose_auto_tune,
            .layout_transpose_benchmark_repetitions = o.layout_transpose_benchmark_repetitions,
            .layout_transpose_benchmark_sweeps = o.layout_transpose_benchmark_sweeps,
            .precision = o.precision,
            .block_width = o.block_width,
        };
        return c;
    }

    std::string_view format_name(pipeline::MatrixFileFormat f)
    {
        switch (f)
        {
        case pipeline::MatrixFileFormat::auto_detect: return "auto";
        case pipeline::MatrixFileFormat::mat: return "mat";
        case pipeline::MatrixFileFormat::txt: return "txt";
        }
        return "unknown";
    }

    std::string_view layout_name(LayoutTransposeMode m)
    {
        switch (m)
        {
        case LayoutTransposeMode::auto_select: return "auto";
        case LayoutTransposeMode::force_enable: return "on";
        case LayoutTransposeMode::force_disable: return "off";
        }
        return "unknown";
    }

    void print_dry_run_config(const CliOptions &o, bool banner)
    {
        if (banner)
            std::cout << "Dry run: pipeline was not executed.\n";
        std::cout << "input           : " << o.input_path.string() << '\n'
                  << "output          : " << o.output_path.string() << '\n'
                  << "input-format    : " << format_name(o.input_format) << '\n'
                  << "output-format   : " << format_name(o.output_format) << '\n'
                  << "epsilon         : " << o.epsilon << '\n'
                  << "max-sweeps      : " << o.max_sweeps << '\n'
                  << "threads-per-blk : " << o.threads_per_block << '\n'
                  << "layout-mode     : " << layout_name(o.layout_transpose_mode) << '\n'
                  << "layout-min-cols : " << o.layout_transpose_min_columns << '\n'
                  << "layout-min-elem : " << o.layout_transpose_min_elements << '\n'
                  << "layout-auto-tune: " << (o.layout_transpose_auto_tune ? "true" : "false") << '\n'
                  << "layout-bench-rep: " << o.layout_transpose_benchmark_repetitions << '\n'
                  << "layout-bench-swp: " << o.layout_transpose_benchmark_sweeps << '\n'
                  << "queue-capacity  : " << o.queue_capacity << '\n'
                  << "force-overwrite : " << (o.force_overwrite ? "true" : "false") << '\n'
                  << "precision       : " << (o.precision == JacobiPrecision::mixed ? "mixed" : "fp64") << '\n'
                  << "block-width     : " << o.block_width << '\n'
                  << "report-errors   : " << (o.report_errors ? "true" : "false") << '\n';
    }

    void print_text_report(const pipeline::PipelineReport &r, double ms)
    {
        std::cout << "Pipeline completed.\n"
                  << "testcases       : " << r.testcase_count << '\n'
                  << "emitted-matrices: " << r.emitted_matrix_count << '\n'
                  << "total-sweeps    : " << r.total_sweeps << '\n'
                  << "layout-mode     : " << layout_name(r.layout_transpose_mode) << '\n'
                  << "layout-min-cols : " << r.layout_transpose_min_columns << '\n'
                  << "layout-min-elem : " << r.layout_transpose_min_elements << '\n'
                  << "layout-auto-tune: " << (r.layout_transpose_auto_tuned ? "true" : "false") << '\n'
                  << "layout-best-spd : " << std::fixed << std::setprecision(3)
                  << r.layout_transpose_estimated_best_speedup << '\n'
                  << "elapsed-ms      : " << std::fixed << std::setprecision(3) << ms << '\n'
                  << "unconverged     : " << r.unconverged_count << '\n';
        if (r.errors_reported && !r.matrix_errors.empty())
        {
            double ou = 0.0, ov = 0.0, res = 0.0, se = -1.0;
            for (const auto &e : r.matrix_errors)
            {
                ou = std::max(ou, e.orth_u);
                ov = std::max(ov, e.orth_v);
                res = std::max(res, e.residual);
                se = std::max(se, e.sigma_err_norm);
            }
            std::cout << std::scientific << std::setprecision(3) << "max-orth-u      : " << ou << '\n'
                      << "max-orth-v      : " << ov << '\n'
                      << "max-residual    : " << res << '\n';
            if (se >= 0.0)
                std::cout << "max-sigma-err   : " << se << '\n';
            std::cout << std::defaultfloat;
        }
    }

    void print_json_report(const pipeline::PipelineReport &r, double ms)
    {
        std::cout << "{\n"
                  << "  \"testcase_count\": " << r.testcase_count << ",\n"
                  << "  \"emitted_matrix_count\": " << r.emitted_matrix_count << ",\n"
                  << "  \"total_sweeps\": " << r.total_sweeps << ",\n"
                  << "  \"layout_transpose_mode\": \"" << layout_name(r.layout_transpose_mode) << "\",\n"
                  << "  \"layout_transpose_min_columns\": " << r.layout_transpose_min_columns << ",\n"
                  << "  \"layout_transpose_min_elements\": " << r.layout_transpose_min_elements << ",\n"
                  << "  \"layout_transpose_auto_tuned\": " << (r.layout_transpose_auto_tuned ? "true" : "false")
                  << ",\n"
                  << "  \"layout_transpose_estimated_best_speedup\": " << std::fixed << std::setprecision(3)
                  << r.layout_transpose_estimated_best_speedup << ",\n"
                  << "  \"elapsed_ms\": " << std::fixed << std::setprecision(3) << ms << ",\n"
                  << "  \"unconverged_count\": " << r.unconverged_count << ",\n"
                  << "  \"matrix_sweeps\": [";
        for (std::size_t k = 0; k < r.matrix_sweeps.size(); ++k)
            std::cout << (k ? ", " : "") << r.matrix_sweeps[k];
        std::cout << "],\n  \"matrix_stop\": [";
        for (std::size_t k = 0; k < r.matrix_converged.size(); ++k)
            std::cout << (k ? ", " : "") << (r.matrix_converged[k] ? "\"converged\"" : "\"max_sweeps\"");
        std::cout << "]";
        if (r.errors_reported)
        {
            // Full precision: these are error norms near the rounding level.
            std::cout << ",\n  \"matrix_errors\": [" << std::setprecision(17) << std::scientific;
            for (std::size_t k = 0; k < r.matrix_errors.size(); ++k)
            {
                const auto &e = r.matrix_errors[k];
                std::cout << (k ? ",\n    " : "\n    ") << "{\"orth_u\": " << e.orth_u << ", \"orth_v\": " << e.orth_v
                          << ", \"residual\": " << e.residual;
                if (e.sigma_err_norm >= 0.0)
                    std::cout << ", \"sigma_err_norm\": " << e.sigma_err_norm
                              << ", \"sigma_err_comp\": " << e.sigma_err_comp;
                std::cout << "}";
            }
            std::cout << "\n  ]" << std::defaultfloat;
        }
        std::cout << "\n}\n";
    }

    int run_main(int argc, char *argv[])
    {
        try
        {
            const std::string exe =
                (argc > 0 && argv[0]) ? std::filesystem::path(argv[0]).filename().string() : "jacobi-svd-amd";
            const ArgParser parser;
            const ParseResult parsed = parser.parse(argc, argv);
            if (parsed.action == ParseAction::help)
            {
                std::cout << parser.help_message(exe);
                return 0;
            }
            if (parsed.action == ParseAction::version)
            {
                std::cout << "jacobi-svd-amd CLI v0.1.0\n";
                return 0;
            }
            if (parsed.options.dry_run)
            {
                print_dry_run_config(parsed.options, true);
                return 0;
            }
            if (parsed.options.print_config)
            {
                std::cout << "Effective configuration:\n";
                print_dry_run_config(parsed.options, false);
                std::cout << '\n';
            }
            const pipeline::PipelineConfig config = make_pipeline_config(parsed.options);
            const auto t0 = std::chrono::steady_clock::now();
            const pipeline::PipelineReport report = pipeline::run_pipeline(config);
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (!parsed.options.quiet)
                print_text_report(report, ms);
            if (parsed.options.json_report)
                print_json_report(report, ms);
            return 0;
        }
        catch (const CliArgumentError &e)
        {
            std::cerr << "Argument error: " << e.what() << "\nUse --help for usage.\n";
            return 2;
        }
        catch (const std::exception &e)
        {
            std::cerr << "Execution failed: " << e.what() << '\n';
            return 1;
        }
    }
} // namespace jacobi::svd::cli
