// CLI parser tests.  The first four restate the reference's own CLI tests
// (tests/cli_parser_tests.cu:74-141 of kleedaisuki/jacobi-svd-cuda-opt); the
// rest cover the other normalisation and validation rules.  Host only: no
// device call is made.
#include "cli.hpp"

#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace fs = std::filesystem;
using jacobi::svd::cli::ArgParser;
using jacobi::svd::cli::CliArgumentError;
using jacobi::svd::cli::ParseAction;
using jacobi::svd::cli::ParseResult;
using Fmt = jacobi::svd::pipeline::MatrixFileFormat;

namespace
{
    int failures = 0;

    void check(bool ok, const char *what)
    {
        if (!ok)
        {
            std::cerr << "FAILED: " << what << '\n';
            ++failures;
        }
    }

    ParseResult parse(std::vector<std::string> args)
    {
        std::vector<char *> argv;
        for (std::string &a : args)
            argv.push_back(a.data());
        return ArgParser().parse(static_cast<int>(argv.size()), argv.data());
    }

    bool rejects(std::vector<std::string> args, const std::string &needle = "")
    {
        try
        {
            (void)parse(std::move(args));
        }
        catch (const CliArgumentError &e)
        {
            return needle.empty() || std::string(e.what()).find(needle) != std::string::npos;
        }
        return false;
    }

    fs::path touch(const std::string &name, const std::string &content = "1 0\n0 1\n")
    {
        const fs::path p = fs::temp_directory_path() / name;
        std::ofstream(p, std::ios::binary) << content;
        return p;
    }
} // namespace

int main()
{
    // reference: --help short-circuits validation.
    check(parse({"jsvd", "--help"}).action == ParseAction::help, "help action");
    // reference: --dry-run names the default output <stem>.svd.txt and infers txt.
    {
        const fs::path in = touch("jsvd-cli-parser-input.txt");
        const ParseResult r = parse({"jsvd", "--dry-run", "--input", in.string()});
        fs::remove(in);
        check(r.options.dry_run, "dry-run flag");
        check(r.options.output_path.filename() == "jsvd-cli-parser-input.svd.txt", "default output name");
        check(r.options.input_format == Fmt::auto_detect, "input format stays auto");
        check(r.options.output_format == Fmt::txt, "output format inferred from input");
    }
    // reference: attached short-option values -fmat -e1e-8 -s64.
    {
        const fs::path in = touch("jsvd-cli-parser-input.mat", "placeholder");
        const ParseResult r = parse({"jsvd", "--dry-run", "-fmat", "-e1e-8", "-s64", in.string()});
        fs::remove(in);
        check(r.options.input_format == Fmt::mat && r.options.output_format == Fmt::mat, "-f sets both formats");
        check(r.options.epsilon == 1.0e-8, "-e value");
        check(r.options.max_sweeps == 64, "-s value");
    }
    // reference: --max-sweeps 0 is rejected.
    check(rejects({"jsvd", "--dry-run", "--max-sweeps", "0", "input.mat"}), "max-sweeps 0 rejected");

    // Further rules of the same parser.
    check(parse({"jsvd", "-v"}).action == ParseAction::version, "version action");
    check(rejects({"jsvd", "--bogus"}, "Unknown option: --bogus"), "unknown long option");
    check(rejects({"jsvd", "-Z"}, "Unknown short option: -Z"), "unknown short option");
    check(rejects({"jsvd", "--epsilon"}, "requires a value"), "missing value");
    check(rejects({"jsvd", "--force=1"}, "does not take a value"), "flag with value");
    check(rejects({"jsvd", "--dry-run"}, "Missing input file"), "missing input");
    check(rejects({"jsvd", "--dry-run", "--epsilon", "-1", "a.mat"}, "positive finite"), "negative epsilon");
    check(rejects({"jsvd", "--dry-run", "--epsilon", "abc", "a.mat"}, "Invalid floating-point"), "bad epsilon");
    check(rejects({"jsvd", "--dry-run", "--queue-capacity", "0", "a.mat"}, "must be positive"), "zero queue");
    check(rejects({"jsvd", "--format", "csv", "a.mat"}, "expected one of {auto, mat, txt}"), "bad format");
    check(rejects({"jsvd", "--layout-transpose-mode", "maybe", "a.mat"}, "{auto, on, off}"), "bad layout mode");
    check(rejects({"jsvd", "--precision", "fp16", "a.mat"}, "{fp64, mixed}"), "bad precision");
    check(rejects({"jsvd", "--block-width", "8", "a.mat"}, "{0, 16, 32}"), "bad block width");
    check(rejects({"jsvd", "--dry-run", "a.bin"}, "Cannot infer input format"), "unknown extension");
    check(rejects({"jsvd", "--dry-run", "missing-input-file.mat"}, "Input file does not exist"), "missing file");
    check(rejects({"jsvd", "a.mat", "b.mat", "c.mat"}, "Unexpected positional argument: c.mat"), "extra positional");
    {
        const fs::path in = touch("jsvd-cli-same.mat", "x");
        check(rejects({"jsvd", "--dry-run", in.string(), in.string()}, "must be different"), "input == output");
        const fs::path out = touch("jsvd-cli-exists.txt", "x");
        check(rejects({"jsvd", "--dry-run", in.string(), out.string()}, "Use --force"), "existing output");
        const ParseResult r = parse({"jsvd", "--dry-run", "--force", "--output=" + out.string(), "--input=" + in.string(),
                                     "--precision", "mixed", "-c8", "-t", "512", "--block-width=32"});
        check(r.options.force_overwrite && r.options.queue_capacity == 8 && r.options.threads_per_block == 512,
              "inline values and flags");
        // Known extensions leave both formats on auto; the pipeline resolves them.
        check(r.options.output_format == Fmt::auto_detect && r.options.input_format == Fmt::auto_detect, "formats");
        check(r.options.precision == jacobi::svd::JacobiPrecision::mixed, "precision mixed");
        check(r.options.block_width == 32, "block width");
        const ParseResult r2 = parse({"jsvd", "--dry-run", in.string(), "--output", (fs::temp_directory_path() / "jsvd-noext").string()});
        check(r2.options.output_path.extension() == ".mat" && r2.options.output_format == Fmt::mat,
              "extension appended from input format");
        fs::remove(in);
        fs::remove(out);
    }
    if (failures)
    {
        std::cerr << failures << " CLI parser check(s) failed\n";
        return 1;
    }
    std::cout << "cli_parser_tests passed\n";
    return 0;
}
