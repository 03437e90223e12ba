#pragma once
// Text codec (reference include/jacobi/svd/io/txt_file.cuh): whitespace-separated
// rows, matrices separated by blank lines, written with max_digits10 digits.
#include "jacobi/svd/io/matrix.cuh"

#include <filesystem>
#include <memory>
#include <span>
#include <vector>

namespace jacobi::svd::io
{
    class TxtReader;
    class TxtWriter;

    struct TxtFilePolicy final
    {
        using Reader = TxtReader;
        using Writer = TxtWriter;
        [[nodiscard]] static Reader open_reader(const std::filesystem::path &path);
        [[nodiscard]] static Writer open_writer(const std::filesystem::path &path);
        // Returns false at end of file; throws std::invalid_argument on a bad
        // token or a row with a different column count.
        [[nodiscard]] static bool read_next(Reader &reader, Matrix &matrix);
        static void write_next(Writer &writer, const Matrix &matrix);
        static void flush(Writer &writer);
        [[nodiscard]] static std::vector<Matrix> read(const std::filesystem::path &path);
        static void write(const std::filesystem::path &path, std::span<const Matrix> matrices);
    };

    class TxtReader final
    {
    public:
        explicit TxtReader(const std::filesystem::path &path);
        ~TxtReader();
        TxtReader(const TxtReader &) = delete;
        TxtReader &operator=(const TxtReader &) = delete;
        TxtReader(TxtReader &&other) noexcept;
        TxtReader &operator=(TxtReader &&other) noexcept;

    private:
        struct Impl;
        std::unique_ptr<Impl> impl_;
        friend struct TxtFilePolicy;
    };

    class TxtWriter final
    {
    public:
        explicit TxtWriter(const std::filesystem::path &path);
        ~TxtWriter();
        TxtWriter(const TxtWriter &) = delete;
        TxtWriter &operator=(const TxtWriter &) = delete;
        TxtWriter(TxtWriter &&other) noexcept;
        TxtWriter &operator=(TxtWriter &&other) noexcept;

    private:
        struct Impl;
        std::unique_ptr<Impl> impl_;
        friend struct TxtFilePolicy;
    };
} // namespace jacobi::svd::io
