#pragma once
// Streaming *.mat codec (reference include/jacobi/svd/io/mat_file.cuh):
// a file is a sequence of records, each a MatMetaData header and a payload.
#include "jacobi/svd/io/matrix.cuh"

#include <filesystem>
#include <memory>
#include <span>
#include <vector>

namespace jacobi::svd::io
{
    class MatReader;
    class MatWriter;

    struct MatFilePolicy final
    {
        using Reader = MatReader;
        using Writer = MatWriter;
        [[nodiscard]] static Reader open_reader(const std::filesystem::path &path);
        [[nodiscard]] static Writer open_writer(const std::filesystem::path &path);
        // Returns false at a clean end of file; throws std::runtime_error on a
        // truncated header or payload.
        [[nodiscard]] static bool read_next(Reader &reader, Matrix &matrix);
        static void write_next(Writer &writer, const Matrix &matrix);
        static void flush(Writer &writer);
        [[nodiscard]] static std::vector<Matrix> read(const std::filesystem::path &path);
        static void write(const std::filesystem::path &path, std::span<const Matrix> matrices);
    };

    class MatReader final
    {
    public:
        explicit MatReader(const std::filesystem::path &path);
        ~MatReader();
        MatReader(const MatReader &) = delete;
        MatReader &operator=(const MatReader &) = delete;
        MatReader(MatReader &&other) noexcept;
        MatReader &operator=(MatReader &&other) noexcept;

    private:
        struct Impl;
        std::unique_ptr<Impl> impl_;
        friend struct MatFilePolicy;
    };

    class MatWriter final
    {
    public:
        explicit MatWriter(const std::filesystem::path &path);
        ~MatWriter();
        MatWriter(const MatWriter &) = delete;
        MatWriter &operator=(const MatWriter &) = delete;
        MatWriter(MatWriter &&other) noexcept;
        MatWriter &operator=(MatWriter &&other) noexcept;

    private:
        struct Impl;
        std::unique_ptr<Impl> impl_;
        friend struct MatFilePolicy;
    };
} // namespace jacobi::svd::io
