#pragma once
// Policy-based matrix streams (reference include/jacobi/svd/io/matrix_stream.cuh).
#include "jacobi/svd/io/mat_file.cuh"
#include "jacobi/svd/io/txt_file.cuh"

#include <concepts>
#include <filesystem>

namespace jacobi::svd::io
{
    template <typename Policy>
    concept MatrixInputPolicy = requires(const std::filesystem::path &path, typename Policy::Reader &reader,
                                         Matrix &matrix) {
        { Policy::open_reader(path) } -> std::same_as<typename Policy::Reader>;
        { Policy::read_next(reader, matrix) } -> std::same_as<bool>;
    };

    template <typename Policy>
    concept MatrixOutputPolicy = requires(const std::filesystem::path &path, typename Policy::Writer &writer,
                                          const Matrix &matrix) {
        { Policy::open_writer(path) } -> std::same_as<typename Policy::Writer>;
        { Policy::write_next(writer, matrix) } -> std::same_as<void>;
        { Policy::flush(writer) } -> std::same_as<void>;
    };

    template <MatrixInputPolicy Policy>
    class MatrixInputStream final
    {
    public:
        explicit MatrixInputStream(const std::filesystem::path &path) : reader_(Policy::open_reader(path)) {}
        bool read_one(Matrix &matrix)
        {
            if (eof_)
                return false;
            eof_ = !Policy::read_next(reader_, matrix);
            return !eof_;
        }
        MatrixInputStream &operator>>(Matrix &matrix)
        {
            (void)read_one(matrix);
            return *this;
        }
        [[nodiscard]] bool eof() const noexcept { return eof_; }

    private:
        typename Policy::Reader reader_;
        bool eof_ = false;
    };

    template <MatrixOutputPolicy Policy>
    class MatrixOutputStream final
    {
    public:
        explicit MatrixOutputStream(const std::filesystem::path &path) : writer_(Policy::open_writer(path)) {}
        void write_one(const Matrix &matrix) { Policy::write_next(writer_, matrix); }
        MatrixOutputStream &operator<<(const Matrix &matrix)
        {
            write_one(matrix);
            return *this;
        }
        void flush() { Policy::flush(writer_); }

    private:
        typename Policy::Writer writer_;
    };
} // namespace jacobi::svd::io
