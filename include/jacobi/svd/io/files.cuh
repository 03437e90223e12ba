#pragma once
// Whole-file helpers (reference include/jacobi/svd/io/files.cuh).
#include "jacobi/svd/io/matrix.cuh"

#include <filesystem>
#include <span>
#include <vector>

namespace jacobi::svd::io
{
    [[nodiscard]] std::vector<Matrix> read_mat_file(const std::filesystem::path &path);
    void write_mat_file(const std::filesystem::path &path, std::span<const Matrix> matrices);
    [[nodiscard]] std::vector<Matrix> read_txt_file(const std::filesystem::path &path);
    void write_txt_file(const std::filesystem::path &path, std::span<const Matrix> matrices);
} // namespace jacobi::svd::io
