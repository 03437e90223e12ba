// jsvd_io.cpp -- *.mat and *.txt matrix stream codecs (host only).
//
// Formats (the reference's wire formats, include/jacobi/svd/io/mat_metadata.cuh:11-22,
// src/infrastructure/mat_file.cu:91-197, src/infrastructure/txt_file.cu:95-195):
//  * .mat: records of {u64 rows, u64 columns} followed by rows*columns
//    doubles, all in network (big-endian) byte order, row-major;
//  * .txt: one matrix row per line, values separated by whitespace, matrices
//    separated by blank lines; written with max_digits10 significant digits.
// Byte swapping runs over fixed-size chunks so large payloads are decoded
// while they stream through a bounded buffer.
#include "jacobi/svd/io/io.cuh"

#include <algorithm>
#include <bit>
#include <cctype>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <locale>
#include <sstream>
#include <stdexcept>
#include <string>

namespace jacobi::svd::io
{
    namespace
    {
        constexpr std::size_t kChunk = 1 << 15; // doubles per conversion chunk

        std::uint64_t swap_if_little(std::uint64_t x)
        {
            if constexpr (std::endian::native == std::endian::little)
                return __builtin_bswap64(x);
            return x;
        }

        std::size_t checked_mul(std::size_t a, std::size_t b)
        {
            if (a != 0 && b > std::numeric_limits<std::size_t>::max() / a)
                throw std::overflow_error("Size multiplication overflow.");
            return a * b;
        }

        void validate_layout(const Matrix &m)
        {
            if (m.values.size() != checked_mul(m.rows, m.columns))
                throw std::invalid_argument("Matrix payload size does not match rows * columns.");
        }

        bool blank(const std::string &line)
        {
            return std::all_of(line.begin(), line.end(), [](unsigned char c) { return std::isspace(c) != 0; });
        }

        std::vector<double> parse_row(const std::string &line)
        {
            std::istringstream in(line);
            in.imbue(std::locale::classic());
            std::vector<double> row;
            double x = 0.0;
            while (in >> x)
                row.push_back(x);
            if (in.fail() && !in.eof())
                throw std::invalid_argument("Invalid numeric token in text matrix row.");
            return row;
        }
    } // namespace

    struct MatReader::Impl
    {
        std::ifstream in;
        std::uint64_t size = 0; // file size, so a record's payload is checked before it is allocated
        explicit Impl(const std::filesystem::path &p) : in(p, std::ios::binary)
        {
            if (!in)
                throw std::runtime_error("Failed to open mat matrix file for reading.");
            in.seekg(0, std::ios::end);
            size = static_cast<std::uint64_t>(in.tellg());
            in.seekg(0, std::ios::beg);
        }
    };

    struct MatWriter::Impl
    {
        std::ofstream out;
        explicit Impl(const std::filesystem::path &p) : out(p, std::ios::binary | std::ios::trunc)
        {
            if (!out)
                throw std::runtime_error("Failed to open mat matrix file for writing.");
        }
    };

    MatReader::MatReader(const std::filesystem::path &path) : impl_(std::make_unique<Impl>(path)) {}
    MatReader::~MatReader() = default;
    MatReader::MatReader(MatReader &&) noexcept = default;
    MatReader &MatReader::operator=(MatReader &&) noexcept = default;
    MatWriter::MatWriter(const std::filesystem::path &path) : impl_(std::make_unique<Impl>(path)) {}
    MatWriter::~MatWriter() = default;
    MatWriter::MatWriter(MatWriter &&) noexcept = default;
    MatWriter &MatWriter::operator=(MatWriter &&) noexcept = default;

    MatFilePolicy::Reader MatFilePolicy::open_reader(const std::filesystem::path &path)
    {
        return MatReader(path);
    }
    MatFilePolicy::Writer MatFilePolicy::open_writer(const std::filesystem::path &path)
    {
        return MatWriter(path);
    }

    bool MatFilePolicy::read_next(Reader &reader, Matrix &matrix)
    {
        if (!reader.impl_)
            throw std::runtime_error("MatReader is not initialized.");
        std::ifstream &in = reader.impl_->in;
        std::uint64_t header[2];
        in.read(reinterpret_cast<char *>(header), sizeof(header));
        const std::streamsize got = in.gcount();
        if (got == 0 && in.eof())
            return false;
        if (got != static_cast<std::streamsize>(sizeof(header)))
            throw std::runtime_error("Truncated *.mat header.");
        const std::uint64_t rows = swap_if_little(header[0]), cols = swap_if_little(header[1]);
        if (rows > std::numeric_limits<std::size_t>::max() || cols > std::numeric_limits<std::size_t>::max())
            throw std::overflow_error("Matrix dimensions exceed size_t.");
        const std::size_t count = checked_mul(static_cast<std::size_t>(rows), static_cast<std::size_t>(cols));
        const std::size_t payload = checked_mul(count, sizeof(double));
        // The payload must be in the file before anything is allocated for it
        // (as src/infrastructure/mat_file.cu:123-129 checks the remaining
        // bytes first): a corrupt header then fails with the truncation
        // error instead of a huge allocation.
        const std::uint64_t here = static_cast<std::uint64_t>(in.tellg());
        if (here > reader.impl_->size || payload > reader.impl_->size - here)
            throw std::runtime_error("Truncated *.mat payload.");
        matrix.rows = static_cast<std::size_t>(rows);
        matrix.columns = static_cast<std::size_t>(cols);
        matrix.values.resize(count);
        std::uint64_t *raw = reinterpret_cast<std::uint64_t *>(matrix.values.data());
        for (std::size_t off = 0; off < count; off += kChunk)
        {
            const std::size_t n = std::min(kChunk, count - off);
            in.read(reinterpret_cast<char *>(raw + off), static_cast<std::streamsize>(n * sizeof(double)));
            if (in.gcount() != static_cast<std::streamsize>(n * sizeof(double)))
                throw std::runtime_error("Truncated *.mat payload.");
            for (std::size_t i = 0; i < n; ++i)
                raw[off + i] = swap_if_little(raw[off + i]);
        }
        return true;
    }

    void MatFilePolicy::write_next(Writer &writer, const Matrix &matrix)
    {
        if (!writer.impl_)
            throw std::runtime_error("MatWriter is not initialized.");
        validate_layout(matrix);
        std::ofstream &out = writer.impl_->out;
        const std::uint64_t header[2] = {swap_if_little(static_cast<std::uint64_t>(matrix.rows)),
                                         swap_if_little(static_cast<std::uint64_t>(matrix.columns))};
        out.write(reinterpret_cast<const char *>(header), sizeof(header));
        std::vector<std::uint64_t> buf(std::min(kChunk, matrix.values.size()));
        for (std::size_t off = 0; off < matrix.values.size(); off += kChunk)
        {
            const std::size_t n = std::min(kChunk, matrix.values.size() - off);
            for (std::size_t i = 0; i < n; ++i)
                buf[i] = swap_if_little(std::bit_cast<std::uint64_t>(matrix.values[off + i]));
            out.write(reinterpret_cast<const char *>(buf.data()), static_cast<std::streamsize>(n * sizeof(double)));
        }
        if (!out)
            throw std::runtime_error("Failed while writing mat matrix file.");
    }

    void MatFilePolicy::flush(Writer &writer)
    {
        if (!writer.impl_)
            throw std::runtime_error("MatWriter is not initialized.");
        writer.impl_->out.flush();
    }

    std::vector<Matrix> MatFilePolicy::read(const std::filesystem::path &path)
    {
        Reader reader = open_reader(path);
        std::vector<Matrix> out;
        Matrix m;
        while (read_next(reader, m))
            out.push_back(std::move(m));
        return out;
    }

    void MatFilePolicy::write(const std::filesystem::path &path, std::span<const Matrix> matrices)
    {
        Writer writer = open_writer(path);
        for (const Matrix &m : matrices)
            write_next(writer, m);
        flush(writer);
    }

    struct TxtReader::Impl
    {
        std::ifstream in;
        explicit Impl(const std::filesystem::path &p) : in(p, std::ios::in | std::ios::binary)
        {
            if (!in)
                throw std::runtime_error("Failed to open text matrix file for reading.");
            in.imbue(std::locale::classic());
        }
    };

    struct TxtWriter::Impl
    {
        std::ofstream out;
        bool wrote = false;
        explicit Impl(const std::filesystem::path &p) : out(p, std::ios::out | std::ios::binary | std::ios::trunc)
        {
            if (!out)
                throw std::runtime_error("Failed to open text matrix file for writing.");
            out.imbue(std::locale::classic());
            out << std::setprecision(std::numeric_limits<double>::max_digits10);
        }
    };

    TxtReader::TxtReader(const std::filesystem::path &path) : impl_(std::make_unique<Impl>(path)) {}
    TxtReader::~TxtReader() = default;
    TxtReader::TxtReader(TxtReader &&) noexcept = default;
    TxtReader &TxtReader::operator=(TxtReader &&) noexcept = default;
    TxtWriter::TxtWriter(const std::filesystem::path &path) : impl_(std::make_unique<Impl>(path)) {}
    TxtWriter::~TxtWriter() = default;
    TxtWriter::TxtWriter(TxtWriter &&) noexcept = default;
    TxtWriter &TxtWriter::operator=(TxtWriter &&) noexcept = default;

    TxtFilePolicy::Reader TxtFilePolicy::open_reader(const std::filesystem::path &path)
    {
        return TxtReader(path);
    }
    TxtFilePolicy::Writer TxtFilePolicy::open_writer(const std::filesystem::path &path)
    {
        return TxtWriter(path);
    }

    bool TxtFilePolicy::read_next(Reader &reader, Matrix &matrix)
    {
        if (!reader.impl_)
            throw std::runtime_error("TxtReader is not initialized.");
        std::ifstream &in = reader.impl_->in;
        std::vector<double> values;
        std::size_t rows = 0, cols = 0;
        bool started = false;
        std::string line;
        while (std::getline(in, line))
        {
            if (blank(line))
            {
                if (!started)
                    continue;
                break;
            }
            std::vector<double> row = parse_row(line);
            if (row.empty())
                throw std::invalid_argument("Text matrix row cannot be empty when line is non-blank.");
            if (!started)
            {
                started = true;
                cols = row.size();
            }
            else if (row.size() != cols)
                throw std::invalid_argument("Inconsistent column count in text matrix block.");
            values.insert(values.end(), row.begin(), row.end());
            ++rows;
        }
        if (!started)
        {
            if (!in.eof())
                throw std::runtime_error("Failed while reading text matrix file.");
            return false;
        }
        matrix.rows = rows;
        matrix.columns = cols;
        matrix.values = std::move(values);
        return true;
    }

    void TxtFilePolicy::write_next(Writer &writer, const Matrix &matrix)
    {
        if (!writer.impl_)
            throw std::runtime_error("TxtWriter is not initialized.");
        validate_layout(matrix);
        auto &st = *writer.impl_;
        if (st.wrote)
            st.out << '\n';
        for (std::size_t r = 0; r < matrix.rows; ++r)
        {
            for (std::size_t c = 0; c < matrix.columns; ++c)
            {
                if (c > 0)
                    st.out << ' ';
                st.out << matrix.values[r * matrix.columns + c];
            }
            st.out << '\n';
        }
        if (!st.out)
            throw std::runtime_error("Failed while writing text matrix file.");
        st.wrote = true;
    }

    void TxtFilePolicy::flush(Writer &writer)
    {
        if (!writer.impl_)
            throw std::runtime_error("TxtWriter is not initialized.");
        writer.impl_->out.flush();
    }

    std::vector<Matrix> TxtFilePolicy::read(const std::filesystem::path &path)
    {
        Reader reader = open_reader(path);
        std::vector<Matrix> out;
        Matrix m;
        while (read_next(reader, m))
            out.push_back(std::move(m));
        return out;
    }

    void TxtFilePolicy::write(const std::filesystem::path &path, std::span<const Matrix> matrices)
    {
        Writer writer = open_writer(path);
        for (const Matrix &m : matrices)
            write_next(writer, m);
        flush(writer);
    }

    std::vector<Matrix> read_mat_file(const std::filesystem::path &path)
    {
        return MatFilePolicy::read(path);
    }
    void write_mat_file(const std::filesystem::path &path, std::span<const Matrix> matrices)
    {
        MatFilePolicy::write(path, matrices);
    }
    std::vector<Matrix> read_txt_file(const std::filesystem::path &path)
    {
        return TxtFilePolicy::read(path);
    }
    void write_txt_file(const std::filesystem::path &path, std::span<const Matrix> matrices)
    {
        TxtFilePolicy::write(path, matrices);
    }
} // namespace jacobi::svd::io
