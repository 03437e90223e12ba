// C++ matrix-stream codec tests (host only).  Round trips of .mat and .txt,
// the reference's error cases (truncated header / payload, ragged text rows,
// bad tokens), the dispatch reader's paths that reserve no pinned memory
// (errors, end of file, empty records), and a dump of an externally written
// .mat file (argv[1]) that the Python test compares with what it wrote.
#include "jacobi/svd/io/io.cuh"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <span>
#include <stdexcept>
#include <vector>

namespace fs = std::filesystem;
using jacobi::svd::io::Matrix;

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

    template <typename F>
    bool throws(F f, const std::string &needle)
    {
        try
        {
            f();
        }
        catch (const std::exception &e)
        {
            return std::string(e.what()).find(needle) != std::string::npos;
        }
        return false;
    }

    std::vector<Matrix> sample()
    {
        std::vector<Matrix> ms;
        ms.push_back({2, 3, {1.0, -2.5, 3.25, 1e-300, -0.0, 7.0}});
        ms.push_back({1, 1, {std::acos(-1.0)}});
        Matrix big{40, 17, {}};
        for (std::size_t i = 0; i < 40 * 17; ++i)
            big.values.push_back(std::sin(double(i)) * std::pow(10.0, double(int(i % 13) - 6)));
        ms.push_back(big);
        return ms;
    }

    bool same(const std::vector<Matrix> &a, const std::vector<Matrix> &b)
    {
        if (a.size() != b.size())
            return false;
        for (std::size_t k = 0; k < a.size(); ++k)
            if (a[k].rows != b[k].rows || a[k].columns != b[k].columns || a[k].values != b[k].values)
                return false;
        return true;
    }
    void write_bytes(const fs::path &p, const std::vector<unsigned char> &bytes)
    {
        std::ofstream out(p, std::ios::binary | std::ios::trunc);
        out.write(reinterpret_cast<const char *>(bytes.data()), static_cast<std::streamsize>(bytes.size()));
    }

    // Big-endian record header.
    void put_header(std::vector<unsigned char> &b, std::uint64_t rows, std::uint64_t cols)
    {
        for (std::uint64_t v : {rows, cols})
            for (int k = 7; k >= 0; --k)
                b.push_back(static_cast<unsigned char>(v >> (8 * k)));
    }

    // MatDispatchReader / PinnedHostTaskBuffer without a device: every case
    // here fails or ends before pinned memory would be reserved (reference
    // src/infrastructure/mat_dispatch_reader.cu:17-84 messages).
    void dispatch_checks(const fs::path &mat)
    {
        using namespace jacobi::svd::io;
        check(throws([] { MatDispatchReader r("/nonexistent/jsvd.mat"); },
                     "Failed to open *.mat file for dispatch reading."),
              "dispatch: open failure");
        {
            PinnedHostTaskBuffer b;
            b.reserve(0, 0);
            check(b.capacity_bytes() == 0 && b.input_bytes().empty() && b.workspace_bytes().empty() &&
                      b.input_size_bytes() == 0 && b.workspace_size_bytes() == 0,
                  "dispatch: empty buffer reserves nothing");
            PinnedHostTaskBuffer moved(std::move(b));
            check(moved.capacity_bytes() == 0, "dispatch: move of an empty buffer");
        }
        write_bytes(mat, {});
        {
            MatDispatchReader r(mat);
            MatDispatchTask t;
            t.sequence_index = 7;
            check(!r.read_next(t) && t.sequence_index == 7, "dispatch: empty file is end of file");
        }
        // Records with no payload: 0 x 3, then 2 x 0.
        {
            std::vector<unsigned char> b;
            put_header(b, 0, 3);
            put_header(b, 2, 0);
            write_bytes(mat, b);
            MatDispatchReader r(mat);
            MatDispatchTask t;
            bool ok = r.read_next(t) && t.sequence_index == 0 && t.rows == 0 && t.columns == 3 &&
                      t.buffer.input_bytes().empty() && t.buffer.capacity_bytes() == 0;
            const Matrix d = decode_dispatch_task_matrix(t);
            ok = ok && d.rows == 0 && d.columns == 3 && d.values.empty();
            ok = ok && r.read_next(t) && t.sequence_index == 1 && t.rows == 2 && t.columns == 0;
            check(ok && !r.read_next(t), "dispatch: empty records and their sequence indices");
        }
        fs::resize_file(mat, 10);
        check(throws([&] { MatDispatchReader r(mat); MatDispatchTask t; (void)r.read_next(t); },
                     "Truncated *.mat header while dispatch reading."),
              "dispatch: truncated header");
        // 2 x 3 header with three of its six doubles: found before the
        // buffer is reserved, so the task keeps no allocation.
        {
            std::vector<unsigned char> b;
            put_header(b, 2, 3);
            b.resize(b.size() + 24, 0);
            write_bytes(mat, b);
            MatDispatchReader r(mat);
            MatDispatchTask t;
            check(throws([&] { (void)r.read_next(t); }, "Truncated *.mat payload while dispatch reading.") &&
                      t.buffer.capacity_bytes() == 0,
                  "dispatch: truncated payload, nothing reserved");
        }
        {
            std::vector<unsigned char> b;
            put_header(b, std::uint64_t(1) << 40, std::uint64_t(1) << 20);
            b.resize(b.size() + 24, 0);
            write_bytes(mat, b);
            MatDispatchReader r(mat);
            std::size_t rows = 0, cols = 0;
            check(throws([&] { (void)r.read_header(rows, cols); }, "Truncated *.mat payload while dispatch reading."),
                  "dispatch: oversized header");
        }
        // The two-step API: order and destination size are enforced.
        {
            std::vector<unsigned char> b;
            put_header(b, 1, 2);
            b.resize(b.size() + 16, 0);
            write_bytes(mat, b);
            MatDispatchReader r(mat);
            std::byte dst[16];
            check(throws([&] { (void)r.read_payload(std::span<std::byte>(dst, 16)); }, "No *.mat header is pending."),
                  "dispatch: payload before header");
            std::size_t rows = 0, cols = 0;
            check(r.read_header(rows, cols) && rows == 1 && cols == 2, "dispatch: read_header");
            check(throws([&] { (void)r.read_header(rows, cols); }, "has not been read"), "dispatch: header twice");
            check(throws([&] { (void)r.read_payload(std::span<std::byte>(dst, 8)); }, "destination size"),
                  "dispatch: destination size");
            check(r.read_payload(std::span<std::byte>(dst, 16)) == 0 && !r.read_header(rows, cols),
                  "dispatch: payload then end of file");
        }
        {
            MatDispatchTask t;
            t.rows = 1;
            t.columns = 1;
            check(throws([&] { (void)decode_dispatch_task_matrix(t); },
                         "Dispatch payload size does not match rows * columns."),
                  "dispatch: decode size check");
        }
    }
} // namespace

int main(int argc, char **argv)
{
    if (argc == 2)
    {
        // Dump mode: rows cols then values, one matrix per line, %.17g.
        for (const Matrix &m : jacobi::svd::io::read_mat_file(argv[1]))
        {
            std::printf("%zu %zu", m.rows, m.columns);
            for (double x : m.values)
                std::printf(" %.17g", x);
            std::printf("\n");
        }
        return 0;
    }
    const fs::path dir = fs::temp_directory_path();
    const std::vector<Matrix> ms = sample();
    const fs::path mat = dir / "jsvd-io-test.mat", txt = dir / "jsvd-io-test.txt";
    jacobi::svd::io::write_mat_file(mat, ms);
    check(same(jacobi::svd::io::read_mat_file(mat), ms), ".mat round trip is exact");
    check(fs::file_size(mat) == 3 * 16 + 8 * (6 + 1 + 40 * 17), ".mat size = headers + payloads");
    jacobi::svd::io::write_txt_file(txt, ms);
    check(same(jacobi::svd::io::read_txt_file(txt), ms), ".txt round trip is exact (max_digits10)");
    {
        jacobi::svd::io::MatrixInputStream<jacobi::svd::io::MatFilePolicy> in(mat);
        Matrix m;
        int count = 0;
        while (in.read_one(m))
            ++count;
        check(count == 3 && in.eof(), "MatrixInputStream reads every record");
    }
    // Truncations.
    fs::resize_file(mat, 10);
    check(throws([&] { (void)jacobi::svd::io::read_mat_file(mat); }, "Truncated *.mat header."), "truncated header");
    jacobi::svd::io::write_mat_file(mat, ms);
    fs::resize_file(mat, 16 + 8 * 3);
    check(throws([&] { (void)jacobi::svd::io::read_mat_file(mat); }, "Truncated *.mat payload."), "truncated payload");
    // A header that claims 2^40 x 2^20 doubles (8 EiB) in a 40-byte file is
    // a truncation, found before anything is allocated for the payload.
    {
        const unsigned char header[16] = {0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 16, 0, 0};
        std::ofstream out(mat, std::ios::binary | std::ios::trunc);
        out.write(reinterpret_cast<const char *>(header), 16);
        out.write("0123456789abcdef01234567", 24);
    }
    check(throws([&] { (void)jacobi::svd::io::read_mat_file(mat); }, "Truncated *.mat payload."),
          "oversized header");
    // Text errors.
    std::ofstream(txt) << "1 2\n3\n";
    check(throws([&] { (void)jacobi::svd::io::read_txt_file(txt); }, "Inconsistent column count"), "ragged rows");
    std::ofstream(txt) << "1 x\n";
    check(throws([&] { (void)jacobi::svd::io::read_txt_file(txt); }, "Invalid numeric token"), "bad token");
    std::ofstream(txt) << "\n\n1 2\n3 4\n\n\n5\n";
    {
        const auto r = jacobi::svd::io::read_txt_file(txt);
        check(r.size() == 2 && r[0].rows == 2 && r[1].columns == 1, "blank lines separate blocks");
    }
    check(throws([&] { jacobi::svd::io::write_mat_file(mat, std::vector<Matrix>{{2, 2, {1.0}}}); },
                 "payload size"),
          "layout check on write");
    dispatch_checks(mat);
    fs::remove(mat);
    fs::remove(txt);
    if (failures)
    {
        std::cerr << failures << " IO check(s) failed\n";
        return 1;
    }
    std::cout << "io_tests passed\n";
    return 0;
}
