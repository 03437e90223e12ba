// MatDispatchReader / PinnedHostTaskBuffer with pinned memory (needs the HIP
// runtime and a GPU for hipHostMalloc).  Capacity growth and region views
// (reference src/infrastructure/pinned_host_task_buffer.cu:28-96,
// io_detail.cuh grow_capacity: 64 KiB, then doubling), moves, record reading
// with a workspace, the two-step reading into one staging buffer, and decode
// against the written matrices.
#include "jacobi/svd/io/io.cuh"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <string>
#include <vector>

namespace fs = std::filesystem;
using namespace jacobi::svd::io;

namespace
{
    int failures = 0;
    void check(bool ok, const std::string &what)
    {
        if (!ok)
        {
            std::cerr << "FAILED: " << what << '\n';
            ++failures;
        }
    }

    std::vector<Matrix> sample()
    {
        std::vector<Matrix> ms;
        ms.push_back({2, 3, {1.0, -2.5, 3.25, 1e-300, -0.0, 7.0}});
        ms.push_back({1, 1, {std::acos(-1.0)}});
        Matrix big{300, 40, {}};
        for (std::size_t i = 0; i < 300 * 40; ++i)
            big.values.push_back(std::sin(double(i)) * std::pow(10.0, double(int(i % 13) - 6)));
        ms.push_back(big); // 96000 payload bytes: more than the first 64 KiB
        return ms;
    }

    bool same(const Matrix &a, const Matrix &b)
    {
        return a.rows == b.rows && a.columns == b.columns &&
               std::memcmp(a.values.data(), b.values.data(), a.values.size() * sizeof(double)) == 0 &&
               a.values.size() == b.values.size();
    }
} // namespace

int main()
{
    {
        PinnedHostTaskBuffer b;
        b.reserve(100, 28);
        check(b.capacity_bytes() == 65536 && b.input_size_bytes() == 100 && b.workspace_size_bytes() == 28,
              "first reserve: 64 KiB");
        check(b.input_bytes().size() == 100 && b.workspace_bytes().size() == 28 &&
                  b.workspace_bytes().data() == b.input_bytes().data() + 100,
              "workspace follows the input region");
        hipPointerAttribute_t attr{};
        check(hipPointerGetAttributes(&attr, b.input_bytes().data()) == hipSuccess && attr.type == hipMemoryTypeHost,
              "the buffer is HIP pinned host memory");
        std::memset(b.mutable_input_bytes().data(), 0x5a, 100);
        std::memset(b.mutable_workspace_bytes().data(), 0xa5, 28);
        check(b.input_bytes()[99] == std::byte{0x5a} && b.workspace_bytes()[0] == std::byte{0xa5}, "regions hold data");
        b.reserve(70000, 0);
        check(b.capacity_bytes() == 131072, "growth doubles 64 KiB -> 128 KiB");
        const std::byte *p = b.input_bytes().data();
        b.reserve(10, 10);
        check(b.capacity_bytes() == 131072 && b.input_bytes().data() == p && b.input_size_bytes() == 10 &&
                  b.workspace_size_bytes() == 10,
              "a smaller reserve keeps the allocation");
        b.reserve((std::size_t(1) << 20) + 1, 0);
        check(b.capacity_bytes() == (std::size_t(1) << 21), "growth to the next power-of-two multiple");
        PinnedHostTaskBuffer c(std::move(b));
        check(c.capacity_bytes() == (std::size_t(1) << 21) && b.capacity_bytes() == 0 && b.input_bytes().empty(),
              "move construction");
        PinnedHostTaskBuffer d;
        d.reserve(8, 0);
        d = std::move(c);
        check(d.capacity_bytes() == (std::size_t(1) << 21) && c.capacity_bytes() == 0, "move assignment");
    }

    const fs::path mat = fs::temp_directory_path() / "jsvd-dispatch-test.mat";
    const std::vector<Matrix> ms = sample();
    write_mat_file(mat, ms);
    {
        MatDispatchReader r(mat);
        MatDispatchTask t;
        std::size_t k = 0;
        while (r.read_next(t, 64))
        {
            const std::size_t bytes = ms[k].values.size() * sizeof(double);
            check(t.sequence_index == k && t.rows == ms[k].rows && t.columns == ms[k].columns,
                  "task " + std::to_string(k) + " header");
            check(t.buffer.input_size_bytes() == bytes && t.buffer.workspace_size_bytes() == 64,
                  "task " + std::to_string(k) + " regions");
            check(same(decode_dispatch_task_matrix(t), ms[k]), "task " + std::to_string(k) + " decodes exactly");
            // The input region is the payload as stored: network byte order.
            std::uint64_t first = 0;
            std::memcpy(&first, t.buffer.input_bytes().data(), 8);
            std::uint64_t want = 0;
            std::memcpy(&want, ms[k].values.data(), 8);
            check(first == __builtin_bswap64(want), "task " + std::to_string(k) + " keeps network byte order");
            ++k;
        }
        check(k == ms.size(), "every record read");
        check(t.buffer.capacity_bytes() == 131072, "the reused task grew once, to 128 KiB");
    }
    // Two-step reading: payloads back to back in one staging buffer.
    {
        MatDispatchReader r(mat);
        PinnedHostTaskBuffer stage;
        stage.reserve(8 * (6 + 1 + 300 * 40), 0);
        std::size_t rows = 0, cols = 0, off = 0, k = 0;
        bool ok = true;
        while (r.read_header(rows, cols))
        {
            const std::size_t bytes = rows * cols * 8;
            ok = ok && r.read_payload(stage.mutable_input_bytes().subspan(off, bytes)) == k;
            ok = ok && rows == ms[k].rows && cols == ms[k].columns;
            for (std::size_t i = 0; i < rows * cols; ++i)
            {
                std::uint64_t x = 0, y = 0;
                std::memcpy(&x, stage.input_bytes().data() + off + 8 * i, 8);
                std::memcpy(&y, &ms[k].values[i], 8);
                ok = ok && x == __builtin_bswap64(y);
            }
            off += bytes;
            ++k;
        }
        check(ok && k == 3 && off == stage.input_size_bytes(), "two-step reading into one staging buffer");
    }
    fs::remove(mat);
    if (failures)
    {
        std::cerr << failures << " dispatch check(s) failed\n";
        return 1;
    }
    std::cout << "dispatch_tests passed\n";
    return 0;
}
