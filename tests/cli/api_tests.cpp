// C++ API tests (needs a GPU): jacobi::svd::one_sided_jacobi_svd with the
// reference's exceptions and messages, DeviceMatrix, and the autotune
// report.  Numeric results are checked against closed forms and invariants.
#include "jacobi/svd/domain/kernels.cuh"

#include <cmath>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

using namespace jacobi::svd;

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

    template <typename F>
    bool invalid(F f, const std::string &message)
    {
        try
        {
            f();
        }
        catch (const std::invalid_argument &e)
        {
            return message == e.what();
        }
        return false;
    }

    // Deterministic N(0,1)-ish entries (sum of uniforms from an LCG).
    std::vector<double> matrix(std::size_t m, std::size_t n, std::uint64_t seed)
    {
        std::vector<double> a(m * n);
        for (double &x : a)
        {
            double s = 0.0;
            for (int k = 0; k < 12; ++k)
            {
                seed = seed * 6364136223846793005ull + 1442695040888963407ull;
                s += double(seed >> 11) * (1.0 / 9007199254740992.0);
            }
            x = s - 6.0;
        }
        return a;
    }
} // namespace

int run_checks();

int main()
{
    try
    {
        return run_checks();
    }
    catch (const std::exception &e)
    {
        std::cerr << "FAILED: unexpected exception: " << e.what() << '\n';
        return 1;
    }
}

int run_checks()
{
    const std::vector<double> a23(6, 1.0);
    check(invalid([&] { (void)one_sided_jacobi_svd(a23, 2, 3); },
                  "One-sided Jacobi SVD requires rows >= columns in this implementation."),
          "rows < columns");
    check(invalid([&] { (void)one_sided_jacobi_svd(a23, 0, 3); }, "Input matrix shape must be non-zero."), "zero rows");
    check(invalid([&] { (void)one_sided_jacobi_svd(a23, 4, 2); }, "Input buffer size mismatch."), "size mismatch");
    JacobiSvdConfig bad;
    bad.epsilon = 0.0;
    check(invalid([&] { (void)one_sided_jacobi_svd(a23, 3, 2, bad); }, "epsilon must be positive."), "epsilon");
    bad = JacobiSvdConfig{};
    bad.max_sweeps = 0;
    check(invalid([&] { (void)one_sided_jacobi_svd(a23, 3, 2, bad); }, "max_sweeps must be positive."), "sweeps");

    // Diagonal input: no rotation, one (verifying) sweep, sigma unsorted.
    const std::vector<double> d = {2, 0, 0, 0, -5, 0, 0, 0, 1, 0, 0, 0};
    const JacobiSvdResult rd = one_sided_jacobi_svd(d, 4, 3);
    check(rd.sweeps == 1 && rd.sigma == std::vector<double>({2, 5, 1}), "diagonal input");
    check(rd.u.size() == 12 && rd.v.size() == 9 && rd.rows == 4 && rd.columns == 3, "result shapes");

    // Random 70 x 40 (a shape for the LDS classes) and 300 x 150 (the
    // persistent kernel): reconstruction and orthogonality.
    for (auto [m, n] : {std::pair<std::size_t, std::size_t>{70, 40}, {300, 150}})
    {
        const std::vector<double> a = matrix(m, n, 42 + m);
        JacobiSvdConfig cfg;
        cfg.epsilon = 1e-13;
        cfg.max_sweeps = 30;
        const JacobiSvdResult r = one_sided_jacobi_svd(a, m, n, cfg);
        double err = 0.0, norm = 0.0, orth = 0.0;
        for (std::size_t i = 0; i < m; ++i)
            for (std::size_t j = 0; j < n; ++j)
            {
                double x = 0.0;
                for (std::size_t k = 0; k < n; ++k)
                    x += r.u[i * n + k] * r.sigma[k] * r.v[j * n + k];
                err += (x - a[i * n + j]) * (x - a[i * n + j]);
                norm += a[i * n + j] * a[i * n + j];
            }
        for (std::size_t p = 0; p < n; ++p)
            for (std::size_t q = 0; q < n; ++q)
            {
                double x = 0.0;
                for (std::size_t k = 0; k < n; ++k)
                    x += r.v[k * n + p] * r.v[k * n + q];
                orth += (x - (p == q)) * (x - (p == q));
            }
        check(std::sqrt(err / norm) < 1e-13, "residual " + std::to_string(m) + "x" + std::to_string(n));
        check(std::sqrt(orth) < 1e-12, "V orthogonality " + std::to_string(m) + "x" + std::to_string(n));
        check(r.sweeps > 1 && r.sweeps <= 30, "sweeps in range");
    }

    // DeviceMatrix: copies and move semantics.
    {
        DeviceMatrix dm(3, 2);
        check(dm.rows() == 3 && dm.columns() == 2 && dm.size() == 6 && dm.bytes() == 48, "DeviceMatrix shape");
        const std::vector<double> h = {1, 2, 3, 4, 5, 6};
        dm.copy_from_host(h);
        DeviceMatrix moved(std::move(dm));
        check(moved.copy_to_host() == h && dm.data() == nullptr && dm.size() == 0, "DeviceMatrix move");
        DeviceMatrix other;
        other = std::move(moved);
        check(other.copy_to_host() == h, "DeviceMatrix move assignment");
        bool threw = false;
        try
        {
            other.copy_from_host(std::vector<double>(5));
        }
        catch (const std::invalid_argument &)
        {
            threw = true;
        }
        check(threw, "DeviceMatrix size check");
        // The reference's messages (src/domain/device_matrix.cu:50-52, 99-101).
        check(invalid([&] { other.copy_from_host(std::vector<double>(5)); },
                      "Host data size does not match device matrix shape."),
              "copy_from_host message");
        check(invalid([&] { other.reset(0, 4); }, "DeviceMatrix dimensions must be positive."), "reset(0, n)");
        check(invalid([&] { DeviceMatrix z(3, 0); }, "DeviceMatrix dimensions must be positive."), "ctor (m, 0)");
        // A rejected reset leaves the old allocation in place.
        check(other.rows() == 3 && other.copy_to_host() == h, "reset(0, n) keeps the matrix");
    }

    // A device failure surfaces as CudaError with the reference's message
    // shape (src/domain/cuda_check.cuh:26-28): 2^20 x 2^20 doubles = 8 TiB
    // is far beyond one GPU's HBM, so hipMalloc reports out of memory.
    {
        std::string what;
        try
        {
            DeviceMatrix huge(std::size_t(1) << 20, std::size_t(1) << 20);
        }
        catch (const CudaError &e)
        {
            what = e.what();
        }
        check(what.rfind("HIP call failed: hipMalloc(", 0) == 0 && what.find(" @ ") != std::string::npos &&
                  what.find(", code=") != std::string::npos && what.find(", message=") != std::string::npos,
              "DeviceMatrix out-of-memory -> CudaError: '" + what + "'");
        // The device is still usable afterwards.
        DeviceMatrix after(2, 2);
        after.copy_from_host(std::vector<double>{1, 2, 3, 4});
        check(after.copy_to_host() == std::vector<double>({1, 2, 3, 4}), "device usable after OOM");
    }

    // Stop reason (appended JacobiSvdResult::converged): one sweep cannot
    // orthogonalise a random 30 x 30; thirty can.
    {
        const std::vector<double> a = matrix(30, 30, 7);
        JacobiSvdConfig cfg;
        cfg.epsilon = 1e-13;
        cfg.max_sweeps = 1;
        const JacobiSvdResult r1 = one_sided_jacobi_svd(a, 30, 30, cfg);
        check(!r1.converged && r1.sweeps == 1, "max_sweeps = 1 reported as not converged");
        cfg.max_sweeps = 30;
        const JacobiSvdResult r30 = one_sided_jacobi_svd(a, 30, 30, cfg);
        check(r30.converged && r30.sweeps < 30, "converged solve reported as converged");
        check(rd.converged, "diagonal input converged");
    }

    const LayoutTransposeAutoTuneReport rep = auto_tune_layout_transpose_threshold();
    check(rep.executed && rep.sample_count > 0 && rep.estimated_best_speedup >= 1.0, "autotune report");

    if (failures)
    {
        std::cerr << failures << " API check(s) failed\n";
        return 1;
    }
    std::cout << "api_tests passed\n";
    return 0;
}
