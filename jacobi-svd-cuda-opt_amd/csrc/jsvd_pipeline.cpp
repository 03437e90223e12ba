// jsvd_pipeline.cpp -- the batched pipeline (reference src/application/pipeline.cu:25-139,
// src/application/output_stage.cuh:20-276, src/application/kernel_stage.cuh:30-58).
//
// The reference hands every matrix to a host thread pool and makes one
// synchronous solver call per matrix.  Here matrices are gathered into
// device batches: a batch is packed into pinned staging, copied to HBM,
// solved by one work-list plan, copied back and written out.  Batches rotate
// over `max_queued_results` slots, each with its own HIP stream, so reading
// and packing the next batch overlaps the device work of the previous ones;
// the oldest slot is always drained first, which keeps the output in input
// order (U, Sigma as 1 x n, V for every testcase, as the reference writes).
// A *.mat input is read with MatDispatchReader straight into the slot's
// pinned staging, payloads back to back in network byte order as they are
// stored in the file; one device pass swaps the bytes after the copy to
// HBM, so the host never decodes a value.
// A slot whose next batch has the same shapes as its last one replays a
// captured hipGraph of copy-in, solve and copy-out instead of re-issuing the
// launches.
#include "jacobi/svd/application/pipeline.cuh"
#include "jacobi/svd/io/io.cuh"
#include "jsvd_c.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace jsvd
{
    hipError_t launch_bswap64(void *p, int64_t count, hipStream_t stream);
}

namespace jacobi::svd::pipeline
{
    namespace
    {
        constexpr std::size_t kBatchMatrices = 4096;
        constexpr std::size_t kBatchElements = std::size_t(32) << 20; // 256 MiB of FP64 input

        void hip_check(hipError_t e, const char *what)
        {
            if (e == hipSuccess)
                return;
            throw CudaError((std::string("HIP call failed: ") + what + ", code=" + std::to_string(int(e)) +
                             ", message=" + hipGetErrorString(e))
                                .c_str());
        }

        void rethrow(jsvd_status status)
        {
            if (status == JSVD_OK)
                return;
            const std::string msg = jsvd_last_error();
            if (status == JSVD_EINVAL)
                throw std::invalid_argument(msg);
            throw CudaError(msg.c_str());
        }

        MatrixFileFormat resolve(MatrixFileFormat requested, const std::filesystem::path &path)
        {
            if (requested != MatrixFileFormat::auto_detect)
                return requested;
            std::string ext = path.extension().string();
            std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return std::tolower(c); });
            if (ext == ".mat")
                return MatrixFileFormat::mat;
            if (ext == ".txt")
                return MatrixFileFormat::txt;
            throw std::invalid_argument("Cannot auto-detect file format for path: " + path.string());
        }

        jsvd_config to_c(const JacobiSvdConfig &c)
        {
            jsvd_config out;
            jsvd_config_default(&out);
            out.epsilon = c.epsilon;
            out.max_sweeps = c.max_sweeps;
            out.threads_per_block = c.threads_per_block;
            out.layout_mode = c.layout_transpose_mode == LayoutTransposeMode::force_enable    ? JSVD_LAYOUT_ON
                              : c.layout_transpose_mode == LayoutTransposeMode::force_disable ? JSVD_LAYOUT_OFF
                                                                                               : JSVD_LAYOUT_AUTO;
            out.layout_min_columns = c.layout_transpose_min_columns;
            out.layout_min_elements = c.layout_transpose_min_elements;
            out.layout_auto_tune = c.layout_transpose_auto_tune ? 1 : 0;
            out.layout_bench_reps = c.layout_transpose_benchmark_repetitions;
            out.layout_bench_sweeps = c.layout_transpose_benchmark_sweeps;
            out.precision = c.precision == JacobiPrecision::mixed ? JSVD_PRECISION_MIXED : JSVD_PRECISION_FP64;
            out.block_width = c.block_width;
            out.refine_sweeps = c.refine_sweeps;
            return out;
        }

        template <typename T>
        struct Pinned
        {
            T *ptr = nullptr;
            std::size_t count = 0;
            ~Pinned()
            {
                if (ptr)
                    (void)hipHostFree(ptr);
            }
            bool reserve(std::size_t n)
            {
                if (n <= count)
                    return false;
                if (ptr)
                    hip_check(hipHostFree(ptr), "hipHostFree");
                ptr = nullptr;
                hip_check(hipHostMalloc(reinterpret_cast<void **>(&ptr), std::max<std::size_t>(n, 1) * sizeof(T), 0),
                          "hipHostMalloc");
                count = n;
                return true;
            }
            // Capacity for at least n elements, keeping the first `keep`;
            // grows geometrically.  Returns true when the buffer moved.
            bool grow_keep(std::size_t n, std::size_t keep)
            {
                if (n <= count)
                    return false;
                const std::size_t cap = std::max(n, 2 * count);
                T *p = nullptr;
                hip_check(hipHostMalloc(reinterpret_cast<void **>(&p), cap * sizeof(T), 0), "hipHostMalloc");
                if (ptr)
                {
                    std::memcpy(p, ptr, keep * sizeof(T));
                    hip_check(hipHostFree(ptr), "hipHostFree");
                }
                ptr = p;
                count = cap;
                return true;
            }
        };

        template <typename T>
        struct Device
        {
            T *ptr = nullptr;
            std::size_t count = 0;
            ~Device()
            {
                if (ptr)
                    (void)hipFree(ptr);
            }
            bool reserve(std::size_t n)
            {
                if (n <= count)
                    return false;
                if (ptr)
                    hip_check(hipFree(ptr), "hipFree");
                ptr = nullptr;
                hip_check(hipMalloc(reinterpret_cast<void **>(&ptr), std::max<std::size_t>(n, 1) * sizeof(T)),
                          "hipMalloc");
                count = n;
                return true;
            }
        };

        // One in-flight device batch.
        struct Slot
        {
            hipStream_t stream = nullptr;
            hipEvent_t done = nullptr;
            jsvd_plan *plan = nullptr;
            hipGraphExec_t graph = nullptr;
            std::vector<int64_t> m, n;
            std::size_t a_elems = 0, v_elems = 0, s_elems = 0;
            bool capturable = false;
            bool metrics = false; // error metrics requested (PipelineConfig::report_errors)
            bool raw_be = false;  // staging holds *.mat payloads in network byte order
            // The batch being staged: shapes and elements so far.
            std::vector<int64_t> bm, bn;
            std::size_t staged = 0;
            bool in_moved = false; // h_in reallocated since the last submit
            Pinned<double> h_in, h_u, h_s, h_v, h_met;
            Pinned<int32_t> h_sw;  // sweeps, then the stop reasons
            Device<double> d_a, d_u, d_s, d_v, d_met;
            Device<int32_t> d_sw;
            bool busy = false;

            Slot()
            {
                hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
                hip_check(hipEventCreateWithFlags(&done, hipEventDisableTiming), "hipEventCreateWithFlags");
            }
            ~Slot()
            {
                if (stream)
                    (void)hipStreamSynchronize(stream);
                drop_plan();
                if (done)
                    (void)hipEventDestroy(done);
                if (stream)
                    (void)hipStreamDestroy(stream);
            }
            void drop_plan()
            {
                if (graph)
                    (void)hipGraphExecDestroy(graph);
                graph = nullptr;
                if (plan)
                    jsvd_plan_destroy(plan);
                plan = nullptr;
            }

            void begin()
            {
                bm.clear();
                bn.clear();
                staged = 0;
            }

            // Room for one more testcase of rows x cols in the staging;
            // returns where its payload goes.
            std::byte *stage(std::size_t rows, std::size_t cols)
            {
                const std::size_t elems = rows * cols;
                in_moved |= h_in.grow_keep(staged + elems, staged);
                double *at = h_in.ptr + staged;
                staged += elems;
                bm.push_back(static_cast<int64_t>(rows));
                bn.push_back(static_cast<int64_t>(cols));
                return reinterpret_cast<std::byte *>(at);
            }

            void issue()
            {
                hip_check(hipMemcpyAsync(d_a.ptr, h_in.ptr, a_elems * sizeof(double), hipMemcpyHostToDevice, stream),
                          "hipMemcpyAsync(H2D A)");
                if (raw_be)
                    hip_check(jsvd::launch_bswap64(d_a.ptr, static_cast<int64_t>(a_elems), stream),
                              "bswap64 launch");
                const std::size_t count = m.size();
                rethrow(jsvd_plan_execute_status_f64(plan, d_a.ptr, d_u.ptr, d_s.ptr, d_v.ptr, d_sw.ptr,
                                                     d_sw.ptr + count, stream));
                if (metrics)
                    rethrow(jsvd_plan_metrics_f64(plan, d_a.ptr, d_u.ptr, d_s.ptr, d_v.ptr, d_met.ptr, stream));
                hip_check(hipMemcpyAsync(h_u.ptr, d_u.ptr, a_elems * sizeof(double), hipMemcpyDeviceToHost, stream),
                          "hipMemcpyAsync(D2H U)");
                hip_check(hipMemcpyAsync(h_s.ptr, d_s.ptr, s_elems * sizeof(double), hipMemcpyDeviceToHost, stream),
                          "hipMemcpyAsync(D2H sigma)");
                hip_check(hipMemcpyAsync(h_v.ptr, d_v.ptr, v_elems * sizeof(double), hipMemcpyDeviceToHost, stream),
                          "hipMemcpyAsync(D2H V)");
                hip_check(hipMemcpyAsync(h_sw.ptr, d_sw.ptr, 2 * count * sizeof(int32_t), hipMemcpyDeviceToHost,
                                         stream),
                          "hipMemcpyAsync(D2H sweeps)");
                if (metrics)
                    hip_check(hipMemcpyAsync(h_met.ptr, d_met.ptr, 3 * count * sizeof(double), hipMemcpyDeviceToHost,
                                             stream),
                              "hipMemcpyAsync(D2H metrics)");
            }

            // Solves the staged batch (bm, bn, h_in[0 .. staged)).
            void submit(const jsvd_config &cfg)
            {
                if (!plan || bm != m || bn != n)
                {
                    drop_plan();
                    m = bm;
                    n = bn;
                    rethrow(jsvd_plan_create(static_cast<int64_t>(m.size()), m.data(), n.data(), &cfg, &plan));
                    jsvd_plan_info info{};
                    rethrow(jsvd_plan_get_info(plan, &info));
                    a_elems = static_cast<std::size_t>(info.a_elements);
                    v_elems = static_cast<std::size_t>(info.v_elements);
                    s_elems = static_cast<std::size_t>(info.sigma_elements);
                    // A metrics run stays out of graphs (its first call on a
                    // plan uploads the tile list synchronously); every solve
                    // path, the blocked one included, keeps its decisions on
                    // the device and can be captured.
                    capturable = !metrics;
                }
                if (a_elems != staged)
                    throw std::logic_error("Pipeline staging does not match the plan's packed size.");
                bool moved = std::exchange(in_moved, false);
                moved |= h_u.reserve(a_elems);
                moved |= h_s.reserve(s_elems);
                moved |= h_v.reserve(v_elems);
                moved |= h_sw.reserve(2 * m.size());
                moved |= d_a.reserve(a_elems);
                moved |= d_u.reserve(a_elems);
                moved |= d_s.reserve(s_elems);
                moved |= d_v.reserve(v_elems);
                moved |= d_sw.reserve(2 * m.size());
                if (metrics)
                {
                    moved |= h_met.reserve(3 * m.size());
                    moved |= d_met.reserve(3 * m.size());
                }
                if (moved && graph)
                {
                    (void)hipGraphExecDestroy(graph);
                    graph = nullptr;
                }
                if (capturable)
                {
                    if (!graph)
                    {
                        hipGraph_t g = nullptr;
                        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
                        try
                        {
                            issue();
                        }
                        catch (...)
                        {
                            (void)hipStreamEndCapture(stream, &g);
                            if (g)
                                (void)hipGraphDestroy(g);
                            throw;
                        }
                        hip_check(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
                        const hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
                        (void)hipGraphDestroy(g);
                        hip_check(e, "hipGraphInstantiate");
                    }
                    hip_check(hipGraphLaunch(graph, stream), "hipGraphLaunch");
                }
                else
                    issue();
                hip_check(hipEventRecord(done, stream), "hipEventRecord");
                busy = true;
            }

            // Writes the batch's results in input order and appends its
            // per-testcase sweeps, stop reasons and (when requested) error
            // metrics to the report.  `reference` reads the next testcase's
            // reference sigma (1 x n), or is empty.
            void drain(const std::function<void(const io::Matrix &)> &write, PipelineReport &report,
                       const std::function<bool(io::Matrix &)> &reference)
            {
                hip_check(hipEventSynchronize(done), "hipEventSynchronize");
                const std::size_t count = m.size();
                for (std::size_t k = 0; k < count; ++k)
                    if (h_sw.ptr[count + k] == JSVD_STOP_BARRIER_TIMEOUT)
                        throw CudaError(("Testcase[" + std::to_string(report.matrix_sweeps.size() + k) +
                                         "]: a device barrier timed out (the workgroups of a barrier kernel were "
                                         "not all resident); the batch's results are invalid.")
                                            .c_str());
                std::size_t a = 0, v = 0, s = 0;
                io::Matrix ref;
                for (std::size_t k = 0; k < count; ++k)
                {
                    const std::size_t rows = static_cast<std::size_t>(m[k]), cols = static_cast<std::size_t>(n[k]);
                    io::Matrix out;
                    out.rows = rows;
                    out.columns = cols;
                    out.values.assign(h_u.ptr + a, h_u.ptr + a + rows * cols);
                    write(out);
                    out.rows = 1;
                    out.values.assign(h_s.ptr + s, h_s.ptr + s + cols);
                    write(out);
                    if (metrics)
                    {
                        PipelineReport::MatrixErrors e;
                        e.orth_u = h_met.ptr[3 * k];
                        e.orth_v = h_met.ptr[3 * k + 1];
                        e.residual = h_met.ptr[3 * k + 2];
                        if (reference)
                        {
                            const std::size_t index = report.matrix_sweeps.size();
                            if (!reference(ref))
                                throw std::invalid_argument("Reference sigma stream ends before testcase[" +
                                                            std::to_string(index) + "].");
                            if (ref.values.size() != cols)
                                throw std::invalid_argument("Reference sigma of testcase[" + std::to_string(index) +
                                                            "] does not have " + std::to_string(cols) + " values.");
                            std::vector<double> got(out.values), want(ref.values);
                            std::sort(got.begin(), got.end(), std::greater<double>());
                            std::sort(want.begin(), want.end(), std::greater<double>());
                            double worst = 0.0, comp = 0.0;
                            for (std::size_t j = 0; j < cols; ++j)
                            {
                                const double d = std::abs(got[j] - want[j]);
                                worst = std::max(worst, d);
                                if (want[j] > 0.0)
                                    comp = std::max(comp, d / want[j]);
                            }
                            e.sigma_err_norm = want[0] > 0.0 ? worst / want[0] : worst;
                            e.sigma_err_comp = comp;
                        }
                        report.matrix_errors.push_back(e);
                    }
                    out.rows = cols;
                    out.values.assign(h_v.ptr + v, h_v.ptr + v + cols * cols);
                    write(out);
                    a += rows * cols;
                    s += cols;
                    v += cols * cols;
                    report.total_sweeps += static_cast<std::size_t>(h_sw.ptr[k]);
                    report.matrix_sweeps.push_back(h_sw.ptr[k]);
                    const bool converged = h_sw.ptr[count + k] == JSVD_STOP_CONVERGED;
                    report.matrix_converged.push_back(converged);
                    report.unconverged_count += converged ? 0 : 1;
                }
                busy = false;
            }
        };
    } // namespace

    JacobiSvdPipeline::JacobiSvdPipeline(PipelineConfig config) : config_(std::move(config)) {}

    PipelineReport JacobiSvdPipeline::run() const
    {
        if (config_.input_path.empty())
            throw std::invalid_argument("Pipeline input path is empty.");
        if (config_.output_path.empty())
            throw std::invalid_argument("Pipeline output path is empty.");
        const MatrixFileFormat in_fmt = resolve(config_.input_format, config_.input_path);
        const MatrixFileFormat out_fmt = resolve(config_.output_format, config_.output_path);

        const JacobiSvdConfig &kc = config_.kernel_config;
        PipelineReport report;
        report.layout_transpose_mode = kc.layout_transpose_mode;
        report.layout_transpose_min_columns = kc.layout_transpose_min_columns;
        report.layout_transpose_min_elements = kc.layout_transpose_min_elements;
        jsvd_config cfg = to_c(kc);
        if (kc.layout_transpose_auto_tune)
        {
            // Reference pipeline.cu:39-48: tune before the run, report the result.
            jsvd_autotune_report tuned{};
            rethrow(jsvd_autotune(&cfg, &tuned));
            report.layout_transpose_auto_tuned = tuned.executed != 0;
            report.layout_transpose_min_columns = tuned.recommended_min_columns;
            report.layout_transpose_min_elements = tuned.recommended_min_elements;
            report.layout_transpose_estimated_best_speedup = tuned.estimated_best_speedup;
            cfg.layout_min_columns = tuned.recommended_min_columns;
            cfg.layout_min_elements = tuned.recommended_min_elements;
        }

        std::unique_ptr<io::MatDispatchReader> mat_in;
        std::unique_ptr<io::TxtReader> txt_in;
        if (in_fmt == MatrixFileFormat::mat)
            mat_in = std::make_unique<io::MatDispatchReader>(config_.input_path);
        else
            txt_in = std::make_unique<io::TxtReader>(config_.input_path);
        // Reference sigma for the error report: a matrix stream of 1 x n rows
        // in the input's formats (extension or the input format).
        std::function<bool(io::Matrix &)> reference;
        std::unique_ptr<io::MatReader> ref_mat;
        std::unique_ptr<io::TxtReader> ref_txt;
        if (config_.report_errors && !config_.reference_sigma_path.empty())
        {
            MatrixFileFormat rf = in_fmt;
            try
            {
                rf = resolve(MatrixFileFormat::auto_detect, config_.reference_sigma_path);
            }
            catch (const std::invalid_argument &)
            {
                // No .mat / .txt extension: the input's format.
            }
            if (rf == MatrixFileFormat::mat)
            {
                ref_mat = std::make_unique<io::MatReader>(config_.reference_sigma_path);
                reference = [&](io::Matrix &x) { return io::MatFilePolicy::read_next(*ref_mat, x); };
            }
            else
            {
                ref_txt = std::make_unique<io::TxtReader>(config_.reference_sigma_path);
                reference = [&](io::Matrix &x) { return io::TxtFilePolicy::read_next(*ref_txt, x); };
            }
        }
        report.errors_reported = config_.report_errors;
        std::function<void(const io::Matrix &)> write;
        std::function<void()> flush;
        std::unique_ptr<io::MatWriter> mat_out;
        std::unique_ptr<io::TxtWriter> txt_out;
        if (out_fmt == MatrixFileFormat::mat)
        {
            mat_out = std::make_unique<io::MatWriter>(config_.output_path);
            write = [&](const io::Matrix &x) { io::MatFilePolicy::write_next(*mat_out, x); };
            flush = [&]() { io::MatFilePolicy::flush(*mat_out); };
        }
        else
        {
            txt_out = std::make_unique<io::TxtWriter>(config_.output_path);
            write = [&](const io::Matrix &x) { io::TxtFilePolicy::write_next(*txt_out, x); };
            flush = [&]() { io::TxtFilePolicy::flush(*txt_out); };
        }

        const std::size_t window = std::clamp<std::size_t>(config_.max_queued_results, 1, 8);
        std::vector<std::unique_ptr<Slot>> slots;
        std::size_t next_slot = 0, submitted = 0, read_count = 0;
        // The slot whose staging receives the batch being read: the oldest
        // one, drained first, while the other slots' batches run.
        Slot *fill = nullptr;
        auto acquire = [&]() -> Slot & {
            if (slots.size() < window)
            {
                slots.push_back(std::make_unique<Slot>());
                slots.back()->metrics = config_.report_errors;
                slots.back()->raw_be = in_fmt == MatrixFileFormat::mat;
            }
            Slot &slot = *slots[next_slot];
            if (slot.busy)
                slot.drain(write, report, reference);
            slot.begin();
            return slot;
        };
        auto submit = [&]() {
            if (!fill || fill->bm.empty())
                return;
            fill->submit(cfg);
            submitted += fill->bm.size();
            next_slot = (next_slot + 1) % window;
            fill = nullptr;
        };
        auto check_shape = [&](std::size_t rows, std::size_t cols) {
            if (rows == 0 || cols == 0)
                throw std::invalid_argument("Testcase[" + std::to_string(read_count) + "] has zero dimension.");
        };
        auto added = [&]() {
            ++read_count;
            if (fill->bm.size() >= kBatchMatrices || fill->staged >= kBatchElements)
                submit();
        };
        if (mat_in)
        {
            std::size_t rows = 0, cols = 0;
            while (mat_in->read_header(rows, cols))
            {
                check_shape(rows, cols);
                if (!fill)
                    fill = &acquire();
                std::byte *at = fill->stage(rows, cols);
                (void)mat_in->read_payload(std::span<std::byte>(at, rows * cols * sizeof(double)));
                added();
            }
        }
        else
        {
            io::Matrix x;
            while (io::TxtFilePolicy::read_next(*txt_in, x))
            {
                check_shape(x.rows, x.columns);
                if (x.values.size() != x.rows * x.columns)
                    throw std::invalid_argument("Testcase[" + std::to_string(read_count) +
                                                "] payload size does not match rows*columns.");
                if (!fill)
                    fill = &acquire();
                std::memcpy(fill->stage(x.rows, x.columns), x.values.data(), x.values.size() * sizeof(double));
                added();
            }
        }
        submit();
        // Drain the remaining slots oldest first.
        for (std::size_t i = 0; i < slots.size(); ++i)
        {
            Slot &slot = *slots[(next_slot + i) % slots.size()];
            if (slot.busy)
                slot.drain(write, report, reference);
        }
        flush();
        report.testcase_count = submitted;
        report.emitted_matrix_count = 3 * submitted;
        return report;
    }

    PipelineReport run_pipeline(const PipelineConfig &config)
    {
        return JacobiSvdPipeline(config).run();
    }
} // namespace jacobi::svd::pipeline
