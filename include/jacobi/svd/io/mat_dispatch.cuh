#pragma once
// Record-at-a-time *.mat reading into pinned host memory (reference
// include/jacobi/svd/io/mat_dispatch.cuh, src/infrastructure/mat_dispatch_reader.cu:8-84).
//
// A task keeps the record's payload exactly as stored in the file (row-major
// doubles in network byte order), so the bytes can go to the device without
// a host-side decode; decode_dispatch_task_matrix() converts a task on the
// host when a Matrix is wanted.
#include "jacobi/svd/io/mat_metadata.cuh"
#include "jacobi/svd/io/matrix.cuh"
#include "jacobi/svd/io/pinned_host_task_buffer.cuh"

#include <cstddef>
#include <filesystem>
#include <fstream>
#include <span>

namespace jacobi::svd::io
{
    struct MatDispatchTask final
    {
        std::size_t sequence_index = 0; // 0, 1, 2, ... in file order
        std::size_t rows = 0;
        std::size_t columns = 0;
        PinnedHostTaskBuffer buffer; // input region: the raw payload; then the workspace
    };

    class MatDispatchReader final
    {
    public:
        // Throws std::runtime_error("Failed to open *.mat file for dispatch reading.").
        explicit MatDispatchReader(const std::filesystem::path &path);

        // Reads the next record into task (buffer reserved for the payload
        // plus workspace_bytes).  Returns false at a clean end of file.
        // Throws std::runtime_error on a truncated header or payload; the
        // payload is checked against the bytes left in the file before any
        // memory is reserved for it.
        [[nodiscard]] bool read_next(MatDispatchTask &task, std::size_t workspace_bytes = 0);

        // Appended (no reference counterpart): the same record read in two
        // steps, so a caller can place payloads where it wants them, e.g.
        // back to back in one batch staging buffer.  read_header() consumes
        // the next header, checks that its payload is in the file, and
        // returns false at a clean end of file.  read_payload() then reads
        // that payload (destination.size() must be rows * columns * 8) and
        // returns the record's sequence index.
        [[nodiscard]] bool read_header(std::size_t &rows, std::size_t &columns);
        std::size_t read_payload(std::span<std::byte> destination);

    private:
        std::ifstream input_;
        std::size_t next_sequence_index_ = 0;
        std::size_t pending_payload_bytes_ = 0; // payload of the header last read, not yet consumed
        bool pending_ = false;
    };

    // Row-major FP64 matrix from a task's payload.  Throws
    // std::invalid_argument("Dispatch payload size does not match rows * columns.").
    [[nodiscard]] Matrix decode_dispatch_task_matrix(const MatDispatchTask &task);
} // namespace jacobi::svd::io
