#pragma once
// One page-locked host allocation split into an input region followed by a
// workspace region (reference include/jacobi/svd/io/pinned_host_task_buffer.cuh,
// src/infrastructure/pinned_host_task_buffer.cu:28-96).  Memory comes from
// hipHostMalloc, so hipMemcpyAsync from the input region is a direct DMA.
// reserve() grows the capacity by doubling from 64 KiB; a grown buffer does
// not keep its old contents.  A reserve() that fits the current capacity
// only resets the two region sizes.  A failed allocation throws
// std::runtime_error("hipHostMalloc failed: <HIP error string>").
#include <cstddef>
#include <span>

namespace jacobi::svd::io
{
    class PinnedHostTaskBuffer final
    {
    public:
        PinnedHostTaskBuffer() = default;
        ~PinnedHostTaskBuffer();
        PinnedHostTaskBuffer(const PinnedHostTaskBuffer &) = delete;
        PinnedHostTaskBuffer &operator=(const PinnedHostTaskBuffer &) = delete;
        PinnedHostTaskBuffer(PinnedHostTaskBuffer &&other) noexcept;
        PinnedHostTaskBuffer &operator=(PinnedHostTaskBuffer &&other) noexcept;

        // Input region of input_bytes, then workspace_bytes of workspace.
        void reserve(std::size_t input_bytes, std::size_t workspace_bytes);

        [[nodiscard]] std::span<std::byte> mutable_input_bytes() noexcept;
        [[nodiscard]] std::span<const std::byte> input_bytes() const noexcept;
        [[nodiscard]] std::span<std::byte> mutable_workspace_bytes() noexcept;
        [[nodiscard]] std::span<const std::byte> workspace_bytes() const noexcept;
        [[nodiscard]] std::size_t capacity_bytes() const noexcept;
        [[nodiscard]] std::size_t input_size_bytes() const noexcept;
        [[nodiscard]] std::size_t workspace_size_bytes() const noexcept;

    private:
        void release() noexcept;
        void move_from(PinnedHostTaskBuffer &&other) noexcept;

        std::byte *data_ = nullptr;
        std::size_t capacity_bytes_ = 0;
        std::size_t input_size_bytes_ = 0;
        std::size_t workspace_size_bytes_ = 0;
    };
} // namespace jacobi::svd::io
