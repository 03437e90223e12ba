// jsvd_dispatch.cpp -- pinned task buffers and record-at-a-time *.mat
// reading (reference src/infrastructure/pinned_host_task_buffer.cu:28-96,
// src/infrastructure/mat_dispatch_reader.cu:8-84).  Host only; the
// allocations are hipHostMalloc'd so the pipeline copies payloads to HBM by
// DMA straight from them.
#include "jacobi/svd/io/mat_dispatch.cuh"

#include <hip/hip_runtime_api.h>

#include <bit>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <utility>

namespace jacobi::svd::io
{
    namespace
    {
        constexpr std::size_t kFirstCapacity = std::size_t(64) << 10;

        std::size_t mul_checked(std::size_t a, std::size_t b)
        {
            if (a != 0 && b > std::numeric_limits<std::size_t>::max() / a)
                throw std::overflow_error("Size multiplication overflow.");
            return a * b;
        }

        std::size_t add_checked(std::size_t a, std::size_t b)
        {
            if (a > std::numeric_limits<std::size_t>::max() - b)
                throw std::overflow_error("Size addition overflow.");
            return a + b;
        }

        std::uint64_t from_network(std::uint64_t x)
        {
            if constexpr (std::endian::native == std::endian::little)
                return __builtin_bswap64(x);
            return x;
        }

        // Doubling from 64 KiB until the request fits (the request itself
        // when doubling would overflow).
        std::size_t grown_capacity(std::size_t current, std::size_t required)
        {
            if (required == 0)
                return 0;
            std::size_t c = current == 0 ? kFirstCapacity : current;
            while (c < required)
            {
                if (c > std::numeric_limits<std::size_t>::max() / 2)
                    return required;
                c *= 2;
            }
            return c;
        }

        std::streamsize to_streamsize(std::size_t bytes)
        {
            if (bytes > static_cast<std::size_t>(std::numeric_limits<std::streamsize>::max()))
                throw std::overflow_error("Byte size exceeds streamsize range.");
            return static_cast<std::streamsize>(bytes);
        }
    } // namespace

    PinnedHostTaskBuffer::~PinnedHostTaskBuffer()
    {
        release();
    }

    PinnedHostTaskBuffer::PinnedHostTaskBuffer(PinnedHostTaskBuffer &&other) noexcept
    {
        move_from(std::move(other));
    }

    PinnedHostTaskBuffer &PinnedHostTaskBuffer::operator=(PinnedHostTaskBuffer &&other) noexcept
    {
        if (this != &other)
        {
            release();
            move_from(std::move(other));
        }
        return *this;
    }

    void PinnedHostTaskBuffer::reserve(std::size_t input_bytes, std::size_t workspace_bytes)
    {
        const std::size_t need = add_checked(input_bytes, workspace_bytes);
        if (need > capacity_bytes_)
        {
            const std::size_t cap = grown_capacity(capacity_bytes_, need);
            void *p = nullptr;
            const hipError_t e = hipHostMalloc(&p, cap, hipHostMallocDefault);
            if (e != hipSuccess)
            {
                (void)hipGetLastError(); // consume it: the failure is reported here
                throw std::runtime_error(std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
            }
            release();
            data_ = static_cast<std::byte *>(p);
            capacity_bytes_ = cap;
        }
        input_size_bytes_ = input_bytes;
        workspace_size_bytes_ = workspace_bytes;
    }

    std::span<std::byte> PinnedHostTaskBuffer::mutable_input_bytes() noexcept
    {
        return {data_, input_size_bytes_};
    }

    std::span<const std::byte> PinnedHostTaskBuffer::input_bytes() const noexcept
    {
        return {data_, input_size_bytes_};
    }

    std::span<std::byte> PinnedHostTaskBuffer::mutable_workspace_bytes() noexcept
    {
        return {data_ == nullptr ? nullptr : data_ + input_size_bytes_, workspace_size_bytes_};
    }

    std::span<const std::byte> PinnedHostTaskBuffer::workspace_bytes() const noexcept
    {
        return {data_ == nullptr ? nullptr : data_ + input_size_bytes_, workspace_size_bytes_};
    }

    std::size_t PinnedHostTaskBuffer::capacity_bytes() const noexcept
    {
        return capacity_bytes_;
    }

    std::size_t PinnedHostTaskBuffer::input_size_bytes() const noexcept
    {
        return input_size_bytes_;
    }

    std::size_t PinnedHostTaskBuffer::workspace_size_bytes() const noexcept
    {
        return workspace_size_bytes_;
    }

    void PinnedHostTaskBuffer::release() noexcept
    {
        if (data_ != nullptr)
            (void)hipHostFree(data_);
        data_ = nullptr;
        capacity_bytes_ = input_size_bytes_ = workspace_size_bytes_ = 0;
    }

    void PinnedHostTaskBuffer::move_from(PinnedHostTaskBuffer &&other) noexcept
    {
        data_ = std::exchange(other.data_, nullptr);
        capacity_bytes_ = std::exchange(other.capacity_bytes_, 0);
        input_size_bytes_ = std::exchange(other.input_size_bytes_, 0);
        workspace_size_bytes_ = std::exchange(other.workspace_size_bytes_, 0);
    }

    MatDispatchReader::MatDispatchReader(const std::filesystem::path &path)
        : input_(path, std::ios::in | std::ios::binary)
    {
        if (!input_)
            throw std::runtime_error("Failed to open *.mat file for dispatch reading.");
    }

    bool MatDispatchReader::read_header(std::size_t &rows, std::size_t &columns)
    {
        if (pending_)
            throw std::logic_error("The payload of the previous *.mat header has not been read.");
        MatMetaData h{};
        input_.read(reinterpret_cast<char *>(&h), sizeof(h));
        const std::streamsize got = input_.gcount();
        if (got == 0 && input_.eof())
            return false;
        if (got != static_cast<std::streamsize>(sizeof(h)))
            throw std::runtime_error("Truncated *.mat header while dispatch reading.");
        const std::uint64_t r = from_network(h.rows), c = from_network(h.columns);
        if (r > std::numeric_limits<std::size_t>::max() || c > std::numeric_limits<std::size_t>::max())
            throw std::overflow_error("Matrix dimensions exceed platform size_t range.");
        const std::size_t bytes = mul_checked(mul_checked(static_cast<std::size_t>(r), static_cast<std::size_t>(c)),
                                              sizeof(double));
        // The payload must be in the file before memory is reserved for it:
        // a corrupt header then fails here instead of pinning gigabytes.
        const std::streampos here = input_.tellg();
        input_.seekg(0, std::ios::end);
        const std::streampos end = input_.tellg();
        input_.seekg(here);
        if (here < 0 || end < here || bytes > static_cast<std::size_t>(end - here))
            throw std::runtime_error("Truncated *.mat payload while dispatch reading.");
        rows = static_cast<std::size_t>(r);
        columns = static_cast<std::size_t>(c);
        pending_payload_bytes_ = bytes;
        pending_ = true;
        return true;
    }

    std::size_t MatDispatchReader::read_payload(std::span<std::byte> destination)
    {
        if (!pending_)
            throw std::logic_error("No *.mat header is pending.");
        if (destination.size() != pending_payload_bytes_)
            throw std::invalid_argument("Dispatch payload destination size does not match rows * columns.");
        if (!destination.empty())
        {
            input_.read(reinterpret_cast<char *>(destination.data()), to_streamsize(destination.size()));
            if (input_.gcount() != static_cast<std::streamsize>(destination.size()))
                throw std::runtime_error("Truncated *.mat payload while dispatch reading.");
        }
        pending_ = false;
        const std::size_t index = next_sequence_index_;
        next_sequence_index_ = add_checked(next_sequence_index_, 1);
        return index;
    }

    bool MatDispatchReader::read_next(MatDispatchTask &task, std::size_t workspace_bytes)
    {
        std::size_t rows = 0, columns = 0;
        if (!read_header(rows, columns))
            return false;
        task.buffer.reserve(pending_payload_bytes_, workspace_bytes);
        task.sequence_index = read_payload(task.buffer.mutable_input_bytes());
        task.rows = rows;
        task.columns = columns;
        return true;
    }

    Matrix decode_dispatch_task_matrix(const MatDispatchTask &task)
    {
        Matrix out;
        out.rows = task.rows;
        out.columns = task.columns;
        const std::size_t count = mul_checked(task.rows, task.columns);
        const std::span<const std::byte> payload = task.buffer.input_bytes();
        if (payload.size() != mul_checked(count, sizeof(double)))
            throw std::invalid_argument("Dispatch payload size does not match rows * columns.");
        out.values.resize(count);
        for (std::size_t i = 0; i < count; ++i)
        {
            std::uint64_t x;
            std::memcpy(&x, payload.data() + i * sizeof(double), sizeof(x));
            out.values[i] = std::bit_cast<double>(from_network(x));
        }
        return out;
    }
} // namespace jacobi::svd::io
