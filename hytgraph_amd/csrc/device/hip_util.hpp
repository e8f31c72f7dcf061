// HIP helpers: status checking that throws hyt::Error, an owning device
// buffer, and 64-lane wavefront primitives for gfx950.
#pragma once

#include "common.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>

namespace hyt
{
    inline void hip_check(hipError_t st, const char *what)
    {
        if (st != hipSuccess)
        {
            (void)hipGetLastError();
            int code = (st == hipErrorOutOfMemory || st == hipErrorMemoryAllocation)
                           ? HYT_ERR_OUT_OF_MEMORY
                           : (st == hipErrorNoDevice || st == hipErrorInvalidDevice) ? HYT_ERR_NO_DEVICE
                                                                                      : HYT_ERR_HIP;
            fail(code, std::string(what) + ": " + hipGetErrorString(st));
        }
    }

    // Fails with HYT_ERR_NO_DEVICE when no device is visible; returns the
    // current device id.
    inline int require_device()
    {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        {
            (void)hipGetLastError();
            fail(HYT_ERR_NO_DEVICE, "no HIP device is available");
        }
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        return dev;
    }

    // Owning device allocation (hipMalloc / hipFree).
    struct DeviceBuffer
    {
        void *ptr = nullptr;
        std::size_t bytes = 0;

        DeviceBuffer() = default;
        explicit DeviceBuffer(std::size_t n, const char *what = "hipMalloc") { alloc(n, what); }
        ~DeviceBuffer() { release(); }
        DeviceBuffer(const DeviceBuffer &) = delete;
        DeviceBuffer &operator=(const DeviceBuffer &) = delete;
        DeviceBuffer(DeviceBuffer &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr, o.bytes = 0; }
        DeviceBuffer &operator=(DeviceBuffer &&o) noexcept
        {
            if (this != &o)
            {
                release();
                ptr = o.ptr;
                bytes = o.bytes;
                o.ptr = nullptr;
                o.bytes = 0;
            }
            return *this;
        }

        void alloc(std::size_t n, const char *what = "hipMalloc")
        {
            release();
            if (n == 0)
                n = 4;
            hip_check(hipMalloc(&ptr, n), what);
            bytes = n;
        }
        // Grows (never shrinks); contents are not preserved.
        void ensure(std::size_t n, const char *what = "hipMalloc")
        {
            if (n > bytes || ptr == nullptr)
                alloc(n, what);
        }
        void release() noexcept
        {
            if (ptr)
                (void)hipFree(ptr);
            ptr = nullptr;
            bytes = 0;
        }
        template <class T>
        T *as() const { return static_cast<T *>(ptr); }
    };

    // Pinned host allocation for small readbacks (counters).
    struct PinnedBuffer
    {
        void *ptr = nullptr;
        explicit PinnedBuffer(std::size_t n)
        {
            hip_check(hipHostMalloc(&ptr, n, hipHostMallocDefault), "hipHostMalloc(counters)");
        }
        ~PinnedBuffer()
        {
            if (ptr)
                (void)hipHostFree(ptr);
        }
        PinnedBuffer(const PinnedBuffer &) = delete;
        PinnedBuffer &operator=(const PinnedBuffer &) = delete;
        template <class T>
        T *as() const { return static_cast<T *>(ptr); }
    };

    // ------------------------------------------------------------ device side
    constexpr int kWave = 64;

    __device__ __forceinline__ int lane_id() { return static_cast<int>(__lane_id()); }

    // Inclusive prefix sum across the 64 lanes of a wavefront.
    __device__ __forceinline__ std::uint32_t wave_incl_scan(std::uint32_t x)
    {
        const int lane = lane_id();
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1)
        {
            std::uint32_t y = __shfl_up(x, d, kWave);
            if (lane >= d)
                x += y;
        }
        return x;
    }

    __device__ __forceinline__ std::uint32_t lanes_below(std::uint64_t mask)
    {
        const int lane = lane_id();
        return static_cast<std::uint32_t>(__popcll(mask & ((1ULL << lane) - 1ULL)));
    }

    // Wavefront-aggregated append: every lane with want=true receives a
    // distinct slot in [*counter, *counter + popcount) with one atomic per
    // wavefront. Must be reached by all active lanes of the wavefront.
    __device__ __forceinline__ std::uint32_t wave_append_slot(bool want, std::uint32_t *counter,
                                                              std::uint64_t &mask_out)
    {
        const std::uint64_t mask = __ballot(want);
        mask_out = mask;
        if (mask == 0)
            return 0;
        const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        std::uint32_t base = 0;
        if (lane_id() == leader)
            base = atomicAdd(counter, static_cast<std::uint32_t>(__popcll(mask)));
        base = __shfl(base, leader, kWave);
        return base + lanes_below(mask);
    }
} // namespace hyt
