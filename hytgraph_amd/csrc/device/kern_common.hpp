// Kernel building blocks shared by the traversal kernels: fills, the
// wavefront load-balanced edge walk, and launch helpers.
#pragma once

#include "dgraph.hpp"

namespace hyt
{
    constexpr int kBlock = 256;                 // 4 wavefronts per workgroup
    constexpr int kWavesPerBlock = kBlock / kWave;
    constexpr int kUnroll = 4;                  // edges in flight per lane
    constexpr unsigned kHeavyGrid = 2048;       // workgroups of the chunk kernels
    // Single-source SSSP and PageRank chunk kernels (RMAT 24 SSSP: 1024
    // workgroups ran 32.5 GTEPS against 31.7 at 2048, 29.9 at 512 and 31.3
    // at 4096; PageRank 0.305 against 0.310 s). The K-lane chunk kernels keep
    // 2048: at 1024 batched K = 8 SSSP fell from 284 to 230 queries/s.
    constexpr unsigned kSingleHeavyGrid = 1024;

    template <class T>
    __global__ void fill_kernel(T *p, std::uint64_t n, T value)
    {
        std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
        for (; i < n; i += stride)
            p[i] = value;
    }

    template <class T>
    inline void launch_fill(T *p, std::uint64_t n, T value, hipStream_t s)
    {
        if (n == 0)
            return;
        std::uint64_t blocks = (n + kBlock - 1) / kBlock;
        if (blocks > 65536)
            blocks = 65536;
        fill_kernel<T><<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(p, n, value);
        hip_check(hipGetLastError(), "launch fill");
    }

    // Counter resets by a fill kernel rather than hipMemsetAsync, so that a
    // sequence of rounds can be captured into a HIP graph (memset nodes have
    // given wrong results on replays). Counters and its round part are whole
    // 32-bit words.
    inline void clear_counters(Counters *c, hipStream_t s)
    {
        static_assert(sizeof(Counters) % 4 == 0, "Counters is whole words");
        launch_fill<std::uint32_t>(reinterpret_cast<std::uint32_t *>(c), sizeof(Counters) / 4, 0U, s);
    }
    inline void clear_round_counters(Counters *c, hipStream_t s)
    {
        static_assert(kRoundCounterBytes % 4 == 0, "the round counters are whole words");
        launch_fill<std::uint32_t>(reinterpret_cast<std::uint32_t *>(c), kRoundCounterBytes / 4, 0U, s);
    }

    inline unsigned blocks_for(std::uint64_t items)
    {
        std::uint64_t b = (items + kBlock - 1) / kBlock;
        return static_cast<unsigned>(b == 0 ? 1 : b);
    }

    // Largest j in [0, 64) with ex[j] <= e, for a non-decreasing ex with
    // ex[0] = 0 <= e (binary search in LDS).
    __device__ __forceinline__ int owner_of(const std::uint32_t *ex, std::uint32_t e)
    {
        int lo = 0;
#pragma unroll
        for (int s = 32; s > 0; s >>= 1)
            if (ex[lo + s] <= e)
                lo += s;
        return lo;
    }
} // namespace hyt
