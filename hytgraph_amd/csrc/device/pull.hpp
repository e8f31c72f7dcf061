// Pull rounds over the incoming CSR (direction-optimised traversal).
//
// When the frontier's out-edges are a large share of E, reading every
// in-edge once is cheaper than pushing: no global atomics, one plain store
// per destination row. Rows are the destinations v = 0..V-1 in order:
//   * light rows (in-degree <= light_max): a wavefront takes 64 consecutive
//     rows and walks the union of their in-edges with the same load-balanced
//     walk as the push kernels. Consecutive lanes hold consecutive in-edges,
//     so the rows form contiguous lane segments: a segmented inclusive scan
//     across the wavefront (6 shuffle steps) leaves each segment's total in
//     its last lane, which folds it into the row's LDS slot - one LDS write per
//     row segment, no LDS atomics. Lane j then finishes row j with one store;
//   * heavy rows come from a static chunk list; a wavefront reduces its chunk
//     in registers and finishes with one global atomic.
// Contributions come only from sources set in the frontier bitmap.
//
// Op supplies (Val = the type identity() returns: u32 SSSP distances, u64
// PageRank mass words):
//   Val identity()
//   bool edge(uint64_t pos, Val &c)             contribution of in-edge pos
//   Val fold(Val a, Val b)                      register fold
//   void finish(uint32_t v, Val acc)            single writer of row v
//   void finish_atomic(uint32_t v, Val acc)     one of several chunks of v
#pragma once

#include "kern_common.hpp"

namespace hyt
{
    __device__ __forceinline__ bool bit_set(const std::uint32_t *bits, std::uint32_t u)
    {
        return (bits[u >> 5] >> (u & 31U)) & 1U;
    }

    // A pull round scans every in-edge of the graph (light rows here, the
    // static chunk list of heavy rows in pull_heavy), so the caller counts E
    // edges for it; the kernels keep no shared counter.
    template <class OffT, class Op>
    __global__ __launch_bounds__(kBlock) void pull_light(const OffT *__restrict__ in_off, Op op, std::uint64_t V,
                                                         std::uint32_t light_max)
    {
        using Val = decltype(op.identity());
        __shared__ std::uint32_t s_ex[kWavesPerBlock][kWave];
        __shared__ OffT s_begin[kWavesPerBlock][kWave];
        __shared__ Val s_acc[kWavesPerBlock][kWave];

        const int lane = lane_id();
        const int wid = static_cast<int>(threadIdx.x) / kWave;
        const std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
        std::uint32_t deg = 0;
        OffT begin = 0;
        if (v < V)
        {
            begin = in_off[v];
            const OffT full = in_off[v + 1] - begin;
            if (full <= light_max)
                deg = static_cast<std::uint32_t>(full);
        }
        const std::uint32_t incl = wave_incl_scan(deg);
        const std::uint32_t total = __shfl(incl, kWave - 1, kWave);
        s_ex[wid][lane] = incl - deg;
        s_begin[wid][lane] = begin;
        s_acc[wid][lane] = op.identity();
        __syncthreads();
        const std::uint32_t *ex = s_ex[wid];
        // (no early return: every thread reaches the second barrier)
        for (std::uint32_t base = 0; base < total; base += kWave * kUnroll)
        {
            Val cc[kUnroll];
            int jj[kUnroll];
#pragma unroll
            for (int k = 0; k < kUnroll; ++k)
            {
                const std::uint32_t e = base + static_cast<std::uint32_t>(k * kWave + lane);
                jj[k] = kWave; // past the last row: never matches a real segment
                cc[k] = op.identity();
                if (e < total)
                {
                    const int j = owner_of(ex, e);
                    jj[k] = j;
                    Val c;
                    if (op.edge(static_cast<std::uint64_t>(s_begin[wid][j]) + (e - ex[j]), c))
                        cc[k] = c;
                }
            }
#pragma unroll
            for (int k = 0; k < kUnroll; ++k)
            {
                Val val = cc[k];
                const int j = jj[k];
#pragma unroll
                for (int d = 1; d < kWave; d <<= 1)
                {
                    const Val y = __shfl_up(val, d, kWave);
                    const int yj = __shfl_up(j, d, kWave);
                    if (lane >= d && yj == j)
                        val = op.fold(val, y);
                }
                const int nj = __shfl_down(j, 1, kWave);
                const bool tail = j < kWave && (lane == kWave - 1 || nj != j);
                if (tail && val != op.identity())
                    s_acc[wid][j] = op.fold(s_acc[wid][j], val);
            }
        }
        __syncthreads();
        if (v < V && deg > 0)
            op.finish(static_cast<std::uint32_t>(v), s_acc[wid][lane]);
    }

    template <class Op>
    __global__ __launch_bounds__(kBlock) void pull_heavy(Op op, const HeavyChunk *chunks, std::uint64_t n)
    {
        using Val = decltype(op.identity());
        const int lane = lane_id();
        const std::uint64_t waves = static_cast<std::uint64_t>(gridDim.x) * kWavesPerBlock;
        for (std::uint64_t c = (static_cast<std::uint64_t>(blockIdx.x) * kBlock + threadIdx.x) / kWave; c < n;
             c += waves)
        {
            const HeavyChunk h = chunks[c];
            Val acc = op.identity();
            for (std::uint32_t base = 0; base < h.count; base += kWave * kUnroll)
            {
                Val cc[kUnroll];
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                {
                    const std::uint32_t e = base + static_cast<std::uint32_t>(k * kWave + lane);
                    cc[k] = op.identity();
                    Val c2;
                    if (e < h.count && op.edge(h.begin + e, c2))
                        cc[k] = c2;
                }
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                    acc = op.fold(acc, cc[k]);
            }
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1)
                acc = op.fold(acc, __shfl_xor(acc, d, kWave));
            if (lane == 0 && acc != op.identity())
                op.finish_atomic(h.value, acc);
        }
    }
} // namespace hyt
