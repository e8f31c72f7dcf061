// PageRank pull rounds over pieces of the incoming CSR held anywhere in
// device memory (the hybrid engine's resident units and streaming ring).
//
// A segment is a range of destination rows [v0, v1) with the in-edges
// [lo, hi) of the graph's incoming CSR, `col` pointing at in-edge lo; a row's
// in-edges outside [lo, hi) belong to other segments (a row longer than a
// window is split over several). Per segment, like pull_light / pull_heavy
// over the whole transpose:
//   * rows whose in-edges in the segment number at most light_max: 64 rows
//     per wavefront, their in-edges walked together, a segmented scan across
//     the wavefront folds each row's shares, and the row's lane adds the
//     total to its residual with one store;
//   * longer rows go to a chunk queue (chunks of `chunk` in-edges, the
//     segment's index in `value`), which pr_pull_seg_heavy reduces one
//     wavefront per chunk, finishing with one atomic per chunk.
// Every row's share is rewritten by the compaction before a round (0 outside
// the frontier), so an in-edge adds the word of share[source] (global
// PageRank: the float share of ops.hpp) whatever its source. The sums are 64-bit integer words, so the totals equal the push
// rounds' bit for bit. Segments of one launch hold disjoint rows, and
// launches on one stream run in order, so the single store is safe.
#pragma once

#include "kern_common.hpp"
#include "ops.hpp"

namespace hyt
{
    struct PullSeg
    {
        std::uint32_t v0, v1;     // destination rows
        std::uint64_t lo, hi;     // in-edge positions of the segment
        const std::uint32_t *col; // device address of in-edge lo
    };

    struct PullQueue
    {
        HeavyChunk *chunks;
        std::uint32_t *count;    // chunks appended (reset before each launch)
        std::uint32_t *overflow; // set when a chunk did not fit
        std::uint64_t cap;
    };

    // Grid: x over the largest segment's rows (kBlock rows per workgroup), y
    // over the segments; segs == nullptr runs the one segment `one`.
    template <class OffT>
    __global__ __launch_bounds__(kBlock) void pr_pull_seg(const OffT *__restrict__ in_off, const PullSeg *segs,
                                                          PullSeg one, const float *__restrict__ share,
                                                          unsigned long long *res, std::uint32_t light_max,
                                                          std::uint32_t chunk, PullQueue pq)
    {
        __shared__ std::uint32_t s_ex[kWavesPerBlock][kWave];
        __shared__ std::uint64_t s_begin[kWavesPerBlock][kWave];
        __shared__ unsigned long long s_acc[kWavesPerBlock][kWave];

        const PullSeg sg = segs != nullptr ? segs[blockIdx.y] : one;
        const int lane = lane_id();
        const int wid = static_cast<int>(threadIdx.x) / kWave;
        const std::uint64_t v = static_cast<std::uint64_t>(sg.v0) + static_cast<std::uint64_t>(blockIdx.x) * kBlock +
                                threadIdx.x;
        std::uint64_t begin = 0, count = 0;
        if (v < sg.v1)
        {
            std::uint64_t b = in_off[v], e = in_off[v + 1];
            b = b < sg.lo ? sg.lo : b;
            e = e > sg.hi ? sg.hi : e;
            if (e > b)
            {
                begin = b;
                count = e - b;
            }
        }
        const std::uint32_t deg = count <= light_max ? static_cast<std::uint32_t>(count) : 0U;

        // rows above light_max: their chunks, one queue atomic per wavefront
        const std::uint32_t nch =
            count > light_max ? static_cast<std::uint32_t>((count + chunk - 1) / chunk) : 0U;
        const std::uint32_t cincl = wave_incl_scan(nch);
        const std::uint32_t ctotal = __shfl(cincl, kWave - 1, kWave);
        if (ctotal != 0)
        {
            std::uint32_t base = 0;
            if (lane == kWave - 1)
                base = atomicAdd(pq.count, ctotal);
            const std::uint64_t slot = static_cast<std::uint64_t>(__shfl(base, kWave - 1, kWave)) + cincl - nch;
            for (std::uint32_t c = 0; c < nch; ++c)
            {
                if (slot + c >= pq.cap)
                {
                    atomicOr(pq.overflow, 1U);
                    break;
                }
                const std::uint64_t clo = static_cast<std::uint64_t>(c) * chunk;
                HeavyChunk h;
                h.value = blockIdx.y;
                h.count = static_cast<std::uint32_t>(count - clo < chunk ? count - clo : chunk);
                h.begin = begin + clo;
                h.row = static_cast<std::uint32_t>(v);
                h.record = 0;
                pq.chunks[slot + c] = h;
            }
        }

        // light rows: the union of the wavefront's in-edges
        const std::uint32_t incl = wave_incl_scan(deg);
        const std::uint32_t total = __shfl(incl, kWave - 1, kWave);
        s_ex[wid][lane] = incl - deg;
        s_begin[wid][lane] = begin;
        s_acc[wid][lane] = 0ULL;
        __syncthreads();
        const std::uint32_t *ex = s_ex[wid];
        const std::uint32_t *col = sg.col;
        const std::uint64_t lo = sg.lo;
        // (no early return: every thread reaches the second barrier)
        for (std::uint32_t base = 0; base < total; base += kWave * kUnroll)
        {
            unsigned long long cc[kUnroll];
            int jj[kUnroll];
#pragma unroll
            for (int k = 0; k < kUnroll; ++k)
            {
                const std::uint32_t e = base + static_cast<std::uint32_t>(k * kWave + lane);
                jj[k] = kWave; // past the last row: matches no segment of rows
                cc[k] = 0ULL;
                if (e < total)
                {
                    const int j = owner_of(ex, e);
                    jj[k] = j;
                    cc[k] = word_of(share[col[s_begin[wid][j] + (e - ex[j]) - lo]]);
                }
            }
#pragma unroll
            for (int k = 0; k < kUnroll; ++k)
            {
                unsigned long long val = cc[k];
                const int j = jj[k];
#pragma unroll
                for (int d = 1; d < kWave; d <<= 1)
                {
                    const unsigned long long y = __shfl_up(val, d, kWave);
                    const int yj = __shfl_up(j, d, kWave);
                    if (lane >= d && yj == j)
                        val += y;
                }
                const int nj = __shfl_down(j, 1, kWave);
                const bool tail = j < kWave && (lane == kWave - 1 || nj != j);
                if (tail && val != 0ULL)
                    s_acc[wid][j] += val;
            }
        }
        __syncthreads();
        if (deg > 0)
        {
            const unsigned long long acc = s_acc[wid][lane];
            if (acc != 0ULL)
                res[v] += acc;
        }
    }

    // The chunk queue of the pr_pull_seg launch before it on the stream.
    __global__ __launch_bounds__(kBlock) void pr_pull_seg_heavy(const PullSeg *segs, PullSeg one,
                                                                const float *__restrict__ share,
                                                                unsigned long long *res, PullQueue pq)
    {
        const int lane = lane_id();
        const std::uint64_t n64 = *pq.count;
        const std::uint64_t n = n64 < pq.cap ? n64 : pq.cap;
        const std::uint64_t waves = static_cast<std::uint64_t>(gridDim.x) * kWavesPerBlock;
        for (std::uint64_t c = (static_cast<std::uint64_t>(blockIdx.x) * kBlock + threadIdx.x) / kWave; c < n;
             c += waves)
        {
            const HeavyChunk h = pq.chunks[c];
            const PullSeg sg = segs != nullptr ? segs[h.value] : one;
            const std::uint32_t *col = sg.col; // in-edge sg.lo
            unsigned long long acc = 0ULL;
            for (std::uint32_t base = 0; base < h.count; base += kWave * kUnroll)
            {
                unsigned long long cc[kUnroll];
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                {
                    const std::uint32_t e = base + static_cast<std::uint32_t>(k * kWave + lane);
                    cc[k] = e < h.count ? word_of(share[col[h.begin + e - sg.lo]]) : 0ULL;
                }
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                    acc += cc[k];
            }
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1)
                acc += __shfl_xor(acc, d, kWave);
            if (lane == 0 && acc != 0ULL)
                atomicAdd(&res[h.row], acc);
        }
    }
} // namespace hyt
