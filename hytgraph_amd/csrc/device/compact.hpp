// Next-frontier compaction: a dense scan over all rows that selects the rows
// an expand pass touched, applies per-row side effects (snapshot values,
// clear marks), and writes them as the next queue.
//
// Each thread scans 16 consecutive rows (4 groups of 4, vector loads); the
// workgroup prefix-sums the per-thread counts with wavefront scans plus an LDS
// pass, and reserves its output range with one atomic. Rows keep ascending
// order within a workgroup, so consecutive queue entries are nearby vertices.
//
// Sel::select4(row0, n, vals[4], aux) -> bit mask of the selected rows among
// [row0, row0+n), n <= 4; vals[j] is the value of row row0+j when selected.
// Sel::whole16() true: Sel::select16(row0, V, vals[4][4], masks[4], aux)
// handles the thread's 16 rows in one call instead.
// Sel::Aux is per-thread side information (e.g. held-back rows) that
// Sel::publish(aux, ctr) reduces across the wavefront and publishes.
// Optionally (non-null pointers) the pass also writes the frontier bitmap
// (16 rows per thread = one u16 word, every word rewritten) and adds the
// out-degree sum of the selected rows to ctr->frontier_edges, which the host
// uses to choose a push or a pull round; and (HubOut) lists the selected
// rows that a per-vertex bitmap marks (rows above the expansion's light_max)
// apart as well, so the pass that queues their chunks reads them alone.
#pragma once

#include "kern_common.hpp"

namespace hyt
{
    // Selector side information for selectors that have none.
    struct NoAux
    {
    };

    // bits[t]: bit j = row 16t + j is listed (the same 16 rows per thread as
    // the compaction); list / count: the list and its length (appended, one
    // atomic per workgroup that has such rows; count on a line of its own).
    struct HubOut
    {
        const std::uint16_t *bits = nullptr;
        std::uint32_t *list = nullptr;
        std::uint32_t *count = nullptr;
    };


    constexpr int kCompactGroups = 4;                         // groups of 4 rows per thread
    constexpr int kCompactRows = 4 * kCompactGroups;          // rows per thread
    constexpr std::uint64_t kCompactRowsPerBlock = static_cast<std::uint64_t>(kBlock) * kCompactRows;

    inline unsigned compact_blocks(std::uint64_t V)
    {
        std::uint64_t b = (V + kCompactRowsPerBlock - 1) / kCompactRowsPerBlock;
        return static_cast<unsigned>(b == 0 ? 1 : b);
    }

    // block_counts == nullptr: rows go straight to q at a position reserved
    // with one atomic per workgroup. Otherwise (ordered mode) workgroup b
    // writes its rows at q[b * kCompactRowsPerBlock ...] and its count to
    // block_counts[b]; gather_ordered() then packs them in vertex order.
    // gate (optional): the pass does nothing when *gate is 0 (a device-driven
    // round whose queue was empty).
    template <class Val, class OffT, class Sel>
    __global__ __launch_bounds__(kBlock) void compact_rows(std::uint64_t V, Sel sel, std::uint32_t *q, Val *qval,
                                                           Counters *ctr, const OffT *off, std::uint16_t *active,
                                                           std::uint32_t *block_counts = nullptr,
                                                           const std::uint32_t *gate = nullptr, HubOut hubs = {},
                                                           std::uint64_t *t_start = nullptr)
    {
        if (gate != nullptr && *gate == 0)
            return; // uniform: every thread reads the same word
        // timed rounds: the device wall clock when the pass starts (the kernels
        // before it on the stream have finished)
        if (t_start != nullptr && blockIdx.x == 0 && threadIdx.x == 0)
            *t_start = static_cast<std::uint64_t>(wall_clock64());
        __shared__ std::uint32_t s_wsum[kWavesPerBlock];
        __shared__ unsigned long long s_dsum[kWavesPerBlock];
        __shared__ std::uint32_t s_base, s_hbase;
        const int lane = lane_id();
        const int wid = static_cast<int>(threadIdx.x) / kWave;
        const std::uint64_t row0 = (static_cast<std::uint64_t>(blockIdx.x) * kBlock + threadIdx.x) * kCompactRows;
        const std::uint64_t tid = static_cast<std::uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
        // the listed-row bits of the thread's 16 rows, read beside the selection's loads
        const std::uint32_t hbits = hubs.bits != nullptr && row0 < V ? hubs.bits[tid] : 0U;

        // Static indexing only (fully unrolled), so everything stays in VGPRs.
        Val gv[kCompactGroups][4];
        std::uint32_t gm[kCompactGroups];
        std::uint32_t n = 0;
        typename Sel::Aux aux{};
        if (sel.whole16())
        {
            // the selector reads its 16 rows at once (no store between the loads)
#pragma unroll
            for (int g = 0; g < kCompactGroups; ++g)
                gm[g] = 0;
            if (row0 < V)
                sel.select16(row0, V, gv, gm, aux);
#pragma unroll
            for (int g = 0; g < kCompactGroups; ++g)
                n += static_cast<std::uint32_t>(__popc(gm[g]));
        }
        else
        {
#pragma unroll
            for (int g = 0; g < kCompactGroups; ++g)
            {
                gm[g] = 0;
                const std::uint64_t r = row0 + 4 * g;
                if (r < V)
                {
                    const std::uint32_t avail = static_cast<std::uint32_t>(V - r < 4 ? V - r : 4);
                    gm[g] = sel.select4(r, avail, gv[g], aux);
                }
                n += static_cast<std::uint32_t>(__popc(gm[g]));
            }
        }

        sel.publish(aux, ctr);
        const std::uint32_t picked = gm[0] | (gm[1] << 4) | (gm[2] << 8) | (gm[3] << 12);
        if (active != nullptr && row0 < V)
            active[tid] = static_cast<std::uint16_t>(picked);
        std::uint32_t hm = picked & hbits;
        const std::uint32_t nh = static_cast<std::uint32_t>(__popc(hm));
        if (off != nullptr)
        {
            unsigned long long dsum = 0;
#pragma unroll
            for (int g = 0; g < kCompactGroups; ++g)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if ((gm[g] >> j) & 1U)
                    {
                        const std::uint64_t r = row0 + 4 * g + j;
                        dsum += static_cast<unsigned long long>(off[r + 1] - off[r]);
                    }
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1)
                dsum += __shfl_xor(dsum, d, kWave);
            if (lane == 0)
                s_dsum[wid] = dsum;
        }

        // rows (low 16 bits) and listed rows (high 16 bits) in one scan: a
        // workgroup selects at most kCompactRowsPerBlock = 4096 rows
        static_assert(kCompactRowsPerBlock < (1U << 16), "packed row counts");
        const std::uint32_t incl2 = wave_incl_scan(n | (nh << 16));
        const std::uint32_t incl = incl2 & 0xFFFFU;
        if (lane == kWave - 1)
            s_wsum[wid] = incl2;
        __syncthreads();
        if (off != nullptr && threadIdx.x == 0)
        {
            // one atomic per workgroup for the out-degree sum
            unsigned long long dsum = 0;
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; ++w)
                dsum += s_dsum[w];
            if (dsum)
                atomicAdd(&ctr->frontier_edges, dsum);
        }
        std::uint32_t wbase2 = 0, total2 = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w)
        {
            if (w < wid)
                wbase2 += s_wsum[w];
            total2 += s_wsum[w];
        }
        const std::uint32_t wbase = wbase2 & 0xFFFFU, total = total2 & 0xFFFFU;
        const std::uint32_t htotal = total2 >> 16;
        if (threadIdx.x == 0)
        {
            if (block_counts != nullptr)
            {
                block_counts[blockIdx.x] = total;
                s_base = static_cast<std::uint32_t>(blockIdx.x * kCompactRowsPerBlock);
            }
            else
                s_base = total ? atomicAdd(&ctr->next_count, total) : 0U;
            s_hbase = htotal ? atomicAdd(hubs.count, htotal) : 0U;
        }
        __syncthreads();
        if (hm != 0U)
        {
            std::uint32_t at = s_hbase + (wbase2 >> 16) + (incl2 >> 16) - nh;
            for (; hm != 0U; hm &= hm - 1U)
                hubs.list[at++] = static_cast<std::uint32_t>(row0) + static_cast<std::uint32_t>(__ffs(hm) - 1);
        }
        std::uint32_t pos = s_base + wbase + incl - n;
#pragma unroll
        for (int g = 0; g < kCompactGroups; ++g)
        {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if ((gm[g] >> j) & 1U)
                {
                    const std::uint32_t at = pos + static_cast<std::uint32_t>(__popc(gm[g] & ((1U << j) - 1U)));
                    q[at] = static_cast<std::uint32_t>(row0 + 4 * g + j);
                    if (qval)
                        qval[at] = gv[g][j];
                }
            pos += static_cast<std::uint32_t>(__popc(gm[g]));
        }
    }
    // Ordered mode, second pass: workgroup b copies its block_counts[b] rows
    // from the staging queue to q[block_offsets[b] ...]; the total lands in
    // ctr->next_count.
    template <class Val>
    __global__ __launch_bounds__(kBlock) void gather_ordered(const std::uint32_t *stage_q, const Val *stage_val,
                                                             const std::uint32_t *block_counts,
                                                             const std::uint32_t *block_offsets, unsigned nblocks,
                                                             std::uint32_t *q, Val *qval, Counters *ctr)
    {
        const unsigned b = blockIdx.x;
        const std::uint32_t n = block_counts[b];
        const std::uint32_t dst = block_offsets[b];
        const std::uint64_t src = static_cast<std::uint64_t>(b) * kCompactRowsPerBlock;
        for (std::uint32_t k = threadIdx.x; k < n; k += kBlock)
        {
            q[dst + k] = stage_q[src + k];
            if (qval)
                qval[dst + k] = stage_val[src + k];
        }
        if (b == nblocks - 1 && threadIdx.x == 0)
            ctr->next_count = dst + n;
    }
} // namespace hyt
