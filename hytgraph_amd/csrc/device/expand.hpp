// Frontier expansion shared by SSSP and delta-push PageRank.
//
// Degree binning for 64-lane wavefronts:
//   * light rows (out-degree <= light_max): a wavefront takes 64 consecutive
//     frontier entries, prefix-sums their degrees across lanes, and walks the
//     union of their edge lists 64 edges at a time, each lane finding its row
//     by binary search over the 64 prefix values in LDS. Consecutive lanes
//     read consecutive column/weight entries of a row (coalesced streaming),
//     and every lane has work regardless of the degree mix.
//   * heavy rows (out-degree > light_max): split into `chunk`-edge pieces
//     appended to a chunk queue (one atomic per wavefront); a second kernel
//     gives each piece to one wavefront, so a hub's edges spread over the GPU.
// The expand kernels append nothing to the next frontier: updates use
// non-returning atomics and mark touched vertices with plain stores; a
// separate compaction pass (compact.hpp) builds the next queue.
//
// Op supplies the per-algorithm parts:
//   uint32_t row_value(uint32_t i, uint32_t u, uint64_t deg)  value carried by row
//   void expanded(uint32_t u, uint32_t value, bool whole_row)  row u walks with value
//   void expanded_chunked(uint32_t u, bool whole_row)  row u goes out as chunks
//   void deferred(uint32_t u)  row u (above light_max) waits this round
//   uint32_t chunk_value(uint32_t u, uint32_t queued, bool record)  value a
//       chunk of row u walks with (queued = the row's value at queue time)
//   bool load(uint64_t pos, uint32_t value, uint32_t& v, uint32_t& c) read edge
//   uint32_t peek(uint32_t v)                          read destination v's state
//   void apply(uint32_t v, uint32_t c, uint32_t seen)  update v (seen = peek(v))
// load returns false to drop the edge (e.g. u32 overflow, flagged by Op).
//
// RowSrc says where a frontier row's edges are: GraphRows reads the CSR
// offsets and clips them to an edge window [lo, hi) (positions relative to
// lo: the whole graph in HBM, a window of the streaming ring, or the mapped
// host graph for zero-copy); PackedRows reads host-compacted rows;
// CachedRows reads the rows of many cached partitions through a segment
// table. Op's column/weight pointers address the same window. get() also
// names the row's value index (the row's index in the frontier arrays).
#pragma once

#include "kern_common.hpp"

namespace hyt
{
    template <class OffT>
    struct GraphRows
    {
        const OffT *off;
        const std::uint32_t *q;
        std::uint64_t lo, hi; // edge window

        __device__ void get(std::uint32_t i, std::uint32_t &u, std::uint32_t &vi, std::uint64_t &begin,
                            std::uint64_t &count, std::uint64_t &full) const
        {
            vi = i;
            u = q[i];
            const std::uint64_t b = off[u], e = off[u + 1];
            full = e - b;
            const std::uint64_t cb = b > lo ? b : lo;
            const std::uint64_t ce = e < hi ? e : hi;
            count = ce > cb ? ce - cb : 0;
            begin = count ? cb - lo : 0;
        }
    };

    struct PackedRows
    {
        const std::uint32_t *rows;  // vertex of packed row i
        const std::uint64_t *index; // packed row i spans [index[i], index[i+1])

        __device__ void get(std::uint32_t i, std::uint32_t &u, std::uint32_t &vi, std::uint64_t &begin,
                            std::uint64_t &count, std::uint64_t &full) const
        {
            vi = i;
            u = rows[i];
            begin = index[i];
            count = index[i + 1] - begin;
            full = count;
        }
    };

    // Resident partitions of one round expanded by one launch. Segment s
    // covers queue rows [qlo, qlo + n) - whole rows of one partition, held in
    // a cache unit so that edge b of the partition sits at unit_pos + (b -
    // edge_begin) of the unit area; rows_before is the exclusive prefix of n.
    struct CacheSeg
    {
        std::uint32_t qlo, rows_before;
        std::uint64_t edge_begin, unit_pos;
    };

    // Segment of launch row i: the last s with rows_before <= i.
    __device__ inline const CacheSeg &seg_of(const CacheSeg *seg, std::uint32_t nseg, std::uint32_t i)
    {
        std::uint32_t lo = 0, hi = nseg;
        while (hi - lo > 1)
        {
            const std::uint32_t mid = (lo + hi) >> 1;
            if (seg[mid].rows_before <= i)
                lo = mid;
            else
                hi = mid;
        }
        return seg[lo];
    }

    template <class OffT>
    struct CachedRows
    {
        const OffT *off;
        const std::uint32_t *q;
        const CacheSeg *seg;
        std::uint32_t nseg;

        __device__ void get(std::uint32_t i, std::uint32_t &u, std::uint32_t &vi, std::uint64_t &begin,
                            std::uint64_t &count, std::uint64_t &full) const
        {
            const CacheSeg &sg = seg_of(seg, nseg, i);
            vi = sg.qlo + (i - sg.rows_before);
            u = q[vi];
            const std::uint64_t b = off[u];
            count = full = off[u + 1] - b;
            begin = sg.unit_pos + (b - sg.edge_begin);
        }
    };

    struct ExpandArgs
    {
        const std::uint32_t *q;   // frontier rows (GraphRows reads them through its own pointer)
        std::uint32_t F;          // their number
        // device-driven rounds: the number of rows is read here at kernel
        // start (the grid then loops over 256-row tiles), F is ignored
        const std::uint32_t *F_dev = nullptr;
        Counters *ctr;
        HeavyChunk *heavy;
        std::uint64_t heavy_cap;
        std::uint32_t light_max;
        std::uint32_t chunk;
        // 0: every row; 1: only queue the chunks of rows above light_max;
        // 2: only walk the rows up to light_max
        std::uint32_t pass = 0;
        // rows above light_max whose value exceeds the bound wait (Op::
        // deferred) instead of going out as chunks; the bound is read on the
        // device when heavy_bound_dev is set (device-driven rounds)
        std::uint32_t heavy_bound = 0xFFFFFFFFU;
        const std::uint32_t *heavy_bound_dev = nullptr;
        // light rows (pass 2) of more than wait_degree edges wait under the
        // same bound (0xFFFFFFFF: light rows never wait)
        std::uint32_t wait_degree = 0xFFFFFFFFU;
    };

    // Processes `kUnroll` x 64 edges per iteration: all loads first, then all
    // updates, so each lane has kUnroll independent memory chains in flight.
    template <class Op, class PosFn>
    __device__ __forceinline__ void walk_edges(const Op &op, std::uint32_t total, PosFn pos_of)
    {
        const int lane = lane_id();
        for (std::uint32_t base = 0; base < total; base += kWave * kUnroll)
        {
            std::uint32_t vv[kUnroll], cc[kUnroll];
            bool ok[kUnroll];
#pragma unroll
            for (int k = 0; k < kUnroll; ++k)
            {
                const std::uint32_t e = base + static_cast<std::uint32_t>(k * kWave + lane);
                ok[k] = false;
                vv[k] = 0;
                cc[k] = 0;
                if (e < total)
                {
                    std::uint64_t pos;
                    std::uint32_t value;
                    pos_of(e, pos, value);
                    ok[k] = op.load(pos, value, vv[k], cc[k]);
                }
            }
            std::uint32_t seen[kUnroll];
#pragma unroll
            for (int k = 0; k < kUnroll; ++k)
                seen[k] = ok[k] ? op.peek(vv[k]) : 0U;
#pragma unroll
            for (int k = 0; k < kUnroll; ++k)
                if (ok[k])
                    op.apply(vv[k], cc[k], seen[k]);
        }
    }

    // Queues the chunks of this workgroup's heavy rows (nch chunks for a
    // thread's row, 0 for light rows; chunk c of a row covers its edges
    // [c * chunk, (c + 1) * chunk) from begin) with one atomic per workgroup:
    // every thread of the block calls it (s_h: kWavesPerBlock + 1 LDS words;
    // s_cex: this wavefront's 64 LDS words). One atomic per wavefront on the
    // single chunk counter serialised in the light kernel's chunk-queuing
    // pass (RMAT 24: ~100 K same-address atomics in a large round). Each
    // wavefront writes its rows' chunks together, 64 slots per step (a lane
    // finds the row of a slot by binary search over the rows' first slots),
    // so a row of thousands of chunks does not leave one lane writing them.
    __device__ __forceinline__ void queue_heavy(std::uint32_t nch, Counters *ctr, HeavyChunk *heavy,
                                                std::uint64_t heavy_cap, std::uint32_t *s_h, std::uint32_t *s_cex,
                                                std::uint32_t value, std::uint64_t count, std::uint64_t begin,
                                                std::uint32_t row, std::uint32_t record, std::uint32_t chunk)
    {
        const int lane = lane_id();
        const int wid = static_cast<int>(threadIdx.x) / kWave;
        const std::uint32_t hincl = wave_incl_scan(nch);
        const std::uint32_t wtotal = __shfl(hincl, kWave - 1, kWave);
        s_cex[lane] = hincl - nch; // the wavefront's first slot of each row
        if (lane == kWave - 1)
            s_h[wid] = hincl;
        __syncthreads();
        std::uint32_t before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w)
        {
            if (w < wid)
                before += s_h[w];
            total += s_h[w];
        }
        if (total == 0)
            return; // uniform across the workgroup
        if (threadIdx.x == 0)
            s_h[kWavesPerBlock] = atomicAdd(&ctr->heavy_count, total);
        __syncthreads();
        const std::uint32_t hbase = s_h[kWavesPerBlock];
        // the round fails on the host when the queue is full; until then the
        // chunk kernel reads min(heavy_count, cap) slots, so the reserved
        // ones that exist hold empty chunks (no edges, no record) rather
        // than stale bytes
        const bool full = static_cast<std::uint64_t>(hbase) + total > heavy_cap;
        if (full && threadIdx.x == 0)
            atomicOr(&ctr->heavy_overflow, 1U);
        const std::uint64_t wbase = static_cast<std::uint64_t>(hbase) + before;
        // a loop uniform across the wavefront: every lane takes part in the
        // shuffles that fetch a slot's row
        for (std::uint32_t base = 0; base < wtotal; base += kWave)
        {
            const std::uint32_t e = base + static_cast<std::uint32_t>(lane);
            const bool ok = e < wtotal;
            const int j = ok ? owner_of(s_cex, e) : 0;
            const std::uint32_t jv = __shfl(value, j, kWave);
            const std::uint64_t jc = __shfl(count, j, kWave);
            const std::uint64_t jb = __shfl(begin, j, kWave);
            const std::uint32_t jr = __shfl(row, j, kWave);
            const std::uint32_t jrec = __shfl(record, j, kWave);
            if (!ok)
                continue;
            const std::uint64_t slot = wbase + e;
            if (full)
            {
                if (slot < heavy_cap)
                {
                    HeavyChunk h{};
                    heavy[slot] = h;
                }
                continue;
            }
            const std::uint64_t lo = static_cast<std::uint64_t>(e - s_cex[j]) * chunk;
            HeavyChunk h;
            h.value = jv;
            h.count = static_cast<std::uint32_t>(jc - lo < chunk ? jc - lo : chunk);
            h.begin = jb + lo;
            h.row = jr;
            h.record = jrec;
            heavy[slot] = h;
        }
    }

    // One 256-row tile of the queue per workgroup iteration; a grid sized
    // blocks_for(F) takes each tile once, a smaller grid (device-driven
    // rounds, F read on the device) loops over them.
    template <class RowSrc, class Op>
    __global__ __launch_bounds__(kBlock) void expand_light(RowSrc rs, Op op, ExpandArgs a)
    {
        __shared__ std::uint32_t s_ex[kWavesPerBlock][kWave];
        __shared__ std::uint64_t s_begin[kWavesPerBlock][kWave];
        __shared__ std::uint32_t s_val[kWavesPerBlock][kWave];
        __shared__ unsigned long long s_edges[kWavesPerBlock];
        __shared__ std::uint32_t s_h[kWavesPerBlock + 1];

        const int lane = lane_id();
        const int wid = static_cast<int>(threadIdx.x) / kWave;
        const std::uint32_t F = a.F_dev != nullptr ? *a.F_dev : a.F;
        const std::uint32_t hbound = a.heavy_bound_dev != nullptr ? *a.heavy_bound_dev : a.heavy_bound;
        const std::uint32_t tiles = (F + kBlock - 1) / kBlock;
        unsigned long long wedges = 0; // this wavefront's light edges (uniform across its lanes)
        // a block-uniform loop: every wavefront meets the same barriers; LDS
        // slices are per wavefront, so a wavefront's next-tile writes only
        // follow its own reads
        for (std::uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x)
        {
            const std::uint32_t i = tile * kBlock + threadIdx.x;
            std::uint32_t deg = 0, value = 0, nch = 0, row = 0;
            std::uint64_t begin = 0, count = 0;
            bool whole = false;
            if (i < F)
            {
                std::uint32_t u, vi;
                std::uint64_t full;
                rs.get(i, u, vi, begin, count, full);
                row = u;
                whole = count == full;
                const bool chunked = count > a.light_max;
                if (chunked && a.pass != 2)
                {
                    value = op.row_value(vi, u, full);
                    if (value > hbound)
                        op.deferred(u);
                    else
                    {
                        nch = static_cast<std::uint32_t>((count + a.chunk - 1) / a.chunk);
                        op.expanded_chunked(u, whole);
                    }
                }
                else if (!chunked && a.pass != 1)
                {
                    value = op.row_value(vi, u, full);
                    if (count > a.wait_degree && value > hbound)
                        op.deferred(u);
                    else
                    {
                        deg = static_cast<std::uint32_t>(count);
                        op.expanded(u, value, whole);
                    }
                }
            }
            // the chunk-queuing passes (block-uniform: the pass is a kernel argument)
            if (a.pass != 2)
                queue_heavy(nch, a.ctr, a.heavy, a.heavy_cap, s_h, s_ex[wid], value, count, begin, row,
                            whole ? 1U : 0U, a.chunk);

            const std::uint32_t incl = wave_incl_scan(deg);
            const std::uint32_t total = __shfl(incl, kWave - 1, kWave);
            s_ex[wid][lane] = incl - deg;
            s_begin[wid][lane] = begin;
            s_val[wid][lane] = value;
            __syncthreads();
            if (total == 0)
                continue;
            wedges += total;
            const std::uint32_t *ex = s_ex[wid];
            const std::uint64_t *bg = s_begin[wid];
            const std::uint32_t *vl = s_val[wid];
            walk_edges(op, total, [&](std::uint32_t e, std::uint64_t &pos, std::uint32_t &val) {
                const int j = owner_of(ex, e);
                pos = bg[j] + (e - ex[j]);
                val = vl[j];
            });
        }
        // one edge-count atomic per workgroup: per-wavefront atomics on the
        // one counter serialise (measured: ~9% of an RMAT-24 SSSP round)
        if (lane == 0)
            s_edges[wid] = wedges;
        __syncthreads();
        if (threadIdx.x == 0)
        {
            unsigned long long sum = 0;
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; ++w)
                sum += s_edges[w];
            if (sum)
                add_edges(a.ctr, sum);
        }
    }

    // The whole graph resident in HBM: window [0, 2^64).
    template <class OffT>
    __host__ __device__ inline GraphRows<OffT> whole_graph(const OffT *off, const std::uint32_t *q)
    {
        return GraphRows<OffT>{off, q, 0, ~0ULL};
    }

    template <class Op>
    __global__ __launch_bounds__(kBlock) void expand_heavy(Op op, ExpandArgs a)
    {
        const int lane = lane_id();
        const std::uint32_t nh = a.ctr->heavy_count;
        const std::uint32_t n = nh < a.heavy_cap ? nh : static_cast<std::uint32_t>(a.heavy_cap);
        const std::uint32_t waves = gridDim.x * kWavesPerBlock;
        unsigned long long wedges = 0; // one edge-count atomic per wavefront
        for (std::uint32_t c = (blockIdx.x * kBlock + threadIdx.x) / kWave; c < n; c += waves)
        {
            const HeavyChunk h = a.heavy[c];
            wedges += h.count;
            const std::uint32_t value = op.chunk_value(h.row, h.value, h.record != 0U);
            walk_edges(op, h.count, [&](std::uint32_t e, std::uint64_t &pos, std::uint32_t &val) {
                pos = h.begin + e;
                val = value;
            });
        }
        if (lane == 0 && wedges)
            add_heavy_edges(a.ctr, wedges);
    }
} // namespace hyt
