// K-lane frontier kernels shared by the batched [V][K] queries and the
// continuous-batching lane sessions.
//
// State is row-major [V][K]: the K query values of a vertex are contiguous,
// so one edge (u, v) is handled by a group of KP = next_pow2(K) lanes that
// read and update K consecutive words of row u and row v - coalesced, and
// the column/weight of the edge are read once for all K queries. A wavefront
// holds G = 64 / KP such groups, i.e. walks G edges per step.
//
// Frontier: a queue of rows plus, per row, the u64 mask of queries for which
// the row is active. During a round, the lanes that improve v OR their bits
// into mark[v] (non-returning atomic, one per edge group); the compaction pass
// scans mark[], moves every non-zero mask into the next queue and clears it.
// A queue may hold one vertex more than once with disjoint masks (a query
// admitted into a running batch appends its source row).
#pragma once

#include "compact.hpp"
#include "expand.hpp"
#include "mass.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace hyt
{
    namespace lanes
    {
        namespace // each translation unit gets its own kernels
        {
        struct HeavyChunkK
        {
            std::uint32_t row;   // SSSP: vertex id; PPR: queue index
            std::uint32_t count; // edges in the chunk
            std::uint64_t begin; // first edge
            std::uint64_t mask;  // active lanes of the row
            std::uint32_t scale; // PPR: damping / outdeg (float bits)
            std::uint32_t pad;
        };

        // Round state of device-driven rounds (rounds captured into a HIP
        // graph and replayed without host round trips): the queue size and
        // the form of the current round are read here by every kernel.
        struct RoundState
        {
            std::uint32_t F;     // rows of the current queue
            std::uint32_t form;  // HYT_LANES_GROUP or HYT_LANES_ITEMS
            std::uint64_t pairs; // (row, query) pairs of the current queue
            std::uint64_t rounds, frontier_total, edges_total, updates_total;
            std::uint32_t done;      // a round found the queue empty (or the round limit)
            std::uint32_t truncated; // stopped by the round limit with work left
        };

        struct KArgs
        {
            const RoundState *rs; // non-null: F and the form come from the device
            std::uint32_t F;
            Counters *ctr;
            HeavyChunkK *heavy;
            std::uint64_t heavy_cap;
            std::uint32_t light_max;
            std::uint32_t chunk;
            std::uint32_t K;
            unsigned long long *mark;
        };

        // Row sources: row i of a launch names its vertex u, its queue index
        // gi (the row of the PPR snapshot), its query mask, the edges to walk
        // (begin, count; positions relative to the Op's column/weight base)
        // and its full out-degree (the PPR share).
        template <class OffT>
        struct QueueRowsK // whole rows of the queue, positions absolute
        {
            using Pos = OffT;
            const OffT *off;
            const std::uint32_t *q;
            const unsigned long long *qmask;

            __device__ void get(std::uint32_t i, std::uint32_t &u, std::uint32_t &gi, std::uint64_t &mask, Pos &begin,
                                std::uint64_t &count, std::uint64_t &deg) const
            {
                gi = i;
                u = q[i];
                mask = qmask[i];
                begin = off[u];
                deg = count = static_cast<std::uint64_t>(off[u + 1] - begin);
            }
        };

        // Queue rows [qa, qa + F), or (rows != nullptr) the rows of a list with
        // their queue indices, clipped to the edge window [lo, hi): one
        // partition in a device slot, a window of the streaming ring, or the
        // mapped host graph.
        template <class OffT>
        struct WindowRowsK
        {
            using Pos = OffT;
            const OffT *off;
            const std::uint32_t *q;
            const unsigned long long *qmask;
            const std::uint32_t *rows; // optional row list
            const std::uint32_t *idx;  // queue index of each listed row
            std::uint32_t qa;
            std::uint64_t lo, hi;

            __device__ void get(std::uint32_t i, std::uint32_t &u, std::uint32_t &gi, std::uint64_t &mask, Pos &begin,
                                std::uint64_t &count, std::uint64_t &deg) const
            {
                if (rows != nullptr)
                {
                    u = rows[i];
                    gi = idx[i];
                }
                else
                {
                    gi = qa + i;
                    u = q[gi];
                }
                mask = qmask[gi];
                const std::uint64_t b = off[u], e = off[u + 1];
                deg = e - b;
                const std::uint64_t cb = b > lo ? b : lo;
                const std::uint64_t ce = e < hi ? e : hi;
                count = ce > cb ? ce - cb : 0;
                begin = static_cast<Pos>(count ? cb - lo : 0);
            }
        };

        // Rows of many cached partitions (segment table, see CacheSeg).
        template <class OffT>
        struct CachedRowsK
        {
            using Pos = std::uint64_t;
            const OffT *off;
            const std::uint32_t *q;
            const unsigned long long *qmask;
            const CacheSeg *seg;
            std::uint32_t nseg;

            __device__ void get(std::uint32_t i, std::uint32_t &u, std::uint32_t &gi, std::uint64_t &mask, Pos &begin,
                                std::uint64_t &count, std::uint64_t &deg) const
            {
                const CacheSeg &sg = seg_of(seg, nseg, i);
                gi = sg.qlo + (i - sg.rows_before);
                u = q[gi];
                mask = qmask[gi];
                const std::uint64_t b = off[u];
                deg = count = off[u + 1] - b;
                begin = sg.unit_pos + (b - sg.edge_begin);
            }
        };

        // Rows packed on the host: row i is vertex rows[i], queue index idx[i],
        // edges [index[i], index[i+1]) of the packed arrays.
        struct PackedRowsK
        {
            using Pos = std::uint64_t;
            const std::uint32_t *rows;
            const std::uint32_t *idx;
            const std::uint64_t *index;
            const unsigned long long *qmask;

            __device__ void get(std::uint32_t i, std::uint32_t &u, std::uint32_t &gi, std::uint64_t &mask, Pos &begin,
                                std::uint64_t &count, std::uint64_t &deg) const
            {
                u = rows[i];
                gi = idx[i];
                mask = qmask[gi];
                begin = index[i];
                deg = count = index[i + 1] - begin;
            }
        };

        // Work split of item-form rows: a row's work is edges x c (its active
        // queries). It is light when the work is at most 8 x light_max (2048
        // items by default), so one wavefront's 64 light rows stay within a few
        // chunks of work whatever the mask density; heavier rows are cut into
        // chunks of max(16, 4 x chunk / c) edges (4096 items by default). The
        // group form keeps the edge split (light <= light_max edges, chunks of
        // `chunk` edges): its rows carry KP lanes each, and finer chunks there
        // measured slower (K = 64 SSSP on RMAT 24: 199 ms -> 1.39 s).
        __device__ __forceinline__ bool light_work(std::uint64_t deg, std::uint32_t w, const KArgs &a)
        {
            return deg * w <= 8ULL * a.light_max;
        }
        __device__ __forceinline__ std::uint32_t chunk_edges(std::uint32_t w, const KArgs &a)
        {
            const std::uint32_t ce = (4U * a.chunk) / (w ? w : 1U);
            return ce < 16U ? 16U : ce;
        }

        // Queues this workgroup's heavy-row chunks (nch per thread) with one
        // atomic per workgroup; every thread of the block calls it (s_h:
        // kWavesPerBlock + 1 LDS words). Per-wavefront atomics on the one
        // chunk counter serialise (see queue_heavy). When the queue cannot
        // take them all, the reserved slots that exist are written as empty
        // chunks and false is returned (uniformly): the rows are then walked
        // by their own wavefronts (slower, never wrong).
        template <class MakeFn>
        __device__ __forceinline__ bool queue_chunks_k(std::uint32_t nch, const KArgs &a, std::uint32_t *s_h,
                                                       MakeFn make)
        {
            const int lane = lane_id();
            const int wid = static_cast<int>(threadIdx.x) / kWave;
            const std::uint32_t hincl = wave_incl_scan(nch);
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
                return true; // uniform across the workgroup
            if (threadIdx.x == 0)
                s_h[kWavesPerBlock] = atomicAdd(&a.ctr->heavy_count, total);
            __syncthreads();
            const std::uint32_t hbase = s_h[kWavesPerBlock];
            const std::uint64_t first = static_cast<std::uint64_t>(hbase) + before + hincl - nch;
            if (static_cast<std::uint64_t>(hbase) + total <= a.heavy_cap)
            {
                for (std::uint32_t c = 0; c < nch; ++c)
                    a.heavy[first + c] = make(c);
                return true;
            }
            for (std::uint32_t c = 0; c < nch; ++c)
                if (first + c < a.heavy_cap)
                {
                    HeavyChunkK h{};
                    h.count = 0;
                    a.heavy[first + c] = h;
                }
            return false;
        }

        template <int KP>
        __device__ __forceinline__ std::uint64_t group_bits(std::uint64_t ballot, int group)
        {
            if constexpr (KP == 64)
                return ballot;
            else
                return (ballot >> (group * KP)) & ((1ULL << KP) - 1ULL);
        }

        // Walks `total` edges, G per step, kUnroll steps in flight.
        template <int KP, class Op, class PosFn>
        __device__ __forceinline__ void walk_edges_k(const Op &op, KArgs &a, std::uint32_t total, PosFn pos_of)
        {
            constexpr int G = kWave / KP;
            const int lane = lane_id();
            const int group = lane / KP;
            const int kl = lane % KP;
            for (std::uint32_t base = 0; base < total; base += G * kUnroll)
            {
                std::uint32_t vv[kUnroll], ww[kUnroll], row[kUnroll], scale[kUnroll];
                typename Op::Val cand[kUnroll], seen[kUnroll];
                std::uint64_t m[kUnroll];
                bool ok[kUnroll];
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                {
                    const std::uint32_t e = base + static_cast<std::uint32_t>(k * G + group);
                    m[k] = 0;
                    vv[k] = ww[k] = row[k] = scale[k] = 0;
                    if (e < total)
                    {
                        std::uint64_t pos;
                        pos_of(e, pos, row[k], m[k], scale[k]);
                        op.load(pos, vv[k], ww[k]);
                    }
                }
                // reads of every in-flight edge first, then the atomics, so no
                // read waits behind an earlier edge's atomic to the same array
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                {
                    const bool act = static_cast<std::uint32_t>(kl) < a.K && ((m[k] >> kl) & 1ULL);
                    cand[k] = seen[k] = 0;
                    ok[k] = act && op.prepare(row[k], scale[k], vv[k], ww[k], kl, cand[k], seen[k]);
                }
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                    if (ok[k])
                        op.commit(vv[k], kl, cand[k], seen[k]);
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                {
                    const bool imp = ok[k] && op.improved(cand[k], seen[k]);
                    const std::uint64_t gb = group_bits<KP>(__ballot(imp), group);
                    if (kl == 0 && gb != 0)
                        atomicOr(&a.mark[vv[k]], static_cast<unsigned long long>(gb));
                }
            }
        }

        // Queue size of this round; false when a device-driven round runs
        // the other form (the kernel then has nothing to do).
        __device__ __forceinline__ bool round_rows(const KArgs &a, std::uint32_t form, std::uint32_t &F)
        {
            if (a.rs == nullptr)
            {
                F = a.F;
                return true;
            }
            F = a.rs->F;
            return a.rs->form == form;
        }

        // Adds the workgroup's edge count (each wavefront's lane 0 holds its
        // own) with one atomic: per-wavefront atomics on the one counter
        // serialise. Every thread of the workgroup must call it.
        __device__ __forceinline__ void add_block_edges(unsigned long long *s_edges, unsigned long long wedges,
                                                        Counters *ctr)
        {
            const int wid = static_cast<int>(threadIdx.x) / kWave;
            if (lane_id() == 0)
                s_edges[wid] = wedges;
            __syncthreads();
            if (threadIdx.x == 0)
            {
                unsigned long long sum = 0;
#pragma unroll
                for (int w = 0; w < kWavesPerBlock; ++w)
                    sum += s_edges[w];
                if (sum)
                    add_edges(ctr, sum);
            }
        }

        template <class RowSrc, int KP, class Op>
        __global__ __launch_bounds__(kBlock) void expandk_light(RowSrc rsrc, Op op, KArgs a)
        {
            using Pos = typename RowSrc::Pos;
            __shared__ std::uint32_t s_ex[kWavesPerBlock][kWave];
            __shared__ Pos s_begin[kWavesPerBlock][kWave];
            __shared__ std::uint32_t s_row[kWavesPerBlock][kWave];
            __shared__ std::uint32_t s_scale[kWavesPerBlock][kWave];
            __shared__ unsigned long long s_mask[kWavesPerBlock][kWave];
            __shared__ unsigned long long s_edges[kWavesPerBlock];
            __shared__ std::uint32_t s_h[kWavesPerBlock + 1];

            std::uint32_t F;
            if (!round_rows(a, HYT_LANES_GROUP, F))
                return;
            unsigned long long wedges = 0; // this wavefront's edges (lane 0)
            const int lane = lane_id();
            const int wid = static_cast<int>(threadIdx.x) / kWave;
            const std::uint32_t ce = a.chunk;
            // a block-uniform loop over 256-row tiles (one tile per block when
            // the grid is sized from F)
            const std::uint32_t tiles = (F + kBlock - 1) / kBlock;
            for (std::uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x)
            {
                const std::uint32_t i = tile * kBlock + threadIdx.x;
                std::uint32_t deg = 0, nch = 0, row = 0, scale = 0;
                std::uint64_t mask = 0, full = 0;
                Pos begin = 0;
                if (i < F)
                {
                    std::uint32_t u, gi;
                    std::uint64_t rdeg;
                    rsrc.get(i, u, gi, mask, begin, full, rdeg);
                    row = Op::kRowIsQueueIndex ? gi : u;
                    scale = op.row_scale(rdeg);
                    if (full <= a.light_max)
                        deg = static_cast<std::uint32_t>(full);
                    else
                        nch = static_cast<std::uint32_t>((full + ce - 1) / ce);
                }
                const bool queued = queue_chunks_k(nch, a, s_h, [&](std::uint32_t c) {
                    const std::uint64_t lo = static_cast<std::uint64_t>(c) * ce;
                    HeavyChunkK h;
                    h.row = row;
                    h.count = static_cast<std::uint32_t>((full - lo < ce) ? (full - lo) : ce);
                    h.begin = static_cast<std::uint64_t>(begin) + lo;
                    h.mask = mask;
                    h.scale = scale;
                    h.pad = 0;
                    return h;
                });
                if (!queued && nch)
                {
                    // chunk queue full: walk the row here (rows of 2^24 edges or
                    // more cannot be - 64 of them must fit a u32 - and are reported)
                    if (full >= (1ULL << 24))
                        atomicOr(&a.ctr->heavy_overflow, 1U);
                    else
                        deg = static_cast<std::uint32_t>(full);
                }
                const std::uint32_t incl = wave_incl_scan(deg);
                const std::uint32_t total = __shfl(incl, kWave - 1, kWave);
                s_ex[wid][lane] = incl - deg;
                s_begin[wid][lane] = begin;
                s_row[wid][lane] = row;
                s_scale[wid][lane] = scale;
                s_mask[wid][lane] = mask;
                __syncthreads();
                if (total == 0)
                    continue; // every wavefront still meets the next tile's barrier
                wedges += total;
                const std::uint32_t *ex = s_ex[wid];
                walk_edges_k<KP>(op, a, total,
                                 [&](std::uint32_t e, std::uint64_t &pos, std::uint32_t &r, std::uint64_t &m,
                                     std::uint32_t &sc) {
                                     const int j = owner_of(ex, e);
                                     pos = static_cast<std::uint64_t>(s_begin[wid][j]) + (e - ex[j]);
                                     r = s_row[wid][j];
                                     m = s_mask[wid][j];
                                     sc = s_scale[wid][j];
                                 });
            }
            add_block_edges(s_edges, wedges, a.ctr);
        }

        template <int KP, class Op>
        __global__ __launch_bounds__(kBlock) void expandk_heavy(Op op, KArgs a)
        {
            std::uint32_t F;
            if (!round_rows(a, HYT_LANES_GROUP, F))
                return;
            const int lane = lane_id();
            const std::uint32_t nh = a.ctr->heavy_count;
            const std::uint32_t n = nh < a.heavy_cap ? nh : static_cast<std::uint32_t>(a.heavy_cap);
            const std::uint32_t waves = gridDim.x * kWavesPerBlock;
            unsigned long long wedges = 0; // one edge-count atomic per wavefront
            for (std::uint32_t c = (blockIdx.x * kBlock + threadIdx.x) / kWave; c < n; c += waves)
            {
                const HeavyChunkK h = a.heavy[c];
                wedges += h.count;
                walk_edges_k<KP>(op, a, h.count,
                                 [&](std::uint32_t e, std::uint64_t &pos, std::uint32_t &r, std::uint64_t &m,
                                     std::uint32_t &sc) {
                                     pos = h.begin + e;
                                     r = h.row;
                                     m = h.mask;
                                     sc = h.scale;
                                 });
            }
            if (lane == 0 && wedges)
                add_edges(a.ctr, wedges);
        }

        // ------------------------------------------------- sparse lane masks
        // When few queries are active per frontier row, a group of KP lanes
        // per edge leaves most lanes idle. The item form gives every lane one
        // (edge, active query) pair instead: a row with c active queries and
        // d edges contributes d*c items, consecutive lanes take the c queries
        // of one edge (same column/weight, neighbouring [V][K] words), and the
        // wavefront stays full whatever the mask density.

        // Chunk-queue capacity of K-lane rounds: every group-form chunk (rows
        // above light_max edges, `chunk`-edge pieces) at the given split, plus
        // V/16 + E/256 for item-form chunks; a full queue only moves rows back
        // into the light walk.
        inline std::uint64_t lane_chunk_capacity(std::uint64_t V, std::uint64_t E, std::uint32_t light_max,
                                                 std::uint32_t chunk)
        {
            return E / chunk + E / (static_cast<std::uint64_t>(light_max) + 1) + E / 256 + V / 16 + 64;
        }

        // Item-form chunks one more lane can add at a chunk length (see
        // lane_chunk_capacity_k): lane sessions and hybrid K-lane runs size
        // their queue as lane_chunk_capacity + K of these.
        inline std::uint64_t lane_item_chunks_per_lane(std::uint64_t E, std::uint32_t chunk)
        {
            return E / (4ULL * chunk) + E / (128ULL * chunk) + 1;
        }

        // Capacity that holds every chunk of a round of K lanes: group-form
        // chunks (E / chunk + one per row above light_max) or item-form
        // chunks, which are max(16, 4 chunk / c) edges for c active queries,
        // so at most E K / (4 chunk) + one per row. Batched calls size their
        // queue with it; a queue below it moves hub rows back into the light
        // walk, where one wavefront takes a whole hub (measured: K = 64 SSSP
        // on RMAT 24 at split 32 / 256, 98 instead of 400 queries/s).
        inline std::uint64_t lane_chunk_capacity_k(std::uint64_t V, std::uint64_t E, std::uint32_t light_max,
                                                   std::uint32_t chunk, std::uint32_t K)
        {
            const std::uint64_t rows = E / (static_cast<std::uint64_t>(light_max) + 1) + 1;
            const std::uint64_t group = E / chunk + rows;
            // (+1/32: the chunk length is 4 chunk / c rounded down)
            const std::uint64_t items = E * K / (4ULL * chunk) + E * K / (128ULL * chunk) + rows;
            return std::max(group, items) + V / 16 + 64;
        }

        // Position of the r-th set bit (0-based) of m; m has more than r set bits.
        __device__ __forceinline__ std::uint32_t nth_set_bit(std::uint64_t m, std::uint32_t r)
        {
            std::uint32_t k = 0;
#pragma unroll
            for (int w = 32; w > 0; w >>= 1)
            {
                const std::uint64_t low = m & ((1ULL << w) - 1ULL);
                const std::uint32_t c = static_cast<std::uint32_t>(__popcll(low));
                if (r >= c)
                {
                    r -= c;
                    m >>= w;
                    k += static_cast<std::uint32_t>(w);
                }
                else
                    m = low;
            }
            return k;
        }

        // Walks `total` items, kUnroll per lane in flight; item_of(e, pos, row,
        // scale, k) names the edge position, row, row scale and query of item e.
        // An improved (v, k) sets bit k of mark[v] unless it is already set.
        template <class Op, class ItemFn>
        __device__ __forceinline__ void walk_items_k(const Op &op, KArgs &a, std::uint32_t total, ItemFn item_of)
        {
            const int lane = lane_id();
            for (std::uint32_t base = 0; base < total; base += kWave * kUnroll)
            {
                std::uint32_t vv[kUnroll], ww[kUnroll], row[kUnroll], scale[kUnroll], kk[kUnroll];
                typename Op::Val cand[kUnroll], seen[kUnroll];
                bool ok[kUnroll];
#pragma unroll
                for (int u = 0; u < kUnroll; ++u)
                {
                    const std::uint32_t e = base + static_cast<std::uint32_t>(u * kWave + lane);
                    ok[u] = e < total;
                    vv[u] = ww[u] = row[u] = scale[u] = kk[u] = 0;
                    if (ok[u])
                    {
                        std::uint64_t pos;
                        item_of(e, pos, row[u], scale[u], kk[u]);
                        op.load(pos, vv[u], ww[u]);
                    }
                }
                // as in walk_edges_k: every read, then every atomic, then the
                // mark words, then the mark atomics
#pragma unroll
                for (int u = 0; u < kUnroll; ++u)
                {
                    cand[u] = seen[u] = 0;
                    ok[u] = ok[u] &&
                            op.prepare(row[u], scale[u], vv[u], ww[u], static_cast<int>(kk[u]), cand[u], seen[u]);
                }
#pragma unroll
                for (int u = 0; u < kUnroll; ++u)
                    if (ok[u])
                        op.commit(vv[u], static_cast<int>(kk[u]), cand[u], seen[u]);
                unsigned long long word[kUnroll];
#pragma unroll
                for (int u = 0; u < kUnroll; ++u)
                {
                    ok[u] = ok[u] && op.improved(cand[u], seen[u]);
                    word[u] = ok[u] ? a.mark[vv[u]] : ~0ULL;
                }
#pragma unroll
                for (int u = 0; u < kUnroll; ++u)
                {
                    const unsigned long long bit = 1ULL << kk[u];
                    if ((word[u] & bit) == 0ULL)
                        atomicOr(a.mark + vv[u], bit);
                }
            }
        }

        template <class RowSrc, class Op>
        __global__ __launch_bounds__(kBlock) void expandk_light_items(RowSrc rsrc, Op op, KArgs a)
        {
            using Pos = typename RowSrc::Pos;
            __shared__ std::uint32_t s_ex[kWavesPerBlock][kWave];
            __shared__ Pos s_begin[kWavesPerBlock][kWave];
            __shared__ std::uint32_t s_row[kWavesPerBlock][kWave];
            __shared__ std::uint32_t s_scale[kWavesPerBlock][kWave];
            __shared__ std::uint32_t s_c[kWavesPerBlock][kWave];
            __shared__ unsigned long long s_mask[kWavesPerBlock][kWave];

            __shared__ unsigned long long s_edges[kWavesPerBlock];
            __shared__ std::uint32_t s_h[kWavesPerBlock + 1];
            std::uint32_t F;
            if (!round_rows(a, HYT_LANES_ITEMS, F))
                return;
            unsigned long long wedges = 0; // this wavefront's edges (lane 0)
            const int lane = lane_id();
            const int wid = static_cast<int>(threadIdx.x) / kWave;
            const std::uint32_t tiles = (F + kBlock - 1) / kBlock;
            for (std::uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x)
            {
                const std::uint32_t i = tile * kBlock + threadIdx.x;
                std::uint32_t items = 0, deg = 0, nch = 0, row = 0, scale = 0, c = 0, ce = 16;
                std::uint64_t mask = 0, full = 0;
                Pos begin = 0;
                if (i < F)
                {
                    std::uint32_t u, gi;
                    std::uint64_t rdeg;
                    rsrc.get(i, u, gi, mask, begin, full, rdeg);
                    c = static_cast<std::uint32_t>(__popcll(mask));
                    row = Op::kRowIsQueueIndex ? gi : u;
                    scale = op.row_scale(rdeg);
                    ce = chunk_edges(c, a);
                    if (light_work(full, c, a))
                    {
                        deg = static_cast<std::uint32_t>(full);
                        items = deg * c;
                    }
                    else
                        nch = static_cast<std::uint32_t>((full + ce - 1) / ce);
                }
                const bool queued = queue_chunks_k(nch, a, s_h, [&](std::uint32_t ch) {
                    const std::uint64_t lo = static_cast<std::uint64_t>(ch) * ce;
                    HeavyChunkK h;
                    h.row = row;
                    h.count = static_cast<std::uint32_t>((full - lo < ce) ? (full - lo) : ce);
                    h.begin = static_cast<std::uint64_t>(begin) + lo;
                    h.mask = mask;
                    h.scale = scale;
                    h.pad = 0;
                    return h;
                });
                if (!queued && nch)
                {
                    // chunk queue full: walk the row here (work of 2^24 items or
                    // more cannot be - 64 rows' work must fit a u32 - and is reported)
                    const std::uint64_t w = full * c;
                    if (w >= (1ULL << 24))
                        atomicOr(&a.ctr->heavy_overflow, 1U);
                    else
                    {
                        deg = static_cast<std::uint32_t>(full);
                        items = static_cast<std::uint32_t>(w);
                    }
                }
                const std::uint32_t incl = wave_incl_scan(items);
                const std::uint32_t total = __shfl(incl, kWave - 1, kWave);
                std::uint32_t dsum = deg;
    #pragma unroll
                for (int d = kWave / 2; d > 0; d >>= 1)
                    dsum += __shfl_xor(dsum, d, kWave);
                s_ex[wid][lane] = incl - items;
                s_begin[wid][lane] = begin;
                s_row[wid][lane] = row;
                s_scale[wid][lane] = scale;
                s_c[wid][lane] = c;
                s_mask[wid][lane] = mask;
                __syncthreads();
                if (total == 0)
                    continue; // every wavefront still meets the next tile's barrier
                wedges += dsum;
                const std::uint32_t *ex = s_ex[wid];
                walk_items_k(op, a, total,
                             [&](std::uint32_t e, std::uint64_t &pos, std::uint32_t &r, std::uint32_t &sc,
                                 std::uint32_t &k) {
                                 const int j = owner_of(ex, e);
                                 const std::uint32_t local = e - ex[j];
                                 const std::uint32_t cj = s_c[wid][j];
                                 const std::uint32_t edge = local / cj;
                                 k = nth_set_bit(s_mask[wid][j], local - edge * cj);
                                 pos = static_cast<std::uint64_t>(s_begin[wid][j]) + edge;
                                 r = s_row[wid][j];
                                 sc = s_scale[wid][j];
                             });
            }
            add_block_edges(s_edges, wedges, a.ctr);
        }

        template <class Op>
        __global__ __launch_bounds__(kBlock) void expandk_heavy_items(Op op, KArgs a)
        {
            std::uint32_t F;
            if (!round_rows(a, HYT_LANES_ITEMS, F))
                return;
            const int lane = lane_id();
            const std::uint32_t nh = a.ctr->heavy_count;
            const std::uint32_t n = nh < a.heavy_cap ? nh : static_cast<std::uint32_t>(a.heavy_cap);
            const std::uint32_t waves = gridDim.x * kWavesPerBlock;
            unsigned long long wedges = 0; // one edge-count atomic per wavefront
            for (std::uint32_t ci = (blockIdx.x * kBlock + threadIdx.x) / kWave; ci < n; ci += waves)
            {
                const HeavyChunkK h = a.heavy[ci];
                wedges += h.count;
                const std::uint32_t c = static_cast<std::uint32_t>(__popcll(h.mask));
                if (c == 0)
                    continue;
                walk_items_k(op, a, h.count * c,
                             [&](std::uint32_t e, std::uint64_t &pos, std::uint32_t &r, std::uint32_t &sc,
                                 std::uint32_t &k) {
                                 const std::uint32_t edge = e / c;
                                 k = nth_set_bit(h.mask, e - edge * c);
                                 pos = h.begin + edge;
                                 r = h.row;
                                 sc = h.scale;
                             });
            }
            if (lane == 0 && wedges)
                add_edges(a.ctr, wedges);
        }

        // Item form when asked, or (auto) when the queue's (row, query) pairs
        // fill under half of the group form's lane slots.
        inline bool use_items(std::uint32_t form, std::uint64_t pairs, std::uint64_t F, std::uint32_t KP)
        {
            if (form == HYT_LANES_GROUP)
                return false;
            if (form == HYT_LANES_ITEMS)
                return true;
            return pairs * 2 < F * static_cast<std::uint64_t>(KP);
        }

        // ---------------------------------------------------------- SSSP-K
        template <int WM>
        struct SsspKOp
        {
            using Val = std::uint32_t; // distance (u32 or float bits)
            static constexpr bool kRowIsQueueIndex = false;
            const std::uint32_t *__restrict__ col;
            const std::uint32_t *__restrict__ w;
            std::uint32_t *dist;
            std::uint32_t K;
            std::uint32_t *overflow;
            float damping = 0.0F; // unused (the PPR snapshot's parameter)

            __device__ std::uint32_t row_scale(std::uint64_t) const { return 0U; }
            __device__ void load(std::uint64_t pos, std::uint32_t &v, std::uint32_t &wt) const
            {
                v = col[pos];
                if constexpr (WM == kWU32 || WM == kWF32)
                    wt = w[pos];
                else if constexpr (WM == kWUnitU32)
                    wt = 1U;
                else
                    wt = 0x3F800000U;
            }
            // Candidate distance of lane kl for v through row u, and v's
            // current value (a value read early may be older and larger; that
            // only costs an extra atomic, it never skips a needed one).
            __device__ bool prepare(std::uint32_t u, std::uint32_t, std::uint32_t v, std::uint32_t wt, int kl,
                                    std::uint32_t &cand, std::uint32_t &seen) const
            {
                const std::uint32_t du = dist[static_cast<std::uint64_t>(u) * K + kl];
                if constexpr (WM == kWU32 || WM == kWUnitU32)
                {
                    const std::uint64_t s = static_cast<std::uint64_t>(du) + wt;
                    if (s >= 0xFFFFFFFFULL)
                    {
                        atomicOr(overflow, 1U);
                        return false;
                    }
                    cand = static_cast<std::uint32_t>(s);
                }
                else
                {
                    cand = __float_as_uint(__fadd_rn(__uint_as_float(du), __uint_as_float(wt)));
                }
                seen = dist[static_cast<std::uint64_t>(v) * K + kl];
                return true;
            }
            __device__ void commit(std::uint32_t v, int kl, std::uint32_t cand, std::uint32_t &seen) const
            {
                if (cand < seen)
                    atomicMin(dist + static_cast<std::uint64_t>(v) * K + kl, cand);
            }
            __device__ bool improved(std::uint32_t cand, std::uint32_t seen) const { return cand < seen; }
        };

        // ----------------------------------------------------------- PPR-K
        // Residuals, ranks and snapshots are mass words (mass.hpp): the
        // snapshot of lane kl at row i is already the share word per out-edge
        // of the residual it moved (pprk_snapshot, computed as in the
        // single-query form), and the returning integer atomic tells exactly
        // one adder that the residual crossed eps.
        struct PprKOp
        {
            using Val = unsigned long long; // mass word
            static constexpr bool kRowIsQueueIndex = true;
            const std::uint32_t *__restrict__ col;
            unsigned long long *res;
            const unsigned long long *snap; // [F][K] share words of the moved residuals
            std::uint32_t K;
            float damping;
            unsigned long long eps; // threshold word

            __device__ std::uint32_t row_scale(std::uint64_t) const { return 0U; }
            __device__ void load(std::uint64_t pos, std::uint32_t &v, std::uint32_t &wt) const
            {
                v = col[pos];
                wt = 0;
            }
            __device__ bool prepare(std::uint32_t i, std::uint32_t, std::uint32_t, std::uint32_t, int kl,
                                    unsigned long long &cand, unsigned long long &) const
            {
                cand = snap[static_cast<std::uint64_t>(i) * K + kl];
                return cand != 0ULL;
            }
            __device__ void commit(std::uint32_t v, int kl, unsigned long long cand, unsigned long long &seen) const
            {
                seen = atomicAdd(&res[static_cast<std::uint64_t>(v) * K + kl], cand);
            }
            __device__ bool improved(unsigned long long cand, unsigned long long seen) const
            {
                return seen <= eps && seen + cand > eps;
            }
        };

        // Moves the active residuals of the frontier rows into rank; the
        // snapshot keeps each moved residual's share word per out-edge.
        template <class OffT>
        __global__ void pprk_snapshot(const std::uint32_t *q, std::uint32_t F, std::uint32_t K,
                                      const unsigned long long *qmask, unsigned long long *res,
                                      unsigned long long *rank, unsigned long long *snap, const OffT *off,
                                      float damping, const RoundState *rs = nullptr)
        {
            const std::uint64_t n = static_cast<std::uint64_t>(rs ? rs->F : F) * K;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (std::uint64_t t = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
                 t += stride)
            {
                const std::uint32_t i = static_cast<std::uint32_t>(t / K);
                const std::uint32_t k = static_cast<std::uint32_t>(t % K);
                const std::uint32_t u = q[i];
                unsigned long long sw = 0ULL;
                if ((qmask[i] >> k) & 1ULL)
                {
                    const std::uint64_t idx = static_cast<std::uint64_t>(u) * K + k;
                    const unsigned long long r = res[idx];
                    res[idx] = 0ULL;
                    rank[idx] += r;
                    sw = share_word(r, row_scale_of(damping, static_cast<std::uint64_t>(off[u + 1] - off[u])));
                }
                snap[t] = sw;
            }
        }

        inline unsigned snapshot_blocks(std::uint64_t n)
        {
            const std::uint64_t b = (n + kBlock - 1) / kBlock;
            return static_cast<unsigned>(b == 0 ? 1 : (b < 65536 ? b : 65536));
        }

        int pow2_at_least(std::uint32_t K)
        {
            int kp = 1;
            while (kp < static_cast<int>(K))
                kp <<= 1;
            return kp < 2 ? 2 : kp;
        }

        template <class T>
        __global__ void seed_rows(std::uint32_t nrows, const std::uint32_t *rows, const unsigned long long *masks,
                                  const std::uint32_t *lane_src, std::uint32_t K, unsigned long long *mark, T *val,
                                  T seed_value)
        {
            const std::uint32_t t = threadIdx.x;
            if (t < nrows)
                mark[rows[t]] = masks[t];
            if (t < K)
                val[static_cast<std::uint64_t>(lane_src[t]) * K + t] = seed_value;
        }

        // Column sums of x = rank + res ([V][K] mass words, written into
        // rank) into partial[block][k], exact (no column passes 2^62).
        __global__ __launch_bounds__(kBlock) void pprk_fold(unsigned long long *rank, const unsigned long long *res,
                                                           std::uint64_t V, std::uint32_t K, std::uint32_t KP,
                                                           unsigned long long *partial)
        {
            __shared__ unsigned long long s[kBlock];
            const std::uint32_t rows_per = kBlock / KP;
            const std::uint32_t k = threadIdx.x % KP;
            const std::uint32_t r0 = threadIdx.x / KP;
            unsigned long long acc = 0;
            if (k < K)
            {
                for (std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * rows_per + r0; v < V;
                     v += static_cast<std::uint64_t>(gridDim.x) * rows_per)
                {
                    const std::uint64_t idx = v * K + k;
                    const unsigned long long x = rank[idx] + res[idx];
                    rank[idx] = x;
                    acc += x;
                }
            }
            s[threadIdx.x] = acc;
            __syncthreads();
            if (threadIdx.x < KP && k < K)
            {
                unsigned long long t = 0;
                for (std::uint32_t r = 0; r < rows_per; ++r)
                    t += s[r * KP + threadIdx.x];
                partial[static_cast<std::uint64_t>(blockIdx.x) * 64 + threadIdx.x] = t;
            }
        }

        __global__ void pprk_total(const unsigned long long *partial, unsigned nblocks, std::uint32_t K,
                                   unsigned long long *total)
        {
            const std::uint32_t k = threadIdx.x;
            if (k >= K)
                return;
            unsigned long long t = 0;
            for (unsigned b = 0; b < nblocks; ++b)
                t += partial[static_cast<std::uint64_t>(b) * 64 + k];
            total[k] = t;
        }

        __global__ void pprk_normalize(const unsigned long long *x, std::uint64_t n, std::uint32_t K,
                                       const unsigned long long *total, float *out)
        {
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; i < n; i += stride)
            {
                const double t = static_cast<double>(total[i % K]);
                out[i] = static_cast<float>(t > 0.0 ? static_cast<double>(x[i]) / t : 0.0);
            }
        }

        struct SeedPlan
        {
            std::vector<std::uint32_t> rows;
            std::vector<unsigned long long> masks;
        };

        inline SeedPlan plan_seeds(const std::uint32_t *src, std::uint32_t K)
        {
            SeedPlan p;
            for (std::uint32_t k = 0; k < K; ++k)
            {
                auto it = std::find(p.rows.begin(), p.rows.end(), src[k]);
                if (it == p.rows.end())
                {
                    p.rows.push_back(src[k]);
                    p.masks.push_back(1ULL << k);
                }
                else
                    p.masks[static_cast<std::size_t>(it - p.rows.begin())] |= 1ULL << k;
            }
            return p;
        }

        // Seeds a batch: the seeds buffer (at least 1 KiB) receives rows u32[64] |
        // masks u64[64] | lane sources u32[64]; mark[row] takes each row's lane
        // mask and val[src_k][k] the seed value (T: u32 distance, u64 mass word).
        template <class T>
        inline void upload_seeds(hipStream_t s, DeviceBuffer &seeds, unsigned long long *mark, const SeedPlan &p,
                                 const std::uint32_t *src, std::uint32_t K, T *val, T seed_value)
        {
            std::vector<unsigned char> blob(64 * 4 + 64 * 8 + 64 * 4, 0);
            std::memcpy(blob.data(), p.rows.data(), p.rows.size() * 4);
            std::memcpy(blob.data() + 256, p.masks.data(), p.masks.size() * 8);
            std::memcpy(blob.data() + 768, src, K * 4);
            unsigned char *d = seeds.as<unsigned char>();
            hip_check(hipMemcpyAsync(d, blob.data(), blob.size(), hipMemcpyHostToDevice, s), "upload seeds");
            seed_rows<T><<<1, 64, 0, s>>>(static_cast<std::uint32_t>(p.rows.size()),
                                       reinterpret_cast<const std::uint32_t *>(d),
                                       reinterpret_cast<const unsigned long long *>(d + 256),
                                       reinterpret_cast<const std::uint32_t *>(d + 768), K, mark, val, seed_value);
            hip_check(hipGetLastError(), "launch seed_rows");
            hip_check(hipStreamSynchronize(s), "seed upload");
        }

        // PPR-K result: x = rank + res per (vertex, query), normalised per query
        // column into dst (device); partial holds (1024 + 1) x 64 words.
        inline void pprk_finish(hipStream_t s, unsigned long long *rank, const unsigned long long *res,
                                std::uint64_t V, std::uint32_t K, unsigned long long *partial, float *dst)
        {
            const std::uint32_t KP = static_cast<std::uint32_t>(pow2_at_least(K));
            const std::uint64_t rows_per = kBlock / KP;
            const std::uint64_t nb64 = (V + rows_per - 1) / rows_per;
            const unsigned nb = static_cast<unsigned>(nb64 < 1024 ? (nb64 ? nb64 : 1) : 1024);
            pprk_fold<<<nb, kBlock, 0, s>>>(rank, res, V, K, KP, partial);
            hip_check(hipGetLastError(), "launch pprk_fold");
            unsigned long long *total = partial + static_cast<std::uint64_t>(1024) * 64;
            pprk_total<<<1, 64, 0, s>>>(partial, nb, K, total);
            hip_check(hipGetLastError(), "launch pprk_total");
            const std::uint64_t VK = V * K;
            const std::uint64_t b = (VK + kBlock - 1) / kBlock;
            pprk_normalize<<<static_cast<unsigned>(b < 65536 ? b : 65536), kBlock, 0, s>>>(rank, VK, K, total, dst);
            hip_check(hipGetLastError(), "launch pprk_normalize");
        }

        // Rows with a non-zero mark become the next queue; the mark moves into
        // qmask and is cleared.
        struct MarkSel
        {
            struct Aux
            {
                unsigned long long lanes = 0; // OR of the selected masks
                std::uint32_t pairs = 0;      // sum of their popcounts
            };
            unsigned long long *mark;
            // optional: OR of every selected mask; with it, the number of
            // (row, lane) pairs selected goes to ctr->frontier_edges
            unsigned long long *lanes_out;

            // no 16-row path
            __device__ bool whole16() const { return false; }
            template <class Vals, class A>
            __device__ void select16(std::uint64_t, std::uint64_t, Vals &, std::uint32_t (&)[4], A &) const
            {
            }
            __device__ void publish(const Aux &aux, Counters *ctr) const
            {
                if (lanes_out == nullptr)
                    return;
                std::uint32_t lo = static_cast<std::uint32_t>(aux.lanes);
                std::uint32_t hi = static_cast<std::uint32_t>(aux.lanes >> 32);
                std::uint32_t pairs = aux.pairs;
#pragma unroll
                for (int d = kWave / 2; d > 0; d >>= 1)
                {
                    lo |= __shfl_xor(lo, d, kWave);
                    hi |= __shfl_xor(hi, d, kWave);
                    pairs += __shfl_xor(pairs, d, kWave);
                }
                // one pair of atomics per workgroup (every thread of the
                // workgroup calls publish): per-wavefront atomics on the two
                // shared words serialise
                __shared__ unsigned long long s_all[kWavesPerBlock];
                __shared__ std::uint32_t s_pairs[kWavesPerBlock];
                const int wid = static_cast<int>(threadIdx.x) / kWave;
                if (lane_id() == 0)
                {
                    s_all[wid] = (static_cast<unsigned long long>(hi) << 32) | lo;
                    s_pairs[wid] = pairs;
                }
                __syncthreads();
                if (threadIdx.x == 0)
                {
                    unsigned long long all = 0, sum = 0;
#pragma unroll
                    for (int w = 0; w < kWavesPerBlock; ++w)
                    {
                        all |= s_all[w];
                        sum += s_pairs[w];
                    }
                    if (all != 0ULL)
                    {
                        atomicOr(lanes_out, all);
                        atomicAdd(&ctr->frontier_edges, sum);
                    }
                }
            }
            __device__ std::uint32_t select4(std::uint64_t row0, std::uint32_t n,
                                             unsigned long long (&vals)[4], Aux &aux) const
            {
                unsigned long long m4[4];
                if (n == 4)
                {
                    const ulonglong2 a = *reinterpret_cast<const ulonglong2 *>(mark + row0);
                    const ulonglong2 b = *reinterpret_cast<const ulonglong2 *>(mark + row0 + 2);
                    m4[0] = a.x, m4[1] = a.y, m4[2] = b.x, m4[3] = b.y;
                }
                else
                {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        m4[j] = static_cast<std::uint32_t>(j) < n ? mark[row0 + j] : 0ULL;
                }
                std::uint32_t sel = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (m4[j])
                    {
                        mark[row0 + j] = 0ULL;
                        vals[j] = m4[j];
                        aux.lanes |= m4[j];
                        aux.pairs += static_cast<std::uint32_t>(__popcll(m4[j]));
                        sel |= 1U << j;
                    }
                return sel;
            }
        };

        } // namespace
    } // namespace lanes
} // namespace hyt
