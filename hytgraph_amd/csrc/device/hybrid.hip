// Hybrid transfer engine: SSSP / delta-push PageRank over a graph whose
// edges stay in pinned, device-mapped host memory (HyTGraph's hybrid transfer
// management, executed rather than modelled).
//
// Pageable host graphs (too large to pin: the boxes pin at most 64 GiB in
// total, an RMAT scale-30 CSR is 146 GB) run without zero-copy: the cost
// model chooses between filter and compaction only (its zero-copy cost taken
// as infinite), filter copies are staged through pinned bounce buffers by the
// host threads while the DMA of the previous chunk runs, and a row too long
// for the packed staging streams through the ring in windows.
//
// Resident in HBM: row offsets, vertex state, the frontier, a partition table,
// and a budgeted edge area holding
//   * a streaming ring: a few buffers of k partitions' edges each; a filter
//     task's consecutive partitions move with one copy per array;
//   * the residency cache: units of one partition's columns (or weights);
//     a partition that streamed in is kept when a unit is free or held by a
//     partition that is inactive this round (least recently active first),
//     so a cyclic scan larger than the cache keeps a stable resident set
//     instead of thrashing;
//   * packed staging for compaction.
//
// A round:
//   1. ordered compaction -> frontier queue sorted by vertex, so partition p's
//      frontier rows form one queue range;
//   2. part_stats kernel -> per partition: frontier rows, active edges,
//      zero-copy requests sum(ceil(deg*d1/m)) and alignment extras;
//   3. host: HyTM cost equations + strict selector (hytm_core.hpp) for every
//      active, non-resident partition; task combiner (filter runs <= k,
//      compaction and zero-copy each one task);
//   4. execute: resident partitions expand from their units; filter tasks
//      stream through the ring (copies on a copy stream, ordered by events,
//      overlapping the compute stream) and are promoted into the cache by a
//      device copy; zero-copy expands straight from the mapped host arrays;
//      compaction packs the active rows' edges on host threads into pinned
//      staging (while the GPU works on the rest) and copies them once.
// The expand kernels are the in-HBM ones with a windowed / packed row source.
//
// PageRank rounds whose frontier out-edges reach pull_percent of E pull
// instead, as in HBM: they read the host graph's incoming CSR (built on host
// threads on first use, kept with the host graph), partitioned by destination
// ranges like the out-edges. Every in-partition is read: resident ones from
// their units in one launch per group, the rest streamed through the ring in
// windows and promoted into the cache (resident ids P + q, so in- and
// out-partitions share the units and the least-recently-active rule). The
// cost model has nothing to choose in such a round: every in-edge is needed,
// so a partition is either resident or copied whole.
#include "host_graph.hpp"
#include "hybrid_lanes.hpp"
#include "lanes.hpp"
#include "ops.hpp"
#include "contribution_order.hpp"
#include "pull_seg.hpp"
#include "transfer_plan.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

namespace hyt
{
    namespace
    {
        using Clock = std::chrono::steady_clock;
        double ms_since(Clock::time_point t0)
        {
            return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        }

        // Partition contributions computed beside the activity counts (the
        // priorities of the contribution order): none; SSSP - the smallest
        // frontier distance of the partition (value bits, u32 or float >= 0,
        // whose unsigned order is the value order); PageRank - the mass the
        // partition pushes, sum of share x out-degree (double).
        enum PrioMode : int
        {
            kPrioNone = 0,
            kPrioMinValue = 1,
            kPrioPushedMass = 2
        };

        // Per-partition activity counts from the sorted frontier: rows, active
        // edges, zero-copy requests, alignment extras (u64 each, 4 per
        // partition), and the partition's contribution (prio, per PrioMode,
        // from the rows' queue values qd). Wavefronts whose rows share one
        // partition add once.
        template <class OffT>
        __global__ __launch_bounds__(kBlock) void part_stats(const std::uint32_t *q, std::uint32_t F, const OffT *off,
                                                             const std::uint32_t *part_vbegin, std::uint32_t P,
                                                             std::uint64_t m, std::uint64_t d1, std::uint64_t align,
                                                             unsigned long long *stats, const std::uint32_t *qd,
                                                             int prio_mode, unsigned long long *prio)
        {
            const std::uint32_t i = blockIdx.x * kBlock + threadIdx.x;
            const bool valid = i < F;
            std::uint32_t p = 0;
            unsigned long long rows = 0, edges = 0, reqs = 0, al = 0;
            if (valid)
            {
                const std::uint32_t u = q[i];
                // partition of u: last p with part_vbegin[p] <= u
                std::uint32_t lo = 0, hi = P; // part_vbegin[lo] <= u < part_vbegin[hi]
                while (hi - lo > 1)
                {
                    const std::uint32_t mid = (lo + hi) >> 1;
                    if (part_vbegin[mid] <= u)
                        lo = mid;
                    else
                        hi = mid;
                }
                p = lo;
                const std::uint64_t b = off[u], e = off[u + 1], deg = e - b;
                rows = 1;
                edges = deg;
                if (deg)
                {
                    reqs = (deg * d1 + m - 1) / m;
                    al = ((b * d1) % align) != 0 ? 1 : 0;
                }
            }
            const std::uint32_t p0 = __shfl(p, 0, kWave);
            const bool same = __ballot(valid && p != p0) == 0ULL;
            if (prio_mode == kPrioMinValue)
            {
                const unsigned long long v = valid ? qd[i] : ~0ULL;
                if (same)
                {
                    unsigned long long mn = v;
#pragma unroll
                    for (int d = kWave / 2; d > 0; d >>= 1)
                    {
                        const unsigned long long o = __shfl_xor(mn, d, kWave);
                        mn = o < mn ? o : mn;
                    }
                    if (lane_id() == 0 && mn != ~0ULL)
                        atomicMin(&prio[p0], mn);
                }
                else if (valid)
                    atomicMin(&prio[p], v);
            }
            else if (prio_mode == kPrioPushedMass)
            {
                double v = valid ? static_cast<double>(__uint_as_float(qd[i])) * static_cast<double>(edges) : 0.0;
                double *dp = reinterpret_cast<double *>(prio);
                if (same)
                {
#pragma unroll
                    for (int d = kWave / 2; d > 0; d >>= 1)
                        v += __shfl_xor(v, d, kWave);
                    if (lane_id() == 0 && v != 0.0)
                        atomicAdd(&dp[p0], v);
                }
                else if (valid && v != 0.0)
                    atomicAdd(&dp[p], v);
            }
            if (same)
            {
#pragma unroll
                for (int d = kWave / 2; d > 0; d >>= 1)
                {
                    rows += __shfl_xor(rows, d, kWave);
                    edges += __shfl_xor(edges, d, kWave);
                    reqs += __shfl_xor(reqs, d, kWave);
                    al += __shfl_xor(al, d, kWave);
                }
                if (lane_id() == 0 && rows)
                {
                    atomicAdd(&stats[4 * p0 + 0], rows);
                    atomicAdd(&stats[4 * p0 + 1], edges);
                    atomicAdd(&stats[4 * p0 + 2], reqs);
                    atomicAdd(&stats[4 * p0 + 3], al);
                }
            }
            else if (valid)
            {
                atomicAdd(&stats[4 * p + 0], rows);
                atomicAdd(&stats[4 * p + 1], edges);
                atomicAdd(&stats[4 * p + 2], reqs);
                atomicAdd(&stats[4 * p + 3], al);
            }
        }

        __global__ void hybrid_seed(std::uint32_t *dist, std::uint8_t *touch, std::uint32_t src)
        {
            if (threadIdx.x == 0 && blockIdx.x == 0)
            {
                dist[src] = 0U;
                touch[src] = 1;
            }
        }

        // Folds queue rows back into the marks (rows are unique after an
        // ordered compaction, and nothing else writes mark meanwhile).
        __global__ void mark_from_queue(const std::uint32_t *q, const unsigned long long *qmask, std::uint32_t F,
                                        unsigned long long *mark)
        {
            const std::uint32_t stride = gridDim.x * blockDim.x;
            for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < F; i += stride)
                mark[q[i]] |= qmask[i];
        }

        // Zero-copy bandwidth probe: streams a mapped host range.
        __global__ void zc_probe(const uint4 *src, std::uint64_t n, unsigned *sink)
        {
            unsigned acc = 0;
            for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
                 i += static_cast<std::uint64_t>(gridDim.x) * blockDim.x)
            {
                const uint4 v = src[i];
                acc ^= v.x ^ v.y ^ v.z ^ v.w;
            }
            if (acc == 0x9E3779B9U)
                sink[0] = acc;
        }

        struct Resident
        {
            std::int64_t col = -1, w = -1; // units holding the columns / weights, -1 = none
        };

        constexpr std::size_t kRingBuffers = 4;
        // staged copies from pageable host graphs: bounce buffers of
        // kBounceBytes each, filled by the host threads while the copy engine
        // moves the previous one
        constexpr std::size_t kBounceBuffers = 4;
        constexpr std::uint64_t kBounceBytes = 64ULL << 20;

        struct EventPair
        {
            hipEvent_t a = nullptr, b = nullptr;
        };

        // [V][K] state of K-lane runs (allocated on first use, grown to the
        // largest K), outside the edge budget like the single-query state.
        struct LaneState
        {
            DeviceBuffer val, rank, snap, mark, qmask, stage_mask, heavy, seeds, lanes, partial;
            std::uint64_t heavy_cap = 0;
            std::unique_ptr<PinnedBuffer> host_lanes; // read-back of `lanes`
            std::uint64_t bytes() const
            {
                return val.bytes + rank.bytes + snap.bytes + mark.bytes + qmask.bytes + stage_mask.bytes +
                       heavy.bytes + seeds.bytes + lanes.bytes + partial.bytes;
            }
        };
    } // namespace

    struct HybridEngine
    {
        // host graph (borrowed; pinned + mapped, or pageable)
        HostGraph *host = nullptr;
        bool weights_f32 = false;
        const std::uint32_t *z_col = nullptr, *z_w = nullptr; // device views of host arrays (pinned graphs)
        // pageable graphs: pinned bounce buffers for the staged filter copies
        bool pageable = false;
        std::uint32_t *bounce[kBounceBuffers] = {};
        hipEvent_t bounce_done[kBounceBuffers] = {};
        bool bounce_busy[kBounceBuffers] = {};
        std::size_t bounce_next = 0;
        // device graph: offsets + vertex workspace (no edges)
        DeviceGraph dg;
        std::vector<hyt_partition> parts;
        DeviceBuffer part_vbegin, part_stats_d, block_counts, block_offsets, scan_temp, stage_q, stage_val;
        std::vector<unsigned long long> part_stats_h;
        // contribution order of the current run (PrioMode) and the partition
        // contributions of the current round
        int prio_mode = kPrioNone;
        DeviceBuffer part_prio_d;
        std::vector<unsigned long long> part_prio_h;
        // edge area
        hyt_hybrid_config cfg{};
        hytm::Params params;
        std::uint64_t edge_bytes = 8; // bytes per edge of the host graph (column + weight)
        // per run: whether weights move (SSSP on a weighted graph) and the
        // cost-model parameters with the entry size of what moves
        bool need_w = true;
        std::uint64_t run_edge_bytes = 8;
        hytm::Params run_params;
        // residency cache
        std::uint64_t unit_edges = 0; // 4-byte entries per unit (= partition_edges)
        // Unit layout: a partition read with its weights holds column unit j
        // < pairs and weight unit j + pairs, so one launch addresses every
        // resident partition (weights at a fixed offset from the columns);
        // runs that read no weights place columns in any unit.
        std::uint64_t units = 0, pairs = 0;
        DeviceBuffer unit_mem;
        std::vector<std::int64_t> owner;          // per unit: partition holding it, -1 = free
        std::vector<Resident> res;                // per partition
        std::vector<std::uint64_t> last_active;   // per partition: stamp of the last round it was active
        std::vector<CacheSeg> seg_h;              // this round's resident segments (host copy)
        DeviceBuffer seg_d;
        std::uint64_t stamp = 0;                  // round stamp (monotone across runs)
        // streaming ring
        std::uint64_t ring_edges = 0; // edges per ring buffer
        DeviceBuffer ring_mem;
        hipEvent_t ring_used[kRingBuffers] = {};
        bool ring_valid[kRingBuffers] = {};
        std::size_t ring_next = 0;
        // packed staging (compaction)
        std::uint64_t pack_cap = 0; // edges
        DeviceBuffer pk_rows, pk_index, pk_col, pk_w, pk_val;
        std::uint32_t *h_rows = nullptr, *h_col = nullptr, *h_w = nullptr, *h_val = nullptr;
        std::uint64_t *h_index = nullptr;
        std::uint32_t *h_q = nullptr, *h_qd = nullptr; // frontier ranges read back for packing
        std::uint64_t h_q_cap = 0;
        hipStream_t copy = nullptr;
        std::vector<EventPair> copy_events;
        std::size_t copy_events_used = 0;
        hipEvent_t ev_pack = nullptr;
        int threads = 1;
        // calibration
        double link_gbs = 0, pack_gbs = 0, zc_gbs = 0;
        // degree-bin split of the current run's expansions
        std::uint32_t light_max = kHybridLightMaxDegree, chunk = kHybridHeavyChunk;
        LaneState lk;
        // PageRank pull rounds: partitions of the incoming CSR (destination
        // ranges; resident ids parts.size() + q), its offsets on the device,
        // the segment table and chunk queue of the pull kernels
        std::vector<hyt_partition> in_parts;
        DeviceBuffer in_off_d, pull_seg_d, pull_heavy, pull_ctr;
        std::vector<PullSeg> pull_seg_h;
        std::uint64_t pull_heavy_cap = 0;
        double transpose_ms = 0; // host transpose build, reported once

        ~HybridEngine()
        {
            if (dg.stream)
                (void)hipStreamSynchronize(dg.stream);
            if (copy)
            {
                (void)hipStreamSynchronize(copy);
                (void)hipStreamDestroy(copy);
            }
            for (hipEvent_t e : ring_used)
                if (e)
                    (void)hipEventDestroy(e);
            for (auto &e : copy_events)
            {
                (void)hipEventDestroy(e.a);
                (void)hipEventDestroy(e.b);
            }
            if (ev_pack)
                (void)hipEventDestroy(ev_pack);
            for (std::size_t b = 0; b < kBounceBuffers; ++b)
            {
                if (bounce_done[b])
                    (void)hipEventDestroy(bounce_done[b]);
                if (bounce[b])
                    (void)hipHostFree(bounce[b]);
            }
            for (void *p : {static_cast<void *>(h_rows), static_cast<void *>(h_col), static_cast<void *>(h_w),
                            static_cast<void *>(h_val), static_cast<void *>(h_index), static_cast<void *>(h_q),
                            static_cast<void *>(h_qd)})
                if (p)
                    (void)hipHostFree(p);
        }

        std::uint64_t state_bytes() const
        {
            return dg.device_bytes() + part_vbegin.bytes + part_stats_d.bytes + block_counts.bytes +
                   block_offsets.bytes + scan_temp.bytes + stage_q.bytes + stage_val.bytes + seg_d.bytes +
                   lk.bytes() + in_off_d.bytes + pull_seg_d.bytes + pull_heavy.bytes + pull_ctr.bytes;
        }
        std::uint64_t edge_area_bytes() const
        {
            return unit_mem.bytes + ring_mem.bytes + pk_rows.bytes + pk_index.bytes + pk_col.bytes + pk_w.bytes +
                   pk_val.bytes;
        }

        EventPair &next_copy_events()
        {
            if (copy_events_used == copy_events.size())
            {
                EventPair e;
                hip_check(hipEventCreate(&e.a), "hipEventCreate");
                hip_check(hipEventCreate(&e.b), "hipEventCreate");
                copy_events.push_back(e);
            }
            return copy_events[copy_events_used++];
        }

        double drain_copy_time()
        {
            double total = 0;
            for (std::size_t i = 0; i < copy_events_used; ++i)
            {
                float ms = 0;
                hip_check(hipEventElapsedTime(&ms, copy_events[i].a, copy_events[i].b), "hipEventElapsedTime");
                total += ms;
            }
            copy_events_used = 0;
            return total;
        }

        // Timed host->device copy on the copy stream.
        void copy_h2d(void *dst, const void *src, std::uint64_t bytes)
        {
            if (bytes == 0)
                return;
            EventPair &e = next_copy_events();
            hip_check(hipEventRecord(e.a, copy), "hipEventRecord");
            hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, copy), "copy partition");
            hip_check(hipEventRecord(e.b, copy), "hipEventRecord");
        }

        // Host graph range -> device: one DMA from a pinned graph; from a
        // pageable graph, kBounceBytes chunks copied by the host threads into
        // the next free bounce buffer, each then moved by the copy engine
        // (the copy stream's last event covers the last chunk).
        void copy_host_graph(void *dst, const void *src, std::uint64_t bytes)
        {
            if (!pageable)
            {
                copy_h2d(dst, src, bytes);
                return;
            }
            for (std::uint64_t off = 0; off < bytes; off += kBounceBytes)
            {
                const std::uint64_t n = std::min<std::uint64_t>(kBounceBytes, bytes - off);
                const std::size_t b = bounce_next;
                bounce_next = (bounce_next + 1) % kBounceBuffers;
                if (bounce_busy[b])
                    hip_check(hipEventSynchronize(bounce_done[b]), "wait bounce buffer");
                const char *from = static_cast<const char *>(src) + off;
                char *to = reinterpret_cast<char *>(bounce[b]);
                const int T = static_cast<int>(std::max<std::uint64_t>(
                    1, std::min<std::uint64_t>(static_cast<std::uint64_t>(threads), n >> 22)));
                if (T == 1)
                    std::memcpy(to, from, n);
                else
                {
                    std::vector<std::thread> pool;
                    for (int t = 0; t < T; ++t)
                        pool.emplace_back([=] {
                            const std::uint64_t lo = (n * t / T) & ~63ULL;
                            const std::uint64_t hi = t + 1 == T ? n : ((n * (t + 1) / T) & ~63ULL);
                            std::memcpy(to + lo, from + lo, hi - lo);
                        });
                    for (auto &th : pool)
                        th.join();
                }
                copy_h2d(static_cast<char *>(dst) + off, bounce[b], n);
                hip_check(hipEventRecord(bounce_done[b], copy), "hipEventRecord");
                bounce_busy[b] = true;
            }
        }

        std::uint32_t *unit_ptr(std::int64_t u) const
        {
            return unit_mem.as<std::uint32_t>() + static_cast<std::uint64_t>(u) * unit_edges;
        }
        std::uint32_t *ring_col(std::size_t r) const
        {
            return ring_mem.as<std::uint32_t>() + r * ring_edges * (edge_bytes / 4);
        }
        std::uint32_t *ring_w(std::size_t r) const { return need_w ? ring_col(r) + ring_edges : nullptr; }

        // Next ring buffer; the copy stream waits until its last kernel is done.
        std::size_t ring_take()
        {
            const std::size_t r = ring_next;
            ring_next = (ring_next + 1) % kRingBuffers;
            if (ring_valid[r])
                hip_check(hipStreamWaitEvent(copy, ring_used[r], 0), "wait ring buffer");
            return r;
        }
        void ring_release(std::size_t r)
        {
            hip_check(hipEventRecord(ring_used[r], dg.stream), "hipEventRecord");
            ring_valid[r] = true;
        }

        // Weights move only for runs that read them.
        // prio: the contribution order of a single-query run (kPrioNone: the
        // round's filter runs go in partition order)
        void begin_run(bool weights, int prio = kPrioNone)
        {
            prio_mode = prio;
            need_w = weights && host->weighted;
            run_edge_bytes = need_w ? 8 : 4;
            run_params = params;
            if (!need_w)
                run_params.entry_bytes = 4;
        }

        void set_split(const hyt_run_options &o, std::uint32_t light_default = kHybridLightMaxDegree,
                       std::uint32_t chunk_default = kHybridHeavyChunk)
        {
            light_max = o.light_max_degree ? o.light_max_degree : light_default;
            chunk = o.heavy_chunk ? o.heavy_chunk : chunk_default;
            require(light_max <= (1U << 20), "light_max_degree must be <= 2^20");
            require(chunk >= 64 && chunk <= (1U << 24), "heavy_chunk must be in [64, 2^24]");
            ensure_workspace(dg, light_max, chunk);
            if (stage_val.ptr == nullptr)
                stage_val.alloc(static_cast<std::uint64_t>(compact_blocks(dg.V)) * kCompactRowsPerBlock * 4,
                                "hipMalloc(stage values)");
        }

        // HBM a lane session leaves allocated besides its lane state: the edge
        // area, offsets, partition tables, compaction staging and the queue.
        std::uint64_t serving_bytes() const
        {
            const std::uint64_t Vp = dg.V + kPad;
            return edge_area_bytes() + dg.off.bytes + part_vbegin.bytes + part_stats_d.bytes + block_counts.bytes +
                   block_offsets.bytes + scan_temp.bytes + stage_q.bytes + seg_d.bytes + Vp * 4 + sizeof(Counters) +
                   in_off_d.bytes + pull_seg_d.bytes + pull_heavy.bytes + pull_ctr.bytes;
        }

        // A lane session uses only the queue and the counters of the
        // single-query workspace; the rest is released (a later single-query
        // call allocates it again).
        void release_single_query_state()
        {
            Workspace &ws = dg.ws;
            for (DeviceBuffer *b : {&ws.dist, &ws.res, &ws.rank, &ws.prev, &ws.touch, &ws.qd, &ws.partial, &ws.active,
                                    &ws.share, &ws.heavy, &ws.hub_bits, &ws.hub_list, &stage_val})
                b->release();
            ws.heavy_cap = 0;
            ws.hub_rows = 0;
            ws.ready = false; // the next ensure_workspace re-creates every array
        }

        bool is_cached(std::uint64_t p) const { return res[p].col >= 0 && (!need_w || res[p].w >= 0); }

        void drop(std::uint64_t p)
        {
            Resident &r = res[p];
            if (r.col >= 0)
                owner[static_cast<std::size_t>(r.col)] = -1;
            if (r.w >= 0)
                owner[static_cast<std::size_t>(r.w)] = -1;
            r = Resident{};
        }

        // Finds units for partition p; partitions active this round (this
        // round's stamp) keep theirs. False when nothing can be freed.
        bool admit(std::uint64_t p)
        {
            if (res[p].col >= 0)
                drop(p); // cached without the weights this run needs
            Resident &r = res[p];
            if (need_w)
            {
                // the pair whose holders were active least recently (free first)
                std::int64_t best = -1;
                std::uint64_t best_key = ~0ULL;
                for (std::uint64_t j = 0; j < pairs && best_key != 0; ++j)
                {
                    std::uint64_t key = 0;
                    bool ok = true;
                    for (const std::int64_t o : {owner[j], owner[j + pairs]})
                    {
                        if (o < 0)
                            continue;
                        if (last_active[static_cast<std::size_t>(o)] >= stamp)
                        {
                            ok = false;
                            break;
                        }
                        key = std::max<std::uint64_t>(key, last_active[static_cast<std::size_t>(o)] + 1);
                    }
                    if (ok && key < best_key)
                    {
                        best_key = key;
                        best = static_cast<std::int64_t>(j);
                    }
                }
                if (best < 0)
                    return false;
                const std::size_t j = static_cast<std::size_t>(best);
                for (const std::size_t k : {j, j + pairs})
                    if (owner[k] >= 0)
                        drop(static_cast<std::uint64_t>(owner[k]));
                r.col = best;
                r.w = best + static_cast<std::int64_t>(pairs);
                owner[j] = owner[j + pairs] = static_cast<std::int64_t>(p);
                return true;
            }
            // columns only: a free unit, else a resident's weight unit (it keeps
            // its columns), else the column unit of the least recently active
            std::int64_t best = -1;
            std::uint64_t best_key = ~0ULL;
            for (std::uint64_t k = 0; k < units && best_key != 0; ++k)
            {
                const std::int64_t o = owner[k];
                std::uint64_t key;
                if (o < 0)
                    key = 0;
                else if (res[static_cast<std::size_t>(o)].w == static_cast<std::int64_t>(k))
                    key = 1;
                else if (last_active[static_cast<std::size_t>(o)] < stamp)
                    key = 2 + last_active[static_cast<std::size_t>(o)];
                else
                    continue;
                if (key < best_key)
                {
                    best_key = key;
                    best = static_cast<std::int64_t>(k);
                }
            }
            if (best < 0)
                return false;
            const std::size_t k = static_cast<std::size_t>(best);
            const std::int64_t o = owner[k];
            if (o >= 0)
            {
                Resident &ro = res[static_cast<std::size_t>(o)];
                if (ro.w == best)
                {
                    ro.w = -1;
                    owner[k] = -1;
                }
                else
                    drop(static_cast<std::uint64_t>(o));
            }
            r.col = best;
            owner[k] = static_cast<std::int64_t>(p);
            return true;
        }
    };
} // namespace hyt

struct hyt_hybrid
{
    hyt::HybridEngine e;
};

namespace hyt
{
    namespace
    {
        void calibrate(HybridEngine &E)
        {
            hipStream_t s = E.dg.stream;
            // H2D peak: copy up to 256 MiB of the host column array (a
            // pageable graph: through the staged path its filter copies take,
            // timed on the host)
            const std::uint64_t bytes = std::min<std::uint64_t>(256ULL << 20, E.host->E * 4);
            DeviceBuffer &dst = E.unit_mem.bytes >= bytes ? E.unit_mem : E.ring_mem;
            if (E.pageable && bytes >= (4ULL << 20) && dst.bytes >= bytes)
            {
                double sec = 0;
                for (int rep = 0; rep < 2; ++rep)
                {
                    const auto t0 = Clock::now();
                    E.copy_host_graph(dst.ptr, E.host->col, bytes);
                    hip_check(hipStreamSynchronize(E.copy), "sync");
                    sec = std::chrono::duration<double>(Clock::now() - t0).count();
                }
                (void)E.drain_copy_time();
                E.link_gbs = bytes / sec / 1e9;
            }
            else if (bytes >= (4ULL << 20) && dst.bytes >= bytes)
            {
                hipEvent_t a, b;
                hip_check(hipEventCreate(&a), "hipEventCreate");
                hip_check(hipEventCreate(&b), "hipEventCreate");
                for (int rep = 0; rep < 2; ++rep)
                {
                    hip_check(hipEventRecord(a, s), "hipEventRecord");
                    hip_check(hipMemcpyAsync(dst.ptr, E.host->col, bytes, hipMemcpyHostToDevice, s), "copy");
                    hip_check(hipEventRecord(b, s), "hipEventRecord");
                    hip_check(hipEventSynchronize(b), "sync");
                }
                float ms = 0;
                hip_check(hipEventElapsedTime(&ms, a, b), "elapsed");
                E.link_gbs = bytes / (ms * 1e-3) / 1e9;
                // zero-copy streaming read of the same range
                DeviceBuffer sink(64);
                for (int rep = 0; rep < 2; ++rep)
                {
                    hip_check(hipEventRecord(a, s), "hipEventRecord");
                    zc_probe<<<1024, kBlock, 0, s>>>(reinterpret_cast<const uint4 *>(E.z_col), bytes / 16,
                                                     sink.as<unsigned>());
                    hip_check(hipGetLastError(), "launch zc_probe");
                    hip_check(hipEventRecord(b, s), "hipEventRecord");
                    hip_check(hipEventSynchronize(b), "sync");
                }
                hip_check(hipEventElapsedTime(&ms, a, b), "elapsed");
                E.zc_gbs = bytes / (ms * 1e-3) / 1e9;
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
            }
            // host packing: gather rows of ~16 edges into pinned staging
            if (E.pack_cap >= (1ULL << 20) && E.host->E > 0)
            {
                const std::uint64_t V = E.host->V;
                const std::uint64_t want = std::min<std::uint64_t>(E.pack_cap / 2, 8ULL << 20);
                const auto t0 = Clock::now();
                std::uint64_t moved = 0, at = 0;
                for (std::uint64_t v = 0; v < V && at < want; v += 7)
                {
                    const std::uint64_t b = E.host->off[v], n = E.host->off[v + 1] - b;
                    if (at + n > want)
                        break;
                    std::memcpy(E.h_col + at, E.host->col + b, n * 4);
                    at += n;
                }
                moved = at * 4;
                const double sec = std::chrono::duration<double>(Clock::now() - t0).count();
                if (sec > 0 && moved > 0)
                    E.pack_gbs = moved / sec / 1e9 * E.threads; // threads pack disjoint rows in parallel
            }
            // Recalibrated model: a "TLP" of m*MR bytes takes RTT at the measured
            // link rate; host compaction throughput as measured.
            if (E.link_gbs > 0)
                E.params.rtt = static_cast<double>(E.params.request_bytes * E.params.requests_per_tlp) /
                               (E.link_gbs * 1e9);
            if (E.pack_gbs > 0)
                E.params.cpu_bytes_per_second = E.pack_gbs * 1e9;
        }

        template <class OffT>
        void upload_offsets(HybridEngine &E)
        {
            const std::uint64_t V = E.host->V;
            E.dg.off.alloc((V + 1) * sizeof(OffT), "hipMalloc(row offsets)");
            if (sizeof(OffT) == 8)
                hip_check(hipMemcpy(E.dg.off.ptr, E.host->off, (V + 1) * 8, hipMemcpyHostToDevice), "copy offsets");
            else
            {
                std::vector<std::uint32_t> o(V + 1);
                for (std::uint64_t v = 0; v <= V; ++v)
                    o[v] = static_cast<std::uint32_t>(E.host->off[v]);
                hip_check(hipMemcpy(E.dg.off.ptr, o.data(), (V + 1) * 4, hipMemcpyHostToDevice), "copy offsets");
            }
        }

        void create(HybridEngine &E, HostGraph *host, bool f32, const hyt_hybrid_config &cfg)
        {
            require(host != nullptr, "null host graph");
            require(host->memory == HYT_HOST_PINNED || host->memory == HYT_HOST_PAGEABLE, "unknown host memory kind");
            E.pageable = host->memory == HYT_HOST_PAGEABLE;
            require(!E.pageable || cfg.force_engine != HYT_ENGINE_ZERO_COPY,
                    "zero-copy needs a pinned host graph (a pageable graph runs filter and compaction only)");
            require(host->V > 0 && host->V <= 0xFFFFFFFFULL, "vertex count must be in [1, 2^32-1]");
            require(cfg.partition_edges > 0, "partition_edges must be > 0");
            E.host = host;
            E.weights_f32 = f32;
            E.cfg = cfg;
            E.params = params_from(cfg.cost);
            hytm::validate(E.params);
            E.edge_bytes = host->weighted ? 8 : 4;
            E.threads = cfg.host_threads ? static_cast<int>(cfg.host_threads)
                                         : static_cast<int>(std::min(16U, std::max(1U, std::thread::hardware_concurrency())));
            DeviceGraph &g = E.dg;
            g.device = require_device();
            g.V = host->V;
            g.E = host->E;
            for (std::uint64_t v = 0; v < host->V; ++v)
                g.max_degree = std::max<std::uint64_t>(g.max_degree, host->off[v + 1] - host->off[v]);
            require((cfg.flags & ~static_cast<std::uint32_t>(HYT_GRAPH_OFF64)) == 0, "unknown hybrid flags");
            g.off64 = host->E >= 0xFFFFFFFFULL || (cfg.flags & HYT_GRAPH_OFF64) != 0;
            g.weighted = host->weighted;
            g.weights_f32 = f32;
            hip_check(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipStreamCreateWithFlags(&E.copy, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreate(&g.ev0), "hipEventCreate");
            hip_check(hipEventCreate(&g.ev1), "hipEventCreate");
            hip_check(hipEventCreateWithFlags(&E.ev_pack, hipEventDisableTiming), "hipEventCreate");
            g.host_ctr = std::make_unique<PinnedBuffer>(sizeof(Counters) + sizeof(RoundCtl));
            if (E.pageable)
            {
                for (std::size_t b = 0; b < kBounceBuffers; ++b)
                {
                    hip_check(hipHostMalloc(reinterpret_cast<void **>(&E.bounce[b]), kBounceBytes, hipHostMallocDefault),
                              "hipHostMalloc(bounce buffer)");
                    hip_check(hipEventCreateWithFlags(&E.bounce_done[b], hipEventDisableTiming), "hipEventCreate");
                }
            }
            else
            {
                void *dp = nullptr;
                hip_check(hipHostGetDevicePointer(&dp, host->col, 0), "hipHostGetDevicePointer(columns)");
                E.z_col = static_cast<const std::uint32_t *>(dp);
                if (host->weighted)
                {
                    hip_check(hipHostGetDevicePointer(&dp, host->w, 0), "hipHostGetDevicePointer(weights)");
                    E.z_w = static_cast<const std::uint32_t *>(dp);
                }
            }
            if (g.off64)
                upload_offsets<std::uint64_t>(E);
            else
                upload_offsets<std::uint32_t>(E);
            ensure_workspace(g, kHybridLightMaxDegree, kHybridHeavyChunk);

            // partitions (target bytes = 4 B per edge, as the original counts)
            E.parts = partition_graph(host->off, host->V, cfg.partition_edges * 4);
            const std::uint64_t P = E.parts.size();
            std::vector<std::uint32_t> vb(P + 1);
            std::uint64_t max_part = 0;
            for (std::uint64_t p = 0; p < P; ++p)
            {
                vb[p] = E.parts[p].vertex_begin;
                max_part = std::max<std::uint64_t>(max_part, E.parts[p].edge_end - E.parts[p].edge_begin);
            }
            vb[P] = static_cast<std::uint32_t>(host->V);
            E.part_vbegin.alloc((P + 1) * 4, "hipMalloc(partition table)");
            hip_check(hipMemcpy(E.part_vbegin.ptr, vb.data(), (P + 1) * 4, hipMemcpyHostToDevice), "copy partitions");
            E.part_stats_d.alloc(P * 4 * 8, "hipMalloc(partition stats)");
            E.part_stats_h.assign(P * 4, 0);
            E.part_prio_d.alloc(P * 8, "hipMalloc(partition contributions)");
            E.part_prio_h.assign(P, 0);
            E.res.assign(P, Resident{});
            E.last_active.assign(P, 0);

            // ordered compaction buffers
            const unsigned nb = compact_blocks(host->V);
            E.block_counts.alloc((nb + 1) * 4, "hipMalloc(block counts)");
            E.block_offsets.alloc((nb + 1) * 4, "hipMalloc(block offsets)");
            hip_check(hipMemset(E.block_counts.ptr, 0, (nb + 1) * 4), "clear block counts");
            std::size_t temp = 0;
            hip_check(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, E.block_counts.as<std::uint32_t>(),
                                                       E.block_offsets.as<std::uint32_t>(), nb + 1, g.stream),
                      "scan size");
            E.scan_temp.alloc(temp + 16, "hipMalloc(scan temp)");
            E.stage_q.alloc(static_cast<std::uint64_t>(nb) * kCompactRowsPerBlock * 4, "hipMalloc(stage queue)");
            E.stage_val.alloc(static_cast<std::uint64_t>(nb) * kCompactRowsPerBlock * 4, "hipMalloc(stage values)");

            // edge area: packed staging, the streaming ring (at most half of
            // what is left), the rest in residency units
            const std::uint64_t budget = cfg.device_budget_bytes;
            E.pack_cap = cfg.pack_edges ? cfg.pack_edges : std::max<std::uint64_t>(4 * cfg.partition_edges, 1ULL << 22);
            // exactly what the packed-staging allocations below take
            const std::uint64_t pack_bytes = (E.pack_cap * 4 + 16) * (host->weighted ? 4 : 3) +
                                             ((E.pack_cap + 1) * 8 + 16);
            require(budget > pack_bytes, "device budget too small for the packed staging area");
            E.unit_edges = std::max<std::uint64_t>(cfg.partition_edges, 1);
            const std::uint64_t avail = budget - pack_bytes;
            const std::uint64_t ring_unit = kRingBuffers * E.unit_edges * E.edge_bytes; // ring of one partition each
            require(avail >= ring_unit, "device budget too small for the streaming ring");
            // ring windows of up to max(k, 8) partitions, within half of what is left
            const std::uint64_t k = std::max<std::uint64_t>(cfg.filter_combine_k ? cfg.filter_combine_k : 4, 8);
            const std::uint64_t ring_k =
                std::max<std::uint64_t>(1, std::min<std::uint64_t>(k, avail / 2 / ring_unit));
            E.ring_edges = ring_k * E.unit_edges;
            E.ring_mem.alloc(ring_k * ring_unit, "hipMalloc(streaming ring)");
            for (hipEvent_t &ev : E.ring_used)
                hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
            E.units = (avail - ring_k * ring_unit) / (E.unit_edges * 4);
            E.pairs = host->weighted ? E.units / 2 : 0;
            if (E.units > 0)
                E.unit_mem.alloc(E.units * E.unit_edges * 4, "hipMalloc(residency units)");
            E.owner.assign(E.units, -1);
            E.seg_h.resize(P);
            E.seg_d.alloc(P * sizeof(CacheSeg), "hipMalloc(cache segments)");
            E.pk_rows.alloc(E.pack_cap * 4 + 16, "hipMalloc(packed rows)");
            E.pk_index.alloc((E.pack_cap + 1) * 8 + 16, "hipMalloc(packed index)");
            E.pk_col.alloc(E.pack_cap * 4 + 16, "hipMalloc(packed columns)");
            if (host->weighted)
                E.pk_w.alloc(E.pack_cap * 4 + 16, "hipMalloc(packed weights)");
            E.pk_val.alloc(E.pack_cap * 4 + 16, "hipMalloc(packed values)");
            auto pin = [](void **p, std::uint64_t bytes) {
                hip_check(hipHostMalloc(p, bytes + 16, hipHostMallocDefault), "hipHostMalloc(staging)");
            };
            pin(reinterpret_cast<void **>(&E.h_rows), E.pack_cap * 4);
            pin(reinterpret_cast<void **>(&E.h_index), (E.pack_cap + 1) * 8);
            pin(reinterpret_cast<void **>(&E.h_col), E.pack_cap * 4);
            if (host->weighted)
                pin(reinterpret_cast<void **>(&E.h_w), E.pack_cap * 4);
            pin(reinterpret_cast<void **>(&E.h_val), E.pack_cap * 4);
            if (cfg.calibrate)
                calibrate(E);
            hip_check(hipStreamSynchronize(g.stream), "hybrid create");
            (void)max_part;
        }

        // ------------------------------------------------------------ rounds
        template <class OffT, class Sel>
        std::uint32_t ordered_compaction(HybridEngine &E, const Sel &sel, Counters &out)
        {
            DeviceGraph &g = E.dg;
            Workspace &ws = g.ws;
            hipStream_t s = g.stream;
            Counters *d_ctr = ws.ctr.as<Counters>();
            const unsigned nb = compact_blocks(g.V);
            compact_rows<std::uint32_t, OffT, Sel><<<nb, kBlock, 0, s>>>(
                g.V, sel, E.stage_q.as<std::uint32_t>(), E.stage_val.as<std::uint32_t>(), d_ctr, g.off.as<OffT>(),
                static_cast<std::uint16_t *>(nullptr), E.block_counts.as<std::uint32_t>());
            hip_check(hipGetLastError(), "launch ordered compaction");
            std::size_t temp = E.scan_temp.bytes;
            hip_check(hipcub::DeviceScan::ExclusiveSum(E.scan_temp.ptr, temp, E.block_counts.as<std::uint32_t>(),
                                                       E.block_offsets.as<std::uint32_t>(), nb + 1, s),
                      "scan block counts");
            gather_ordered<std::uint32_t><<<nb, kBlock, 0, s>>>(
                E.stage_q.as<std::uint32_t>(), E.stage_val.as<std::uint32_t>(), E.block_counts.as<std::uint32_t>(),
                E.block_offsets.as<std::uint32_t>(), nb, ws.q.as<std::uint32_t>(), ws.qd.as<std::uint32_t>(), d_ctr);
            hip_check(hipGetLastError(), "launch gather");
            Counters *h = g.host_ctr->as<Counters>();
            hip_check(hipMemcpyAsync(h, d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, s), "read counters");
            hip_check(hipStreamSynchronize(s), "hybrid round");
            out = *h;
            return h->next_count;
        }

        struct Plan
        {
            std::vector<std::uint64_t> qlo, qn;  // frontier range per partition
            std::vector<std::uint64_t> active;   // partitions with frontier edges, ascending
            std::vector<Task> tasks;             // over non-resident active partitions
            std::vector<std::uint64_t> cached;   // resident active partitions
        };

        template <class OffT>
        void plan_round(HybridEngine &E, std::uint32_t F, Plan &plan, hyt_hybrid_stats &st)
        {
            const auto t0 = Clock::now();
            ++E.stamp;
            DeviceGraph &g = E.dg;
            hipStream_t s = g.stream;
            const std::uint64_t P = E.parts.size();
            hip_check(hipMemsetAsync(E.part_stats_d.ptr, 0, P * 32, s), "clear partition stats");
            if (E.prio_mode != kPrioNone)
                hip_check(hipMemsetAsync(E.part_prio_d.ptr, E.prio_mode == kPrioMinValue ? 0xFF : 0, P * 8, s),
                          "clear partition contributions");
            part_stats<OffT><<<blocks_for(F), kBlock, 0, s>>>(
                g.ws.q.as<std::uint32_t>(), F, g.off.as<OffT>(), E.part_vbegin.as<std::uint32_t>(),
                static_cast<std::uint32_t>(P), E.run_params.request_bytes, E.run_params.entry_bytes,
                E.run_params.alignment_bytes, E.part_stats_d.as<unsigned long long>(), g.ws.qd.as<std::uint32_t>(),
                E.prio_mode, E.part_prio_d.as<unsigned long long>());
            hip_check(hipGetLastError(), "launch part_stats");
            hip_check(hipMemcpyAsync(E.part_stats_h.data(), E.part_stats_d.ptr, P * 32, hipMemcpyDeviceToHost, s),
                      "read partition stats");
            if (E.prio_mode != kPrioNone)
                hip_check(hipMemcpyAsync(E.part_prio_h.data(), E.part_prio_d.ptr, P * 8, hipMemcpyDeviceToHost, s),
                          "read partition contributions");
            hip_check(hipStreamSynchronize(s), "partition stats");
            plan.qlo.assign(P, 0);
            plan.qn.assign(P, 0);
            plan.active.clear();
            plan.cached.clear();
            std::vector<std::int32_t> engines;
            std::vector<std::uint64_t> members;
            std::uint64_t at = 0;
            for (std::uint64_t p = 0; p < P; ++p)
            {
                const unsigned long long *c = &E.part_stats_h[4 * p];
                plan.qlo[p] = at;
                plan.qn[p] = c[0];
                at += c[0];
                if (c[1] == 0)
                    continue; // no frontier edges: nothing to move (original filter_engine active test)
                plan.active.push_back(p);
                E.last_active[p] = E.stamp;
                if (E.is_cached(p))
                {
                    plan.cached.push_back(p);
                    continue;
                }
                int engine;
                if (E.cfg.force_engine >= 0)
                    engine = E.cfg.force_engine;
                else
                {
                    hytm::Metrics m;
                    m.active_vertices = c[0];
                    m.active_edges = c[1];
                    m.total_edges = E.parts[p].edge_end - E.parts[p].edge_begin;
                    m.zc_requests = c[2];
                    m.zc_alignment = c[3];
                    hytm::finish_metrics(m, E.run_params);
                    const hytm::Costs cst = hytm::evaluate(m, E.run_params);
                    engine = cst.engine;
                    // a pageable graph has no zero-copy: the strict selector
                    // with its cost infinite (compaction when T_c < alpha T_f)
                    if (E.pageable && engine == HYT_ENGINE_ZERO_COPY)
                        engine = cst.compaction < E.run_params.alpha * cst.filter ? HYT_ENGINE_COMPACTION
                                                                                  : HYT_ENGINE_FILTER;
                }
                engines.push_back(engine);
                members.push_back(p);
            }
            require(at == F, "partition statistics do not cover the frontier");
            combine_tasks(engines.data(), engines.size(), E.cfg.filter_combine_k ? E.cfg.filter_combine_k : 4,
                          plan.tasks);
            for (auto &t : plan.tasks)
                for (auto &m : t.members)
                    m = members[m]; // task members -> partition ids
            st.tasks += plan.tasks.size();
            st.plan_ms += ms_since(t0);
        }

        // One expand (light + heavy) over a row source and window, on the compute stream.
        template <class RowSrc, class Op>
        void expand(HybridEngine &E, const RowSrc &rs, const Op &op, std::uint32_t F)
        {
            if (F == 0)
                return;
            DeviceGraph &g = E.dg;
            hipStream_t s = g.stream;
            Counters *d_ctr = g.ws.ctr.as<Counters>();
            hip_check(hipMemsetAsync(&d_ctr->heavy_count, 0, 4, s), "reset heavy count");
            ExpandArgs a;
            a.q = nullptr;
            a.F = F;
            a.ctr = d_ctr;
            a.heavy = g.ws.heavy.as<HeavyChunk>();
            a.heavy_cap = g.ws.heavy_cap;
            a.light_max = E.light_max;
            a.chunk = E.chunk;
            expand_light<RowSrc, Op><<<blocks_for(F), kBlock, 0, s>>>(rs, op, a);
            hip_check(hipGetLastError(), "launch hybrid expand_light");
            expand_heavy<Op><<<kHeavyGrid, kBlock, 0, s>>>(op, a);
            hip_check(hipGetLastError(), "launch hybrid expand_heavy");
        }

        // Expander of single-query rounds. A round hands row sets to an
        // expander X:
        //   X.window(qa, F, lo, hi, col, w): queue rows [qa, qa + F), edges
        //       clipped to [lo, hi), col/w address edge lo;
        //   X.list(rows, vals, F, lo, hi, col, w): F rows of a device list with
        //       their row values;
        //   X.packed(rows, index, vals, F, col, w): rows packed on the host;
        //   X.cached(seg, nseg, F, col, w): the rows of the resident
        //       partitions (CacheSeg table; col/w address the unit area).
        // X::kIndexValues: the row values of listed and packed rows are queue
        // indices (K-lane rounds) instead of the rows' frontier values.
        template <class OffT, class Make>
        struct SingleExpander
        {
            static constexpr bool kIndexValues = false;
            HybridEngine &E;
            Make make; // Make(col_base, w_base, value_base) -> Op
            const OffT *off;
            const std::uint32_t *q, *qd;

            void window(std::uint64_t qa, std::uint32_t F, std::uint64_t lo, std::uint64_t hi,
                        const std::uint32_t *col, const std::uint32_t *w) const
            {
                expand(E, GraphRows<OffT>{off, q + qa, lo, hi}, make(col, w, qd + qa), F);
            }
            void list(const std::uint32_t *rows, const std::uint32_t *vals, std::uint32_t F, std::uint64_t lo,
                      std::uint64_t hi, const std::uint32_t *col, const std::uint32_t *w) const
            {
                expand(E, GraphRows<OffT>{off, rows, lo, hi}, make(col, w, vals), F);
            }
            void packed(const std::uint32_t *rows, const std::uint64_t *index, const std::uint32_t *vals,
                        std::uint32_t F, const std::uint32_t *col, const std::uint32_t *w) const
            {
                expand(E, PackedRows{rows, index}, make(col, w, vals), F);
            }
            void cached(const CacheSeg *seg, std::uint32_t nseg, std::uint32_t F, const std::uint32_t *col,
                        const std::uint32_t *w) const
            {
                expand(E, CachedRows<OffT>{off, q, seg, nseg}, make(col, w, qd), F);
            }
        };

        template <class OffT, class Make>
        SingleExpander<OffT, Make> single_expander(HybridEngine &E, Make make)
        {
            return SingleExpander<OffT, Make>{E, make, E.dg.off.template as<OffT>(),
                                              E.dg.ws.q.template as<std::uint32_t>(),
                                              E.dg.ws.qd.template as<std::uint32_t>()};
        }

        template <class OffT, class X>
        void execute_round(HybridEngine &E, const Plan &plan, const X &x, hyt_hybrid_stats &st)
        {
            DeviceGraph &g = E.dg;
            hipStream_t s = g.stream;
            const std::uint32_t *q = g.ws.q.as<std::uint32_t>();
            const std::uint32_t *qd = g.ws.qd.as<std::uint32_t>();
            const bool weighted = E.need_w;

            // maximal runs [p0, p1] of consecutive partition ids in a member list
            auto for_runs = [](const std::vector<std::uint64_t> &m, auto fn) {
                std::size_t i = 0;
                while (i < m.size())
                {
                    std::size_t j = i + 1;
                    while (j < m.size() && m[j] == m[j - 1] + 1)
                        ++j;
                    fn(m[i], m[j - 1]);
                    i = j;
                }
            };
            // frontier rows of partitions [p0, p1]: one range of the sorted queue
            auto rows_of = [&](std::uint64_t p0, std::uint64_t p1) {
                return static_cast<std::uint32_t>(plan.qlo[p1] + plan.qn[p1] - plan.qlo[p0]);
            };

            // 1. compaction tasks read their frontier rows back while the stream is idle
            std::uint64_t crows = 0;
            for (const Task &t : plan.tasks)
                if (t.engine == HYT_ENGINE_COMPACTION)
                    for (std::uint64_t p : t.members)
                        crows += plan.qn[p];
            if (crows)
            {
                if (E.h_q_cap < crows)
                {
                    if (E.h_q)
                        (void)hipHostFree(E.h_q);
                    if (E.h_qd)
                        (void)hipHostFree(E.h_qd);
                    E.h_q = E.h_qd = nullptr;
                    E.h_q_cap = crows + 1024;
                    hip_check(hipHostMalloc(reinterpret_cast<void **>(&E.h_q), E.h_q_cap * 4), "hipHostMalloc");
                    hip_check(hipHostMalloc(reinterpret_cast<void **>(&E.h_qd), E.h_q_cap * 4), "hipHostMalloc");
                }
                std::uint64_t at = 0;
                for (const Task &t : plan.tasks)
                    if (t.engine == HYT_ENGINE_COMPACTION)
                        for_runs(t.members, [&](std::uint64_t p0, std::uint64_t p1) {
                            const std::uint64_t n = rows_of(p0, p1);
                            hip_check(hipMemcpyAsync(E.h_q + at, q + plan.qlo[p0], n * 4, hipMemcpyDeviceToHost, s),
                                      "read frontier");
                            if constexpr (X::kIndexValues)
                                for (std::uint64_t j = 0; j < n; ++j)
                                    E.h_qd[at + j] = static_cast<std::uint32_t>(plan.qlo[p0] + j);
                            else
                                hip_check(hipMemcpyAsync(E.h_qd + at, qd + plan.qlo[p0], n * 4,
                                                         hipMemcpyDeviceToHost, s),
                                          "read frontier values");
                            at += n;
                        });
                hip_check(hipStreamSynchronize(s), "frontier readback");
            }

            // 2. resident partitions: one launch over all of them (segment table)
            if (!plan.cached.empty())
            {
                std::uint64_t rows = 0;
                for (std::size_t k = 0; k < plan.cached.size(); ++k)
                {
                    const std::uint64_t p = plan.cached[k];
                    CacheSeg &sg = E.seg_h[k];
                    sg.qlo = static_cast<std::uint32_t>(plan.qlo[p]);
                    sg.rows_before = static_cast<std::uint32_t>(rows);
                    sg.edge_begin = E.parts[p].edge_begin;
                    sg.unit_pos = static_cast<std::uint64_t>(E.res[p].col) * E.unit_edges;
                    rows += plan.qn[p];
                }
                // seg_h is rewritten only after the round's final synchronisation
                hip_check(hipMemcpyAsync(E.seg_d.ptr, E.seg_h.data(), plan.cached.size() * sizeof(CacheSeg),
                                         hipMemcpyHostToDevice, s),
                          "upload cache segments");
                x.cached(E.seg_d.as<CacheSeg>(), static_cast<std::uint32_t>(plan.cached.size()),
                         static_cast<std::uint32_t>(rows), E.unit_mem.as<std::uint32_t>(),
                         weighted ? E.unit_ptr(static_cast<std::int64_t>(E.pairs)) : nullptr);
                st.parts_cached += plan.cached.size();
            }

            // 3. filter: each run of consecutive filter partitions (across the
            // combiner's tasks, which group at most k of them for the cost
            // model) streams through the ring in windows of ring_edges (one
            // copy per array per window); whole partitions of a window are
            // then promoted into the cache
            std::vector<std::uint64_t> filter_parts;
            for (const Task &t : plan.tasks)
                if (t.engine == HYT_ENGINE_FILTER)
                    filter_parts.insert(filter_parts.end(), t.members.begin(), t.members.end());
            std::sort(filter_parts.begin(), filter_parts.end());
            // the runs in partition order, or in contribution order (the
            // original's ContributionScheduler ordering over the runs' device-
            // computed contributions: SSSP -min frontier distance of the run,
            // PageRank the mass it pushes)
            std::vector<std::pair<std::uint64_t, std::uint64_t>> runs;
            for_runs(filter_parts, [&](std::uint64_t p0, std::uint64_t p1) { runs.emplace_back(p0, p1); });
            if (E.prio_mode != kPrioNone && runs.size() > 1)
            {
                std::vector<ContributionItem> items(runs.size());
                for (std::size_t i = 0; i < runs.size(); ++i)
                {
                    double c = 0.0;
                    if (E.prio_mode == kPrioMinValue)
                    {
                        unsigned long long mn = ~0ULL;
                        for (std::uint64_t p = runs[i].first; p <= runs[i].second; ++p)
                            mn = std::min(mn, E.part_prio_h[p]);
                        std::uint32_t bits = static_cast<std::uint32_t>(mn);
                        if (E.weights_f32)
                        {
                            float f;
                            std::memcpy(&f, &bits, 4);
                            c = -static_cast<double>(f);
                        }
                        else
                            c = -static_cast<double>(bits);
                    }
                    else
                        for (std::uint64_t p = runs[i].first; p <= runs[i].second; ++p)
                        {
                            double v;
                            std::memcpy(&v, &E.part_prio_h[p], 8);
                            c += v;
                        }
                    items[i] = ContributionItem{static_cast<std::size_t>(runs[i].first), c};
                }
                std::size_t moved = 0;
                const std::vector<std::size_t> order = contribution_order(items, true, &moved);
                std::vector<std::pair<std::uint64_t, std::uint64_t>> sorted;
                sorted.reserve(runs.size());
                for (const std::size_t k : order)
                    sorted.push_back(runs[k]);
                runs.swap(sorted);
                st.runs_reordered += moved;
            }
            {
                for (const auto &run : runs)
                {
                    const std::uint64_t p0 = run.first, p1 = run.second;
                    st.parts_filter += p1 - p0 + 1;
                    const std::uint64_t e0 = E.parts[p0].edge_begin, e1 = E.parts[p1].edge_end;
                    std::uint64_t pa = p0;
                    for (std::uint64_t a = e0; a < e1; a += E.ring_edges)
                    {
                        const std::uint64_t b = std::min<std::uint64_t>(e1, a + E.ring_edges);
                        while (E.parts[pa].edge_end <= a)
                            ++pa;
                        std::uint64_t pb = pa;
                        while (pb < p1 && E.parts[pb + 1].edge_begin < b)
                            ++pb;
                        const std::size_t r = E.ring_take();
                        E.copy_host_graph(E.ring_col(r), E.host->col + a, (b - a) * 4);
                        if (weighted)
                            E.copy_host_graph(E.ring_w(r), E.host->w + a, (b - a) * 4);
                        st.bytes_filter += (b - a) * E.run_edge_bytes;
                        hip_check(hipStreamWaitEvent(s, E.copy_events[E.copy_events_used - 1].b, 0), "wait copy");
                        x.window(plan.qlo[pa], rows_of(pa, pb), a, b, E.ring_col(r), E.ring_w(r));
                        if (E.cfg.cache)
                            for (std::uint64_t pp = pa; pp <= pb; ++pp)
                            {
                                const hyt_partition &pt = E.parts[pp];
                                const std::uint64_t n = pt.edge_end - pt.edge_begin;
                                if (pt.edge_begin < a || pt.edge_end > b || n > E.unit_edges)
                                    continue;
                                if (!E.admit(pp))
                                    break; // every resident partition is active this round
                                const Resident &rr = E.res[pp];
                                hip_check(hipMemcpyAsync(E.unit_ptr(rr.col), E.ring_col(r) + (pt.edge_begin - a), n * 4,
                                                         hipMemcpyDeviceToDevice, s),
                                          "promote columns");
                                if (weighted)
                                    hip_check(hipMemcpyAsync(E.unit_ptr(rr.w), E.ring_w(r) + (pt.edge_begin - a), n * 4,
                                                             hipMemcpyDeviceToDevice, s),
                                              "promote weights");
                            }
                        E.ring_release(r);
                    }
                }
            }

            // 4. zero-copy: every partition of the zero-copy task in one launch
            // over the mapped host arrays (segment table after the resident
            // one; a segment's rows sit at their own edge positions)
            for (const Task &t : plan.tasks)
            {
                if (t.engine != HYT_ENGINE_ZERO_COPY || t.members.empty())
                    continue;
                const std::size_t base = plan.cached.size();
                std::uint64_t rows = 0;
                std::size_t nz = 0;
                for (const std::uint64_t p : t.members)
                {
                    // estimated link bytes: m per modelled request
                    st.bytes_zero_copy += E.part_stats_h[4 * p + 2] * E.run_params.request_bytes;
                    CacheSeg &sg = E.seg_h[base + nz++];
                    sg.qlo = static_cast<std::uint32_t>(plan.qlo[p]);
                    sg.rows_before = static_cast<std::uint32_t>(rows);
                    sg.edge_begin = E.parts[p].edge_begin;
                    sg.unit_pos = E.parts[p].edge_begin;
                    rows += plan.qn[p];
                }
                hip_check(hipMemcpyAsync(E.seg_d.as<CacheSeg>() + base, E.seg_h.data() + base, nz * sizeof(CacheSeg),
                                         hipMemcpyHostToDevice, s),
                          "upload zero-copy segments");
                x.cached(E.seg_d.as<CacheSeg>() + base, static_cast<std::uint32_t>(nz), static_cast<std::uint32_t>(rows),
                         E.z_col, weighted ? E.z_w : nullptr);
                st.parts_zero_copy += nz;
            }

            // 5. compaction: pack the rows read back in step 1 on host threads,
            // in batches that fit the packed staging (the GPU works on the
            // expansions queued above meanwhile)
            std::uint64_t r0 = 0;
            for (const Task &t : plan.tasks)
                if (t.engine == HYT_ENGINE_COMPACTION)
                    st.parts_compaction += t.members.size();
            while (r0 < crows)
            {
                const auto tp = Clock::now();
                // the previous batch's kernels must be done with the staging
                hip_check(hipEventSynchronize(E.ev_pack), "wait packed staging");
                {
                    // a row larger than the packed staging is read zero-copy
                    // (a pageable graph: streamed through the ring in windows)
                    const std::uint32_t u = E.h_q[r0];
                    const std::uint64_t d = E.host->off[u + 1] - E.host->off[u];
                    if (d > E.pack_cap && E.pageable)
                    {
                        E.h_rows[0] = u;
                        E.h_val[0] = E.h_qd[r0];
                        E.copy_h2d(E.pk_rows.ptr, E.h_rows, 4);
                        E.copy_h2d(E.pk_val.ptr, E.h_val, 4);
                        const std::uint64_t e0 = E.host->off[u], e1 = E.host->off[u + 1];
                        for (std::uint64_t a = e0; a < e1; a += E.ring_edges)
                        {
                            const std::uint64_t b = std::min<std::uint64_t>(e1, a + E.ring_edges);
                            const std::size_t r = E.ring_take();
                            E.copy_host_graph(E.ring_col(r), E.host->col + a, (b - a) * 4);
                            if (weighted)
                                E.copy_host_graph(E.ring_w(r), E.host->w + a, (b - a) * 4);
                            hip_check(hipStreamWaitEvent(s, E.copy_events[E.copy_events_used - 1].b, 0),
                                      "wait row copy");
                            x.list(E.pk_rows.as<std::uint32_t>(), E.pk_val.as<std::uint32_t>(), 1U, a, b,
                                   E.ring_col(r), weighted ? E.ring_w(r) : nullptr);
                            E.ring_release(r);
                        }
                        hip_check(hipEventRecord(E.ev_pack, s), "hipEventRecord");
                        st.bytes_compaction += d * E.run_edge_bytes;
                        ++r0;
                        continue;
                    }
                    if (d > E.pack_cap)
                    {
                        E.h_rows[0] = u;
                        E.h_val[0] = E.h_qd[r0];
                        E.copy_h2d(E.pk_rows.ptr, E.h_rows, 4);
                        E.copy_h2d(E.pk_val.ptr, E.h_val, 4);
                        hip_check(hipStreamWaitEvent(s, E.copy_events[E.copy_events_used - 1].b, 0), "wait row copy");
                        x.list(E.pk_rows.as<std::uint32_t>(), E.pk_val.as<std::uint32_t>(), 1U, 0, ~0ULL, E.z_col,
                               weighted ? E.z_w : nullptr);
                        hip_check(hipEventRecord(E.ev_pack, s), "hipEventRecord");
                        st.bytes_zero_copy += d * E.run_edge_bytes;
                        ++r0;
                        continue;
                    }
                }
                std::uint64_t r1 = r0, edges = 0;
                while (r1 < crows)
                {
                    const std::uint32_t u = E.h_q[r1];
                    const std::uint64_t d = E.host->off[u + 1] - E.host->off[u];
                    if (r1 > r0 && (d > E.pack_cap || edges + d > E.pack_cap || r1 - r0 >= E.pack_cap))
                        break;
                    edges += d;
                    ++r1;
                }
                const std::uint64_t nr = r1 - r0;
                // prefix of degrees, then parallel copy of the rows' edges
                E.h_index[0] = 0;
                for (std::uint64_t k = 0; k < nr; ++k)
                {
                    const std::uint32_t u = E.h_q[r0 + k];
                    E.h_rows[k] = u;
                    E.h_val[k] = E.h_qd[r0 + k];
                    E.h_index[k + 1] = E.h_index[k] + (E.host->off[u + 1] - E.host->off[u]);
                }
                std::vector<std::thread> pool;
                const int T = std::max(1, std::min<int>(E.threads, static_cast<int>(nr / 64 + 1)));
                for (int ti = 0; ti < T; ++ti)
                    pool.emplace_back([&, ti] {
                        const std::uint64_t lo = nr * ti / T, hi = nr * (ti + 1) / T;
                        for (std::uint64_t k = lo; k < hi; ++k)
                        {
                            const std::uint32_t u = E.h_rows[k];
                            const std::uint64_t b = E.host->off[u], n = E.host->off[u + 1] - b;
                            std::memcpy(E.h_col + E.h_index[k], E.host->col + b, n * 4);
                            if (weighted)
                                std::memcpy(E.h_w + E.h_index[k], E.host->w + b, n * 4);
                        }
                    });
                for (auto &th : pool)
                    th.join();
                st.pack_ms += ms_since(tp);
                E.copy_h2d(E.pk_rows.ptr, E.h_rows, nr * 4);
                E.copy_h2d(E.pk_val.ptr, E.h_val, nr * 4);
                E.copy_h2d(E.pk_index.ptr, E.h_index, (nr + 1) * 8);
                E.copy_h2d(E.pk_col.ptr, E.h_col, edges * 4);
                if (weighted)
                    E.copy_h2d(E.pk_w.ptr, E.h_w, edges * 4);
                st.bytes_compaction += nr * 4 + (nr + 1) * 8 + edges * E.run_edge_bytes;
                hip_check(hipStreamWaitEvent(s, E.copy_events[E.copy_events_used - 1].b, 0), "wait pack copy");
                x.packed(E.pk_rows.as<std::uint32_t>(), E.pk_index.as<std::uint64_t>(), E.pk_val.as<std::uint32_t>(),
                         static_cast<std::uint32_t>(nr), E.pk_col.as<std::uint32_t>(),
                         weighted ? E.pk_w.as<std::uint32_t>() : nullptr);
                hip_check(hipEventRecord(E.ev_pack, s), "hipEventRecord");
                r0 = r1;
            }
        }

        void finish_round_stats(HybridEngine &E, hyt_hybrid_stats &st)
        {
            hip_check(hipStreamSynchronize(E.copy), "copy stream");
            st.copy_ms += E.drain_copy_time();
        }

        // --------------------------------------------- PageRank pull rounds
        // A launch of the pull kernels covers at most this many in-edges (or a
        // ring window, if larger), so their chunk queue has a fixed size.
        constexpr std::uint64_t kPullLaunchEdges = 1ULL << 26;

        // The host graph's incoming CSR built on the device, when the graph
        // and the sort's working arrays fit in free HBM (about 44 bytes per
        // edge, E < 2^31): the in-HBM transpose (a radix sort of (destination,
        // source) keys, so each destination's sources come out ascending, the
        // order of the host build) copied back into the graph's memory kind.
        // False when it does not fit; the host build runs instead.
        template <class OffT>
        bool device_transpose(HybridEngine &E)
        {
            HostGraph &h = *E.host;
            const std::uint64_t V = h.V, Ed = h.E;
            if (h.in_off != nullptr || Ed == 0 || Ed >= (1ULL << 31))
                return false;
            std::size_t free_b = 0, total_b = 0;
            hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
            const std::uint64_t need = Ed * 44 + (V + 1) * 16 + (256ULL << 20);
            if (need > free_b / 10 * 9)
                return false;
            DeviceGraph t;
            t.device = E.dg.device;
            t.V = V;
            t.E = Ed;
            t.off64 = sizeof(OffT) == 8;
            hip_check(hipStreamCreateWithFlags(&t.stream, hipStreamNonBlocking), "hipStreamCreate");
            t.off.alloc((V + 1) * sizeof(OffT), "hipMalloc(transpose offsets)");
            if (sizeof(OffT) == 8)
                hip_check(hipMemcpy(t.off.ptr, h.off, (V + 1) * 8, hipMemcpyHostToDevice), "copy offsets");
            else
            {
                std::vector<std::uint32_t> o(V + 1);
                for (std::uint64_t v = 0; v <= V; ++v)
                    o[v] = static_cast<std::uint32_t>(h.off[v]);
                hip_check(hipMemcpy(t.off.ptr, o.data(), (V + 1) * 4, hipMemcpyHostToDevice), "copy offsets");
            }
            t.col.alloc(Ed * 4, "hipMalloc(transpose columns)");
            hip_check(hipMemcpy(t.col.ptr, h.col, Ed * 4, hipMemcpyHostToDevice), "copy columns");
            build_incoming_csr(t);
            t.col.release();
            std::uint64_t *in_off = nullptr;
            std::uint32_t *in_col = nullptr;
            try
            {
                in_off = static_cast<std::uint64_t *>(host_alloc((V + 1) * 8, h.memory));
                in_col = static_cast<std::uint32_t *>(host_alloc(Ed * 4, h.memory));
                if (sizeof(OffT) == 8)
                    hip_check(hipMemcpy(in_off, t.in_off.ptr, (V + 1) * 8, hipMemcpyDeviceToHost), "read in offsets");
                else
                {
                    std::vector<std::uint32_t> o(V + 1);
                    hip_check(hipMemcpy(o.data(), t.in_off.ptr, (V + 1) * 4, hipMemcpyDeviceToHost),
                              "read in offsets");
                    for (std::uint64_t v = 0; v <= V; ++v)
                        in_off[v] = o[v];
                }
                hip_check(hipMemcpy(in_col, t.in_col.ptr, Ed * 4, hipMemcpyDeviceToHost), "read in columns");
            }
            catch (...)
            {
                host_free(in_off, h.memory);
                host_free(in_col, h.memory);
                throw;
            }
            h.in_off = in_off;
            h.in_col = in_col;
            return true;
        }

        // The incoming CSR on first use: transpose (on the device when it
        // fits, else on host threads), in-partitions, incoming offsets on the
        // device, pull kernel buffers. True when built by this call.
        template <class OffT>
        bool ensure_pull(HybridEngine &E)
        {
            if (!E.in_parts.empty())
                return false;
            const auto t0 = Clock::now();
            {
                std::lock_guard<std::recursive_mutex> lock(E.host->transpose_mu);
                if (!device_transpose<OffT>(E))
                    ensure_host_transpose(*E.host, E.threads);
            }
            E.transpose_ms = ms_since(t0);
            const std::uint64_t V = E.host->V;
            E.in_parts = partition_graph(E.host->in_off, V, E.cfg.partition_edges * 4);
            const std::uint64_t P = E.parts.size(), Pin = E.in_parts.size();
            E.in_off_d.alloc((V + 1) * sizeof(OffT), "hipMalloc(incoming offsets)");
            if (sizeof(OffT) == 8)
                hip_check(hipMemcpy(E.in_off_d.ptr, E.host->in_off, (V + 1) * 8, hipMemcpyHostToDevice),
                          "copy incoming offsets");
            else
            {
                std::vector<std::uint32_t> o(V + 1);
                for (std::uint64_t v = 0; v <= V; ++v)
                    o[v] = static_cast<std::uint32_t>(E.host->in_off[v]);
                hip_check(hipMemcpy(E.in_off_d.ptr, o.data(), (V + 1) * 4, hipMemcpyHostToDevice),
                          "copy incoming offsets");
            }
            E.res.resize(P + Pin, Resident{});
            E.last_active.resize(P + Pin, 0);
            E.pull_seg_h.resize(Pin);
            E.pull_seg_d.alloc(std::max<std::uint64_t>(Pin, 1) * sizeof(PullSeg), "hipMalloc(pull segments)");
            const std::uint64_t L =
                std::max<std::uint64_t>(E.ring_edges, std::min<std::uint64_t>(kPullLaunchEdges, E.host->E));
            // a row of d > light_max in-edges makes ceil(d / chunk) <= d / chunk
            // + d / (light_max + 1) chunks
            E.pull_heavy_cap = L / kPagerankHeavyChunk + L / (kPagerankLightMaxDegree + 1) + 64;
            E.pull_heavy.alloc(E.pull_heavy_cap * sizeof(HeavyChunk), "hipMalloc(pull chunks)");
            E.pull_ctr.alloc(256, "hipMalloc(pull counters)");
            hip_check(hipMemset(E.pull_ctr.ptr, 0, 256), "clear pull counters");
            return true;
        }

        // One pull round: every in-partition with in-edges adds the shares of
        // its in-edges to its rows' residuals; resident ones from their units,
        // the rest through the ring (then promoted into the cache).
        template <class OffT>
        void pull_round(HybridEngine &E, hyt_hybrid_stats &st)
        {
            DeviceGraph &g = E.dg;
            hipStream_t s = g.stream;
            ++E.stamp;
            const std::uint64_t P = E.parts.size();
            const OffT *in_off = E.in_off_d.as<OffT>();
            const float *share = g.ws.share.as<float>();
            unsigned long long *res = g.ws.res.as<unsigned long long>();
            std::uint32_t *ctr = E.pull_ctr.as<std::uint32_t>();
            const PullQueue pq{E.pull_heavy.as<HeavyChunk>(), ctr, ctr + 1, E.pull_heavy_cap};
            const std::uint64_t launch_edges =
                std::max<std::uint64_t>(E.ring_edges, std::min<std::uint64_t>(kPullLaunchEdges, E.host->E));
            auto launch = [&](const PullSeg *segs, const PullSeg &one, std::uint64_t nseg, std::uint64_t rows) {
                if (rows == 0 || nseg == 0)
                    return;
                hip_check(hipMemsetAsync(ctr, 0, 4, s), "reset pull chunk count");
                const dim3 grid(static_cast<unsigned>((rows + kBlock - 1) / kBlock), static_cast<unsigned>(nseg));
                pr_pull_seg<OffT><<<grid, kBlock, 0, s>>>(in_off, segs, one, share, res, kPagerankLightMaxDegree,
                                                          kPagerankHeavyChunk, pq);
                hip_check(hipGetLastError(), "launch pr_pull_seg");
                pr_pull_seg_heavy<<<kHeavyGrid, kBlock, 0, s>>>(segs, one, share, res, pq);
                hip_check(hipGetLastError(), "launch pr_pull_seg_heavy");
            };

            std::vector<std::uint64_t> cached, streamed;
            for (std::uint64_t q = 0; q < E.in_parts.size(); ++q)
            {
                const hyt_partition &pt = E.in_parts[q];
                if (pt.edge_end == pt.edge_begin)
                    continue;
                E.last_active[P + q] = E.stamp;
                (E.is_cached(P + q) ? cached : streamed).push_back(q);
            }

            // resident in-partitions: one segment table, in groups of about a
            // ring window's in-edges (at most launch_edges, 65535 segments),
            // launched one after each streamed window: the compute stream
            // then frees ring buffers as fast as the copies fill them, where
            // one long launch first would leave the copy stream waiting
            struct Group
            {
                std::size_t k0, k1;
                std::uint64_t rows;
            };
            std::vector<Group> groups;
            if (!cached.empty())
            {
                for (std::size_t k = 0; k < cached.size(); ++k)
                {
                    const hyt_partition &pt = E.in_parts[cached[k]];
                    E.pull_seg_h[k] = PullSeg{pt.vertex_begin, pt.vertex_end, pt.edge_begin, pt.edge_end,
                                              E.unit_ptr(E.res[P + cached[k]].col)};
                }
                // pull_seg_h is rewritten only after the round's final synchronisation
                hip_check(hipMemcpyAsync(E.pull_seg_d.ptr, E.pull_seg_h.data(), cached.size() * sizeof(PullSeg),
                                         hipMemcpyHostToDevice, s),
                          "upload pull segments");
                const std::uint64_t group_edges = std::min<std::uint64_t>(E.ring_edges, launch_edges);
                std::size_t k0 = 0;
                while (k0 < cached.size())
                {
                    std::size_t k1 = k0;
                    std::uint64_t edges = 0, rows = 0;
                    while (k1 < cached.size() && k1 - k0 < 65535)
                    {
                        const hyt_partition &pt = E.in_parts[cached[k1]];
                        const std::uint64_t n = pt.edge_end - pt.edge_begin;
                        if (k1 > k0 && edges + n > group_edges)
                            break;
                        edges += n;
                        rows = std::max<std::uint64_t>(rows, pt.vertex_end - pt.vertex_begin);
                        ++k1;
                    }
                    groups.push_back(Group{k0, k1, rows});
                    k0 = k1;
                }
                st.parts_cached += cached.size();
            }
            std::size_t next_group = 0;
            auto launch_group = [&] {
                const Group &gr = groups[next_group++];
                launch(E.pull_seg_d.as<PullSeg>() + gr.k0, PullSeg{}, gr.k1 - gr.k0, gr.rows);
            };

            // the rest: runs of consecutive in-partitions in ring windows that
            // end at a partition boundary (a partition longer than a window is
            // split over several)
            std::size_t i = 0;
            while (i < streamed.size())
            {
                std::size_t j = i + 1;
                while (j < streamed.size() && streamed[j] == streamed[j - 1] + 1)
                    ++j;
                const std::uint64_t q0 = streamed[i], q1 = streamed[j - 1];
                i = j;
                st.parts_filter += q1 - q0 + 1;
                const std::uint64_t e1 = E.in_parts[q1].edge_end;
                std::uint64_t qa = q0;
                for (std::uint64_t a = E.in_parts[q0].edge_begin; a < e1;)
                {
                    while (E.in_parts[qa].edge_end <= a)
                        ++qa;
                    const std::uint64_t lim = std::min<std::uint64_t>(e1, a + E.ring_edges);
                    std::uint64_t qb = qa, b = lim;
                    if (E.in_parts[qa].edge_end <= lim)
                    {
                        while (qb < q1 && E.in_parts[qb + 1].edge_end <= lim)
                            ++qb;
                        b = E.in_parts[qb].edge_end;
                    }
                    const std::size_t r = E.ring_take();
                    E.copy_host_graph(E.ring_col(r), E.host->in_col + a, (b - a) * 4);
                    st.bytes_filter += (b - a) * 4;
                    hip_check(hipStreamWaitEvent(s, E.copy_events[E.copy_events_used - 1].b, 0), "wait copy");
                    const PullSeg one{E.in_parts[qa].vertex_begin, E.in_parts[qb].vertex_end, a, b, E.ring_col(r)};
                    launch(nullptr, one, 1, one.v1 - one.v0);
                    if (E.cfg.cache)
                        for (std::uint64_t qq = qa; qq <= qb; ++qq)
                        {
                            const hyt_partition &pt = E.in_parts[qq];
                            const std::uint64_t n = pt.edge_end - pt.edge_begin;
                            if (pt.edge_begin < a || pt.edge_end > b || n > E.unit_edges || n == 0)
                                continue;
                            if (!E.admit(P + qq))
                                break; // every resident partition is in use this round
                            hip_check(hipMemcpyAsync(E.unit_ptr(E.res[P + qq].col), E.ring_col(r) + (pt.edge_begin - a),
                                                     n * 4, hipMemcpyDeviceToDevice, s),
                                      "promote incoming columns");
                        }
                    E.ring_release(r);
                    // after the release, so the buffer's next copy does not wait for it
                    if (next_group < groups.size())
                        launch_group();
                    a = b;
                }
            }
            while (next_group < groups.size())
                launch_group();
        }

        template <class OffT, int WM>
        void hybrid_sssp_typed(HybridEngine &E, std::uint32_t src, std::uint32_t *dist_out, const hyt_run_options &o,
                               hyt_hybrid_stats &st)
        {
            DeviceGraph &g = E.dg;
            Workspace &ws = g.ws;
            hipStream_t s = g.stream;
            Counters *d_ctr = ws.ctr.as<Counters>();
            launch_fill<std::uint32_t>(ws.dist.as<std::uint32_t>(), g.V, Weights<WM>::kInf, s);
            launch_fill<std::uint32_t>(ws.prev.as<std::uint32_t>(), g.V, Weights<WM>::kInf, s);
            hybrid_seed<<<1, 64, 0, s>>>(ws.dist.as<std::uint32_t>(), ws.touch.as<std::uint8_t>(), src);
            hip_check(hipGetLastError(), "launch seed");
            hip_check(hipMemsetAsync(d_ctr, 0, sizeof(Counters), s), "reset counters");
            SsspSel sel{ws.touch.as<std::uint8_t>(), ws.dist.as<std::uint32_t>(), ws.prev.as<std::uint32_t>(),
                        0xFFFFFFFFU};
            Counters c;
            std::uint32_t F = ordered_compaction<OffT>(E, sel, c);
            auto make = [&](const std::uint32_t *col, const std::uint32_t *w, const std::uint32_t *qd) {
                SsspOp<WM, false> op;
                op.col = col;
                op.w = w;
                op.dist = ws.dist.as<std::uint32_t>();
                op.touch = ws.touch.as<std::uint8_t>();
                op.qd = qd;
                op.overflow = &d_ctr->overflow;
                op.expanded_at = ws.prev.as<std::uint32_t>();
                return op;
            };
            const std::uint64_t max_rounds = o.max_rounds ? o.max_rounds : g.V;
            Plan plan;
            st.run.converged = 0;
            for (std::uint64_t round = 0; round < max_rounds && F > 0; ++round)
            {
                plan_round<OffT>(E, F, plan, st);
                hip_check(hipMemsetAsync(d_ctr, 0, kRoundCounterBytes, s), "reset counters");
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev0, s), "hipEventRecord");
                execute_round<OffT>(E, plan, single_expander<OffT>(E, make), st);
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev1, s), "hipEventRecord");
                Counters cc;
                const std::uint32_t Fn = ordered_compaction<OffT>(E, sel, cc);
                finish_round_stats(E, st);
                if (o.time_kernels)
                {
                    float ms = 0;
                    hip_check(hipEventElapsedTime(&ms, g.ev0, g.ev1), "hipEventElapsedTime");
                    st.run.expand_ms += ms;
                    st.run.expand_launches += 1;
                }
                if (cc.heavy_overflow)
                    fail(HYT_ERR_INTERNAL, "heavy chunk queue overflow");
                if (cc.overflow)
                {
                    st.run.overflow = 1;
                    fail(HYT_ERR_OVERFLOW, "u32 distance sum exceeded 2^32-2");
                }
                st.run.frontier_total += F;
                st.run.edges_total += cc.all_edges();
                st.run.updates_total += Fn;
                st.run.rounds = round + 1;
                F = Fn;
                if (F == 0)
                    st.run.converged = 1;
            }
            hip_check(hipMemcpyAsync(dist_out, ws.dist.ptr, g.V * 4,
                                     o.out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s),
                      "copy distances");
            hip_check(hipStreamSynchronize(s), "hybrid sssp output");
        }

        template <class OffT>
        void hybrid_pagerank_typed(HybridEngine &E, float damping, float eps, float *out, const hyt_run_options &o,
                                   hyt_hybrid_stats &st)
        {
            DeviceGraph &g = E.dg;
            Workspace &ws = g.ws;
            hipStream_t s = g.stream;
            Counters *d_ctr = ws.ctr.as<Counters>();
            ensure_mass_state(g);
            const MassScale m = mass_scale(g.V, damping, eps);
            launch_fill<unsigned long long>(ws.res.as<unsigned long long>(), g.V, m.r0, s);
            launch_fill<unsigned long long>(ws.rank.as<unsigned long long>(), g.V, 0ULL, s);
            hip_check(hipMemsetAsync(d_ctr, 0, sizeof(Counters), s), "reset counters");
            // global PageRank: float shares, as in HBM
            PrSel<OffT, false> sel{ws.res.as<unsigned long long>(), ws.rank.as<unsigned long long>(),
                                   ws.share.as<float>(), g.off.as<OffT>(), damping, m.eps};
            Counters c;
            std::uint32_t F = ordered_compaction<OffT>(E, sel, c);
            auto make = [&](const std::uint32_t *col, const std::uint32_t *, const std::uint32_t *qd) {
                PrOp<false> op;
                op.col = col;
                op.res = ws.res.as<unsigned long long>();
                op.share = nullptr;
                op.qd = qd;
                return op;
            };
            const std::uint64_t max_rounds = o.max_rounds ? o.max_rounds : 100000;
            // direction per round as in HBM: pull when the frontier's out-edges
            // reach pull_percent of E (pull rounds read every in-edge)
            const std::uint32_t dir = o.direction;
            require(dir <= HYT_DIR_PULL, "direction must be auto (0), push (1) or pull (2)");
            const std::uint64_t pct = o.pull_percent ? o.pull_percent : kDefaultPullPercentPagerank;
            const bool can_pull = dir != HYT_DIR_PUSH && g.E > 0;
            unsigned long long Fe = c.frontier_edges;
            Plan plan;
            st.run.converged = F == 0 ? 1 : 0;
            for (std::uint64_t round = 0; round < max_rounds && F > 0; ++round)
            {
                const bool pull = can_pull && (dir == HYT_DIR_PULL || Fe * 100ULL >= pct * g.E);
                if (pull)
                {
                    if (ensure_pull<OffT>(E))
                        st.transpose_ms = E.transpose_ms;
                }
                else
                    plan_round<OffT>(E, F, plan, st);
                hip_check(hipMemsetAsync(d_ctr, 0, kRoundCounterBytes, s), "reset counters");
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev0, s), "hipEventRecord");
                if (pull)
                    pull_round<OffT>(E, st);
                else
                    execute_round<OffT>(E, plan, single_expander<OffT>(E, make), st);
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev1, s), "hipEventRecord");
                Counters cc;
                const std::uint32_t Fn = ordered_compaction<OffT>(E, sel, cc);
                finish_round_stats(E, st);
                if (pull)
                {
                    std::uint32_t pov = 0;
                    hip_check(hipMemcpy(&pov, E.pull_ctr.as<std::uint32_t>() + 1, 4, hipMemcpyDeviceToHost),
                              "read pull overflow");
                    if (pov)
                        fail(HYT_ERR_INTERNAL, "pull chunk queue overflow");
                    st.run.pull_rounds += 1;
                    st.run.pull_edges += g.E;
                    st.run.pull_frontier += F;
                    st.run.pull_active_edges += Fe;
                }
                if (o.time_kernels)
                {
                    float ms = 0;
                    hip_check(hipEventElapsedTime(&ms, g.ev0, g.ev1), "hipEventElapsedTime");
                    st.run.expand_ms += ms;
                    st.run.expand_launches += 1;
                }
                if (cc.heavy_overflow)
                    fail(HYT_ERR_INTERNAL, "heavy chunk queue overflow");
                st.run.frontier_total += F;
                st.run.edges_total += pull ? g.E : cc.all_edges();
                st.run.updates_total += Fn;
                st.run.rounds = round + 1;
                F = Fn;
                Fe = cc.frontier_edges;
                if (F == 0)
                    st.run.converged = 1;
            }
            finalize_ranks(g, out, o.out_on_device != 0, st.run);
        }

        // ------------------------------------------------------ K-lane runs
        // K queries per run with [V][K] state (the kernels of lanes.hpp): the
        // frontier is the sorted queue of rows with their query masks, so the
        // planner sees each row once whatever its mask, and every partition
        // that moves (or is read zero-copy) is walked once for all K queries.

        // Expander of K-lane rounds: row values are queue indices, from which
        // the kernels take the row's mask and PPR snapshot row.
        template <class OffT, int KP, class Op>
        struct LaneExpander
        {
            static constexpr bool kIndexValues = true;
            HybridEngine &E;
            Op op;
            lanes::KArgs a;
            bool items; // item form this round

            static void set_edges(Op &o, const std::uint32_t *col, const std::uint32_t *w)
            {
                o.col = col;
                if constexpr (!std::is_same_v<Op, lanes::PprKOp>)
                    o.w = w;
                else
                    (void)w;
            }
            template <class RowSrc>
            void run(const RowSrc &rows, const std::uint32_t *col, const std::uint32_t *w, std::uint32_t F) const
            {
                if (F == 0)
                    return;
                hipStream_t s = E.dg.stream;
                Op o = op;
                set_edges(o, col, w);
                lanes::KArgs k = a;
                k.F = F;
                hip_check(hipMemsetAsync(&k.ctr->heavy_count, 0, 4, s), "reset heavy count");
                if (items)
                {
                    lanes::expandk_light_items<RowSrc, Op><<<blocks_for(F), kBlock, 0, s>>>(rows, o, k);
                    hip_check(hipGetLastError(), "launch hybrid expandk_light_items");
                    lanes::expandk_heavy_items<Op><<<kHeavyGrid, kBlock, 0, s>>>(o, k);
                    hip_check(hipGetLastError(), "launch hybrid expandk_heavy_items");
                }
                else
                {
                    lanes::expandk_light<RowSrc, KP, Op><<<blocks_for(F), kBlock, 0, s>>>(rows, o, k);
                    hip_check(hipGetLastError(), "launch hybrid expandk_light");
                    lanes::expandk_heavy<KP, Op><<<kHeavyGrid, kBlock, 0, s>>>(o, k);
                    hip_check(hipGetLastError(), "launch hybrid expandk_heavy");
                }
            }
            lanes::WindowRowsK<OffT> queue_rows() const
            {
                lanes::WindowRowsK<OffT> r{};
                r.off = E.dg.off.template as<OffT>();
                r.q = E.dg.ws.q.template as<std::uint32_t>();
                r.qmask = E.lk.qmask.template as<unsigned long long>();
                return r;
            }
            void window(std::uint64_t qa, std::uint32_t F, std::uint64_t lo, std::uint64_t hi,
                        const std::uint32_t *col, const std::uint32_t *w) const
            {
                lanes::WindowRowsK<OffT> r = queue_rows();
                r.qa = static_cast<std::uint32_t>(qa);
                r.lo = lo;
                r.hi = hi;
                run(r, col, w, F);
            }
            void list(const std::uint32_t *rows, const std::uint32_t *vals, std::uint32_t F, std::uint64_t lo,
                      std::uint64_t hi, const std::uint32_t *col, const std::uint32_t *w) const
            {
                lanes::WindowRowsK<OffT> r = queue_rows();
                r.rows = rows;
                r.idx = vals;
                r.lo = lo;
                r.hi = hi;
                run(r, col, w, F);
            }
            void packed(const std::uint32_t *rows, const std::uint64_t *index, const std::uint32_t *vals,
                        std::uint32_t F, const std::uint32_t *col, const std::uint32_t *w) const
            {
                run(lanes::PackedRowsK{rows, vals, index, E.lk.qmask.template as<unsigned long long>()}, col, w, F);
            }
            void cached(const CacheSeg *seg, std::uint32_t nseg, std::uint32_t F, const std::uint32_t *col,
                        const std::uint32_t *w) const
            {
                run(lanes::CachedRowsK<OffT>{E.dg.off.template as<OffT>(), E.dg.ws.q.template as<std::uint32_t>(),
                                             E.lk.qmask.template as<unsigned long long>(), seg, nseg},
                    col, w, F);
            }
        };

        void prepare_lanes(HybridEngine &E, std::uint32_t K, bool ppr)
        {
            DeviceGraph &g = E.dg;
            LaneState &L = E.lk;
            const std::uint64_t VK = g.V * K;
            // SSSP: u32 distances; PPR: u64 residual, rank and snapshot words
            L.val.ensure(VK * (ppr ? 8 : 4), "hipMalloc([V][K] state)");
            if (ppr)
            {
                L.rank.ensure(VK * 8, "hipMalloc([V][K] rank)");
                L.snap.ensure(VK * 8, "hipMalloc([V][K] snapshot)");
            }
            if (L.mark.bytes < (g.V + kPad) * 8)
            {
                // mark stays all-zero between runs (the compaction clears what it reads)
                L.mark.alloc((g.V + kPad) * 8, "hipMalloc(lane marks)");
                L.qmask.alloc((g.V + kPad) * 8, "hipMalloc(queue masks)");
                hip_check(hipMemsetAsync(L.mark.ptr, 0, (g.V + kPad) * 8, g.stream), "clear marks");
            }
            L.stage_mask.ensure(static_cast<std::uint64_t>(compact_blocks(g.V)) * kCompactRowsPerBlock * 8,
                                "hipMalloc(stage masks)");
            const std::uint64_t cap = lanes::lane_chunk_capacity(g.V, g.E, E.light_max, E.chunk) +
                                      static_cast<std::uint64_t>(K) * lanes::lane_item_chunks_per_lane(g.E, E.chunk);
            if (L.heavy.bytes < cap * sizeof(lanes::HeavyChunkK))
                L.heavy.alloc(cap * sizeof(lanes::HeavyChunkK), "hipMalloc(heavy chunks)");
            L.heavy_cap = L.heavy.bytes / sizeof(lanes::HeavyChunkK);
            L.seeds.ensure(64 * 4 * 4, "hipMalloc(seeds)");
            L.lanes.ensure(8, "hipMalloc(lane mask)");
            L.partial.ensure((1024 + 1) * 64 * sizeof(double), "hipMalloc(partials)");
        }

        // Mark compaction in vertex order: rows with a non-zero mark -> sorted
        // queue + masks; ctr->frontier_edges = (row, query) pairs; optionally
        // the OR of the kept masks (E.lk.lanes accumulates it).
        std::uint32_t ordered_compaction_k(HybridEngine &E, Counters &out, std::uint64_t *lanes_or = nullptr)
        {
            DeviceGraph &g = E.dg;
            hipStream_t s = g.stream;
            Counters *d_ctr = g.ws.ctr.as<Counters>();
            const unsigned nb = compact_blocks(g.V);
            lanes::MarkSel sel{E.lk.mark.as<unsigned long long>(), E.lk.lanes.as<unsigned long long>()};
            compact_rows<unsigned long long, std::uint32_t, lanes::MarkSel><<<nb, kBlock, 0, s>>>(
                g.V, sel, E.stage_q.as<std::uint32_t>(), E.lk.stage_mask.as<unsigned long long>(), d_ctr, nullptr,
                nullptr, E.block_counts.as<std::uint32_t>());
            hip_check(hipGetLastError(), "launch ordered mark compaction");
            std::size_t temp = E.scan_temp.bytes;
            hip_check(hipcub::DeviceScan::ExclusiveSum(E.scan_temp.ptr, temp, E.block_counts.as<std::uint32_t>(),
                                                       E.block_offsets.as<std::uint32_t>(), nb + 1, s),
                      "scan block counts");
            gather_ordered<unsigned long long><<<nb, kBlock, 0, s>>>(
                E.stage_q.as<std::uint32_t>(), E.lk.stage_mask.as<unsigned long long>(),
                E.block_counts.as<std::uint32_t>(), E.block_offsets.as<std::uint32_t>(), nb,
                g.ws.q.as<std::uint32_t>(), E.lk.qmask.as<unsigned long long>(), d_ctr);
            hip_check(hipGetLastError(), "launch gather");
            Counters *h = g.host_ctr->as<Counters>();
            hip_check(hipMemcpyAsync(h, d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, s), "read counters");
            if (lanes_or)
                hip_check(hipMemcpyAsync(E.lk.host_lanes->ptr, E.lk.lanes.ptr, 8, hipMemcpyDeviceToHost, s),
                          "read lanes");
            hip_check(hipStreamSynchronize(s), "hybrid K round");
            out = *h;
            if (lanes_or)
                std::memcpy(lanes_or, E.lk.host_lanes->ptr, 8);
            return h->next_count;
        }

        template <class OffT, int KP, class Op>
        void hybrid_k_rounds(HybridEngine &E, const Op &op, std::uint32_t K, bool ppr, const hyt_run_options &o,
                             hyt_hybrid_stats &st)
        {
            DeviceGraph &g = E.dg;
            LaneState &L = E.lk;
            hipStream_t s = g.stream;
            Counters *d_ctr = g.ws.ctr.as<Counters>();
            hip_check(hipMemsetAsync(d_ctr, 0, sizeof(Counters), s), "reset counters");
            Counters c;
            std::uint32_t F = ordered_compaction_k(E, c);
            std::uint64_t pairs = c.frontier_edges;
            lanes::KArgs a{};
            a.rs = nullptr;
            a.ctr = d_ctr;
            a.heavy = L.heavy.as<lanes::HeavyChunkK>();
            a.heavy_cap = L.heavy_cap;
            a.light_max = E.light_max;
            a.chunk = E.chunk;
            a.K = K;
            a.mark = L.mark.as<unsigned long long>();
            const std::uint64_t max_rounds = o.max_rounds ? o.max_rounds : (ppr ? 100000 : g.V);
            Plan plan;
            st.run.converged = F == 0 ? 1 : 0;
            for (std::uint64_t round = 0; round < max_rounds && F > 0; ++round)
            {
                plan_round<OffT>(E, F, plan, st);
                hip_check(hipMemsetAsync(d_ctr, 0, kRoundCounterBytes, s), "reset counters");
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev0, s), "hipEventRecord");
                if (ppr)
                {
                    lanes::pprk_snapshot<OffT><<<lanes::snapshot_blocks(static_cast<std::uint64_t>(F) * K), kBlock, 0,
                                                 s>>>(g.ws.q.as<std::uint32_t>(), F, K, L.qmask.as<unsigned long long>(),
                                                      L.val.as<unsigned long long>(), L.rank.as<unsigned long long>(),
                                                      L.snap.as<unsigned long long>(), g.off.as<OffT>(), op.damping);
                    hip_check(hipGetLastError(), "launch pprk_snapshot");
                }
                const LaneExpander<OffT, KP, Op> x{E, op, a, lanes::use_items(o.lane_form, pairs, F, KP)};
                execute_round<OffT>(E, plan, x, st);
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev1, s), "hipEventRecord");
                Counters cc;
                const std::uint32_t Fn = ordered_compaction_k(E, cc);
                finish_round_stats(E, st);
                if (o.time_kernels)
                {
                    float ms = 0;
                    hip_check(hipEventElapsedTime(&ms, g.ev0, g.ev1), "hipEventElapsedTime");
                    st.run.expand_ms += ms;
                    st.run.expand_launches += 1;
                }
                if (cc.heavy_overflow)
                    fail(HYT_ERR_INTERNAL, "heavy chunk queue overflow");
                if (cc.overflow)
                {
                    st.run.overflow = 1;
                    fail(HYT_ERR_OVERFLOW, "u32 distance sum exceeded 2^32-2");
                }
                st.run.frontier_total += F;
                st.run.edges_total += cc.all_edges();
                st.run.updates_total += Fn;
                st.run.rounds = round + 1;
                F = Fn;
                pairs = cc.frontier_edges;
                if (F == 0)
                    st.run.converged = 1;
            }
        }

        template <class OffT, class Op>
        void hybrid_k_dispatch(HybridEngine &E, const Op &op, std::uint32_t K, bool ppr, const hyt_run_options &o,
                               hyt_hybrid_stats &st)
        {
            switch (lanes::pow2_at_least(K))
            {
            case 2: hybrid_k_rounds<OffT, 2>(E, op, K, ppr, o, st); break;
            case 4: hybrid_k_rounds<OffT, 4>(E, op, K, ppr, o, st); break;
            case 8: hybrid_k_rounds<OffT, 8>(E, op, K, ppr, o, st); break;
            case 16: hybrid_k_rounds<OffT, 16>(E, op, K, ppr, o, st); break;
            case 32: hybrid_k_rounds<OffT, 32>(E, op, K, ppr, o, st); break;
            default: hybrid_k_rounds<OffT, 64>(E, op, K, ppr, o, st); break;
            }
        }

        template <class Op>
        void hybrid_k_run(HybridEngine &E, const Op &op, std::uint32_t K, bool ppr, const hyt_run_options &o,
                          hyt_hybrid_stats &st)
        {
            if (E.dg.off64)
                hybrid_k_dispatch<std::uint64_t>(E, op, K, ppr, o, st);
            else
                hybrid_k_dispatch<std::uint32_t>(E, op, K, ppr, o, st);
        }

        void hybrid_sssp_k(HybridEngine &E, const std::uint32_t *sources, std::uint32_t K, std::uint32_t *dist_out,
                           const hyt_run_options &o, hyt_hybrid_stats &st)
        {
            DeviceGraph &g = E.dg;
            LaneState &L = E.lk;
            prepare_lanes(E, K, false);
            const int wm = g.weight_mode();
            const std::uint32_t inf = (wm == kWF32 || wm == kWUnitF32) ? 0x7F800000U : 0xFFFFFFFFU;
            launch_fill<std::uint32_t>(L.val.as<std::uint32_t>(), g.V * K, inf, g.stream);
            const lanes::SeedPlan p = lanes::plan_seeds(sources, K);
            lanes::upload_seeds<std::uint32_t>(g.stream, L.seeds, L.mark.as<unsigned long long>(), p, sources, K,
                                               L.val.as<std::uint32_t>(), 0U);
            auto run = [&](auto op) {
                op.dist = L.val.as<std::uint32_t>();
                op.K = K;
                op.overflow = &g.ws.ctr.as<Counters>()->overflow;
                hybrid_k_run(E, op, K, false, o, st);
            };
            switch (wm)
            {
            case kWU32: run(lanes::SsspKOp<kWU32>{}); break;
            case kWF32: run(lanes::SsspKOp<kWF32>{}); break;
            case kWUnitU32: run(lanes::SsspKOp<kWUnitU32>{}); break;
            default: run(lanes::SsspKOp<kWUnitF32>{}); break;
            }
            hip_check(hipMemcpyAsync(dist_out, L.val.ptr, g.V * K * 4,
                                     o.out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, g.stream),
                      "copy distances");
            hip_check(hipStreamSynchronize(g.stream), "hybrid sssp output");
        }

        void hybrid_ppr_k(HybridEngine &E, const std::uint32_t *seeds, std::uint32_t K, float damping, float eps,
                          float *out, const hyt_run_options &o, hyt_hybrid_stats &st)
        {
            DeviceGraph &g = E.dg;
            LaneState &L = E.lk;
            prepare_lanes(E, K, true);
            const std::uint64_t VK = g.V * K;
            const MassScale m = mass_scale(1, damping, eps);
            launch_fill<unsigned long long>(L.val.as<unsigned long long>(), VK, 0ULL, g.stream);
            launch_fill<unsigned long long>(L.rank.as<unsigned long long>(), VK, 0ULL, g.stream);
            const lanes::SeedPlan p = lanes::plan_seeds(seeds, K);
            lanes::upload_seeds<unsigned long long>(g.stream, L.seeds, L.mark.as<unsigned long long>(), p, seeds, K,
                                                    L.val.as<unsigned long long>(), m.r0);
            lanes::PprKOp op{};
            op.res = L.val.as<unsigned long long>();
            op.snap = L.snap.as<unsigned long long>();
            op.K = K;
            op.damping = damping;
            op.eps = m.eps;
            hybrid_k_run(E, op, K, true, o, st);
            float *dst = o.out_on_device ? out : L.snap.as<float>();
            lanes::pprk_finish(g.stream, L.rank.as<unsigned long long>(), op.res, g.V, K,
                               L.partial.as<unsigned long long>(), dst);
            if (!o.out_on_device)
                hip_check(hipMemcpyAsync(out, dst, VK * 4, hipMemcpyDeviceToHost, g.stream), "copy ppr");
            hip_check(hipStreamSynchronize(g.stream), "hybrid ppr output");
        }

        // ------------------------------------------------ lane sessions
        template <class OffT, int KP, class Op>
        void lane_round_t(HybridEngine &E, const Op &op, std::uint32_t K, bool ppr, std::uint32_t F,
                          std::uint64_t pairs, std::uint32_t form, hyt_hybrid_stats &st, HybridLaneRound &out)
        {
            DeviceGraph &g = E.dg;
            LaneState &L = E.lk;
            hipStream_t s = g.stream;
            Counters *d_ctr = g.ws.ctr.as<Counters>();
            Plan plan;
            plan_round<OffT>(E, F, plan, st);
            hip_check(hipMemsetAsync(d_ctr, 0, kRoundCounterBytes, s), "reset counters");
            hip_check(hipMemsetAsync(L.lanes.ptr, 0, 8, s), "reset lane mask");
            if (ppr)
            {
                lanes::pprk_snapshot<OffT><<<lanes::snapshot_blocks(static_cast<std::uint64_t>(F) * K), kBlock, 0, s>>>(
                    g.ws.q.as<std::uint32_t>(), F, K, L.qmask.as<unsigned long long>(), L.val.as<unsigned long long>(),
                    L.rank.as<unsigned long long>(), L.snap.as<unsigned long long>(), g.off.as<OffT>(), op.damping);
                hip_check(hipGetLastError(), "launch pprk_snapshot");
            }
            lanes::KArgs a{};
            a.rs = nullptr;
            a.ctr = d_ctr;
            a.heavy = L.heavy.as<lanes::HeavyChunkK>();
            a.heavy_cap = L.heavy_cap;
            a.light_max = E.light_max;
            a.chunk = E.chunk;
            a.K = K;
            a.mark = L.mark.as<unsigned long long>();
            out.items = lanes::use_items(form, pairs, F, KP);
            const LaneExpander<OffT, KP, Op> x{E, op, a, out.items};
            execute_round<OffT>(E, plan, x, st);
            ordered_compaction_k(E, out.c, &out.lanes);
            finish_round_stats(E, st);
        }

        template <class OffT, int KP>
        void lane_round_kind(HybridEngine &E, int kind, std::uint32_t K, std::uint32_t F, std::uint64_t pairs,
                             float damping, float eps, std::uint32_t form, hyt_hybrid_stats &st,
                             HybridLaneRound &out)
        {
            LaneState &L = E.lk;
            if (kind == HYT_QUERY_PPR)
            {
                lanes::PprKOp op{};
                op.res = L.val.as<unsigned long long>();
                op.snap = L.snap.as<unsigned long long>();
                op.K = K;
                op.damping = damping;
                op.eps = mass_scale(1, damping, eps).eps;
                lane_round_t<OffT, KP>(E, op, K, true, F, pairs, form, st, out);
                return;
            }
            auto run = [&](auto op) {
                op.dist = L.val.as<std::uint32_t>();
                op.K = K;
                op.overflow = &E.dg.ws.ctr.as<Counters>()->overflow;
                lane_round_t<OffT, KP>(E, op, K, false, F, pairs, form, st, out);
            };
            switch (E.dg.weight_mode())
            {
            case kWU32: run(lanes::SsspKOp<kWU32>{}); break;
            case kWF32: run(lanes::SsspKOp<kWF32>{}); break;
            case kWUnitU32: run(lanes::SsspKOp<kWUnitU32>{}); break;
            default: run(lanes::SsspKOp<kWUnitF32>{}); break;
            }
        }

        template <class OffT>
        void lane_round_kp(HybridEngine &E, int kind, std::uint32_t K, std::uint32_t F, std::uint64_t pairs,
                           float damping, float eps, std::uint32_t form, hyt_hybrid_stats &st, HybridLaneRound &out)
        {
            switch (lanes::pow2_at_least(K))
            {
            case 2: lane_round_kind<OffT, 2>(E, kind, K, F, pairs, damping, eps, form, st, out); break;
            case 4: lane_round_kind<OffT, 4>(E, kind, K, F, pairs, damping, eps, form, st, out); break;
            case 8: lane_round_kind<OffT, 8>(E, kind, K, F, pairs, damping, eps, form, st, out); break;
            case 16: lane_round_kind<OffT, 16>(E, kind, K, F, pairs, damping, eps, form, st, out); break;
            case 32: lane_round_kind<OffT, 32>(E, kind, K, F, pairs, damping, eps, form, st, out); break;
            default: lane_round_kind<OffT, 64>(E, kind, K, F, pairs, damping, eps, form, st, out); break;
            }
        }
    } // namespace

    HybridEngine &hybrid_engine(hyt_hybrid *h)
    {
        require(h != nullptr, "null hybrid engine");
        return h->e;
    }

    DeviceGraph &hybrid_device_graph(HybridEngine &E) { return E.dg; }

    std::uint64_t hybrid_engine_bytes(HybridEngine &E) { return E.serving_bytes(); }

    std::uint64_t hybrid_lane_state_bytes(HybridEngine &E, int kind, std::uint32_t K)
    {
        const std::uint64_t V = E.dg.V, Vp = V + kPad;
        // PPR: residual, rank and snapshot words (u64); SSSP: u32 distances
        const std::uint64_t per_lane = (kind == HYT_QUERY_PPR ? V * 24 : V * 4) +
                                       lanes::lane_item_chunks_per_lane(E.dg.E, kLaneHeavyChunk) *
                                           sizeof(lanes::HeavyChunkK);
        const std::uint64_t fixed = Vp * 16 + static_cast<std::uint64_t>(compact_blocks(V)) * kCompactRowsPerBlock * 8 +
                                    lanes::lane_chunk_capacity(V, E.dg.E, kLaneLightMaxDegree, kLaneHeavyChunk) *
                                        sizeof(lanes::HeavyChunkK) +
                                    (1024 + 1) * 64 * sizeof(double) + 1024 + 64;
        return fixed + per_lane * K;
    }

    HybridLaneView hybrid_lanes_open(HybridEngine &E, int kind, std::uint32_t K)
    {
        require(kind == HYT_QUERY_SSSP || kind == HYT_QUERY_PPR, "query kind must be SSSP or PPR");
        require(K >= 1 && K <= 64, "lanes must be in [1, 64]");
        hip_check(hipSetDevice(E.dg.device), "hipSetDevice");
        const bool ppr = kind == HYT_QUERY_PPR;
        E.begin_run(!ppr);
        E.set_split(hyt_run_options{}, kLaneLightMaxDegree, kLaneHeavyChunk);
        hip_check(hipStreamSynchronize(E.dg.stream), "lane session");
        E.release_single_query_state();
        prepare_lanes(E, K, ppr);
        LaneState &L = E.lk;
        if (!L.host_lanes)
            L.host_lanes = std::make_unique<PinnedBuffer>(8);
        const std::uint64_t VK = E.dg.V * K;
        const int wm = E.dg.weight_mode();
        if (ppr)
        {
            launch_fill<unsigned long long>(L.val.as<unsigned long long>(), VK, 0ULL, E.dg.stream);
            launch_fill<unsigned long long>(L.rank.as<unsigned long long>(), VK, 0ULL, E.dg.stream);
        }
        else
            launch_fill<std::uint32_t>(L.val.as<std::uint32_t>(), VK,
                                       (wm == kWF32 || wm == kWUnitF32) ? 0x7F800000U : 0xFFFFFFFFU, E.dg.stream);
        hip_check(hipMemsetAsync(L.mark.ptr, 0, L.mark.bytes, E.dg.stream), "clear marks");
        hip_check(hipStreamSynchronize(E.dg.stream), "lane state");
        return HybridLaneView{L.val.ptr, ppr ? L.rank.as<unsigned long long>() : nullptr,
                              L.mark.as<unsigned long long>()};
    }

    HybridLaneRound hybrid_lanes_round(HybridEngine &E, int kind, std::uint32_t K, std::uint32_t F,
                                       std::uint64_t pairs, bool merge, float damping, float eps,
                                       std::uint32_t form, hyt_hybrid_stats &st)
    {
        hip_check(hipSetDevice(E.dg.device), "hipSetDevice");
        DeviceGraph &g = E.dg;
        LaneState &L = E.lk;
        HybridLaneRound out{};
        if (merge)
        {
            if (F > 0)
            {
                mark_from_queue<<<blocks_for(F), kBlock, 0, g.stream>>>(
                    g.ws.q.as<std::uint32_t>(), L.qmask.as<unsigned long long>(), F, L.mark.as<unsigned long long>());
                hip_check(hipGetLastError(), "launch mark_from_queue");
            }
            hip_check(hipMemsetAsync(g.ws.ctr.ptr, 0, sizeof(Counters), g.stream), "reset counters");
            Counters c;
            F = ordered_compaction_k(E, c);
            pairs = c.frontier_edges;
        }
        out.F = F;
        out.pairs = pairs;
        if (F == 0)
            return out;
        if (g.off64)
            lane_round_kp<std::uint64_t>(E, kind, K, F, pairs, damping, eps, form, st, out);
        else
            lane_round_kp<std::uint32_t>(E, kind, K, F, pairs, damping, eps, form, st, out);
        st.run.rounds += 1;
        st.run.frontier_total += F;
        st.run.edges_total += out.c.all_edges();
        return out;
    }
} // namespace hyt

namespace hyt
{
    // Shared front of the K-lane entry points: checks, split, stats, and a
    // clean mark array after an error (the next run expects it all zero).
    template <class Body>
    void hybrid_k_call(hyt_hybrid_t h, const uint32_t *ids, uint32_t K, bool weights, const hyt_run_options *opts,
                       hyt_hybrid_stats *stats, Body body)
    {
        using namespace hyt;
        require(h != nullptr && ids != nullptr, "null engine or sources");
        require(K >= 1 && K <= 64, "K must be in [1, 64]");
        HybridEngine &E = h->e;
        for (uint32_t k = 0; k < K; ++k)
            require(ids[k] < E.dg.V, "source vertex is out of range");
        hyt_run_options o{};
        if (opts)
            o = *opts;
        require(o.lane_form <= HYT_LANES_ITEMS, "lane_form must be 0, 1 or 2");
        require(o.tolerance == 0.0F, "SSSP tolerance > 0 runs on an HBM-resident graph (hyt_sssp)");
        hyt_hybrid_stats st{};
        st.partitions = E.parts.size();
        st.link_gbs_peak = E.link_gbs;
        st.cpu_pack_gbs = E.pack_gbs;
        st.zero_copy_gbs = E.zc_gbs;
        hip_check(hipSetDevice(E.dg.device), "hipSetDevice");
        const auto t0 = std::chrono::steady_clock::now();
        require(o.task_order == 0, "K-query runs keep partition order (task_order 0)");
        E.begin_run(weights);
        E.set_split(o, kLaneLightMaxDegree, kLaneHeavyChunk);
        try
        {
            body(E, o, st);
        }
        catch (...)
        {
            if (E.lk.mark.ptr)
                (void)hipMemsetAsync(E.lk.mark.ptr, 0, E.lk.mark.bytes, E.dg.stream);
            (void)hipStreamSynchronize(E.dg.stream);
            if (stats)
                *stats = st;
            throw;
        }
        st.device_edge_bytes = E.edge_area_bytes();
        st.device_state_bytes = E.state_bytes();
        st.run.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (stats)
            *stats = st;
    }
} // namespace hyt

extern "C" {

int hyt_hybrid_config_default(hyt_hybrid_config *out)
{
    return hyt::guarded("hyt_hybrid_config_default", [&] {
        hyt::require(out != nullptr, "null output");
        hyt_hybrid_config c{};
        c.device_budget_bytes = 4ULL << 30;
        c.partition_edges = 1ULL << 20;
        c.pack_edges = 0;
        c.filter_combine_k = 4;
        c.force_engine = -1;
        c.host_threads = 0;
        c.calibrate = 1;
        c.cache = 1;
        c.cost = hyt::default_cost_params();
        c.cost.entry_bytes = 8; // column + weight travel together
        *out = c;
    });
}

int hyt_hybrid_create(hyt_host_graph_t host, uint32_t weights_f32, const hyt_hybrid_config *config, hyt_hybrid_t *out)
{
    return hyt::guarded("hyt_hybrid_create", [&] {
        hyt::require(host != nullptr && config != nullptr && out != nullptr, "null pointer");
        *out = nullptr;
        hyt::require(config->force_engine >= -1 && config->force_engine <= 2, "force_engine must be -1, 0, 1 or 2");
        auto h = std::make_unique<hyt_hybrid>();
        hyt::create(h->e, hyt::host_graph_of(host), weights_f32 != 0, *config);
        *out = h.release();
    });
}

int hyt_hybrid_free(hyt_hybrid_t h)
{
    return hyt::guarded("hyt_hybrid_free", [&] { delete h; });
}

int hyt_hybrid_info(hyt_hybrid_t h, uint64_t *partitions, hyt_cost_params *params_in_use, hyt_hybrid_stats *cal)
{
    return hyt::guarded("hyt_hybrid_info", [&] {
        hyt::require(h != nullptr, "null engine");
        if (partitions)
            *partitions = h->e.parts.size();
        if (params_in_use)
        {
            const auto &p = h->e.params;
            params_in_use->alpha = p.alpha;
            params_in_use->beta = p.beta;
            params_in_use->gamma = p.gamma;
            params_in_use->request_bytes = p.request_bytes;
            params_in_use->requests_per_tlp = p.requests_per_tlp;
            params_in_use->rtt = p.rtt;
            params_in_use->cpu_bytes_per_second = p.cpu_bytes_per_second;
            params_in_use->entry_bytes = p.entry_bytes;
            params_in_use->index_bytes = p.index_bytes;
            params_in_use->alignment_bytes = p.alignment_bytes;
        }
        if (cal)
        {
            *cal = hyt_hybrid_stats{};
            cal->partitions = h->e.parts.size();
            cal->link_gbs_peak = h->e.link_gbs;
            cal->cpu_pack_gbs = h->e.pack_gbs;
            cal->zero_copy_gbs = h->e.zc_gbs;
            cal->device_edge_bytes = h->e.edge_area_bytes();
            cal->device_state_bytes = h->e.state_bytes();
        }
    });
}

int hyt_hybrid_sssp(hyt_hybrid_t h, uint32_t source, uint32_t *dist, const hyt_run_options *opts,
                    hyt_hybrid_stats *stats)
{
    return hyt::guarded("hyt_hybrid_sssp", [&] {
        using namespace hyt;
        require(h != nullptr && dist != nullptr, "null engine or dist");
        HybridEngine &E = h->e;
        require(source < E.dg.V, "SSSP source vertex is out of range");
        hyt_run_options o{};
        if (opts)
            o = *opts;
        require(o.tolerance == 0.0F, "SSSP tolerance > 0 runs on an HBM-resident graph (hyt_sssp)");
        hyt_hybrid_stats st{};
        st.partitions = E.parts.size();
        st.link_gbs_peak = E.link_gbs;
        st.cpu_pack_gbs = E.pack_gbs;
        st.zero_copy_gbs = E.zc_gbs;
        st.device_edge_bytes = E.edge_area_bytes();
        st.device_state_bytes = E.state_bytes();
        hip_check(hipSetDevice(E.dg.device), "hipSetDevice");
        const auto t0 = std::chrono::steady_clock::now();
        const int wm = E.dg.weight_mode();
        require(o.task_order <= 1, "task_order must be 0 or 1");
        E.begin_run(true, o.task_order ? kPrioMinValue : kPrioNone);
        E.set_split(o);
        auto run = [&](auto off_tag) {
            using OffT = decltype(off_tag);
            switch (wm)
            {
            case kWU32: hybrid_sssp_typed<OffT, kWU32>(E, source, dist, o, st); break;
            case kWF32: hybrid_sssp_typed<OffT, kWF32>(E, source, dist, o, st); break;
            case kWUnitU32: hybrid_sssp_typed<OffT, kWUnitU32>(E, source, dist, o, st); break;
            default: hybrid_sssp_typed<OffT, kWUnitF32>(E, source, dist, o, st); break;
            }
        };
        try
        {
            if (E.dg.off64)
                run(std::uint64_t{});
            else
                run(std::uint32_t{});
        }
        catch (...)
        {
            (void)hipMemsetAsync(E.dg.ws.touch.ptr, 0, E.dg.V + kPad, E.dg.stream);
            (void)hipStreamSynchronize(E.dg.stream);
            if (stats)
                *stats = st;
            throw;
        }
        st.run.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (stats)
            *stats = st;
    });
}

int hyt_hybrid_pagerank(hyt_hybrid_t h, float damping, float epsilon, float *ranks, const hyt_run_options *opts,
                        hyt_hybrid_stats *stats)
{
    return hyt::guarded("hyt_hybrid_pagerank", [&] {
        using namespace hyt;
        require(h != nullptr && ranks != nullptr, "null engine or ranks");
        require(damping >= 0.0F && damping < 1.0F, "PageRank damping_factor must be in [0, 1)");
        require(epsilon > 0.0F && std::isfinite(epsilon), "PageRank epsilon must be finite and positive");
        HybridEngine &E = h->e;
        hyt_run_options o{};
        if (opts)
            o = *opts;
        hyt_hybrid_stats st{};
        st.partitions = E.parts.size();
        st.link_gbs_peak = E.link_gbs;
        st.cpu_pack_gbs = E.pack_gbs;
        st.zero_copy_gbs = E.zc_gbs;
        st.device_edge_bytes = E.edge_area_bytes();
        st.device_state_bytes = E.state_bytes();
        hip_check(hipSetDevice(E.dg.device), "hipSetDevice");
        const auto t0 = std::chrono::steady_clock::now();
        require(o.task_order <= 1, "task_order must be 0 or 1");
        E.begin_run(false, o.task_order ? kPrioPushedMass : kPrioNone);
        E.set_split(o);
        if (E.dg.off64)
            hybrid_pagerank_typed<std::uint64_t>(E, damping, epsilon, ranks, o, st);
        else
            hybrid_pagerank_typed<std::uint32_t>(E, damping, epsilon, ranks, o, st);
        st.run.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (stats)
            *stats = st;
    });
}

int hyt_hybrid_sssp_k(hyt_hybrid_t h, const uint32_t *sources, uint32_t num_sources, uint32_t *dist,
                      const hyt_run_options *opts, hyt_hybrid_stats *stats)
{
    return hyt::guarded("hyt_hybrid_sssp_k", [&] {
        hyt::require(dist != nullptr, "null dist");
        hyt::hybrid_k_call(h, sources, num_sources, true, opts, stats,
                      [&](hyt::HybridEngine &E, const hyt_run_options &o, hyt_hybrid_stats &st) {
                          hyt::hybrid_sssp_k(E, sources, num_sources, dist, o, st);
                      });
    });
}

int hyt_hybrid_ppr_k(hyt_hybrid_t h, const uint32_t *seeds, uint32_t num_seeds, float damping, float epsilon,
                     float *out, const hyt_run_options *opts, hyt_hybrid_stats *stats)
{
    return hyt::guarded("hyt_hybrid_ppr_k", [&] {
        hyt::require(out != nullptr, "null output");
        hyt::require(damping >= 0.0F && damping < 1.0F, "PageRank damping_factor must be in [0, 1)");
        hyt::require(epsilon > 0.0F && std::isfinite(epsilon), "PageRank epsilon must be finite and positive");
        hyt::hybrid_k_call(h, seeds, num_seeds, false, opts, stats,
                      [&](hyt::HybridEngine &E, const hyt_run_options &o, hyt_hybrid_stats &st) {
                          hyt::hybrid_ppr_k(E, seeds, num_seeds, damping, epsilon, out, o, st);
                      });
    });
}

} // extern "C"
