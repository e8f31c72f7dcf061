// Hybrid transfer engine: SSSP / delta-push PageRank over a graph whose
// edges stay in pinned, device-mapped host memory (HyTGraph's hybrid transfer
// management, executed rather than modelled).
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
#include "host_graph.hpp"
#include "ops.hpp"
#include "transfer_plan.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
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

        // Per-partition activity counts from the sorted frontier: rows, active
        // edges, zero-copy requests, alignment extras (u64 each, 4 per
        // partition). Wavefronts whose rows share one partition add once.
        template <class OffT>
        __global__ __launch_bounds__(kBlock) void part_stats(const std::uint32_t *q, std::uint32_t F, const OffT *off,
                                                             const std::uint32_t *part_vbegin, std::uint32_t P,
                                                             std::uint64_t m, std::uint64_t d1, std::uint64_t align,
                                                             unsigned long long *stats)
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
            std::int64_t pos = -1;         // index in the resident list
        };

        constexpr std::size_t kRingBuffers = 4;

        struct EventPair
        {
            hipEvent_t a = nullptr, b = nullptr;
        };
    } // namespace

    struct HybridEngine
    {
        // host graph (borrowed, pinned + mapped)
        HostGraph *host = nullptr;
        bool weights_f32 = false;
        const std::uint32_t *z_col = nullptr, *z_w = nullptr; // device views of host arrays
        // device graph: offsets + vertex workspace (no edges)
        DeviceGraph dg;
        std::vector<hyt_partition> parts;
        DeviceBuffer part_vbegin, part_stats_d, block_counts, block_offsets, scan_temp, stage_q, stage_val;
        std::vector<unsigned long long> part_stats_h;
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
        std::uint64_t units = 0;
        DeviceBuffer unit_mem;
        std::vector<std::uint32_t> free_units;
        std::vector<Resident> res;                // per partition
        std::vector<std::uint64_t> resident;      // partitions holding units
        std::vector<std::uint64_t> last_active;   // per partition: stamp of the last round it was active
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
            for (void *p : {static_cast<void *>(h_rows), static_cast<void *>(h_col), static_cast<void *>(h_w),
                            static_cast<void *>(h_val), static_cast<void *>(h_index), static_cast<void *>(h_q),
                            static_cast<void *>(h_qd)})
                if (p)
                    (void)hipHostFree(p);
        }

        std::uint64_t state_bytes() const { return dg.device_bytes() + part_vbegin.bytes + part_stats_d.bytes + block_counts.bytes + block_offsets.bytes + scan_temp.bytes + stage_q.bytes + stage_val.bytes; }
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
        void begin_run(bool weights)
        {
            need_w = weights && host->weighted;
            run_edge_bytes = need_w ? 8 : 4;
            run_params = params;
            if (!need_w)
                run_params.entry_bytes = 4;
        }

        void set_split(const hyt_run_options &o)
        {
            light_max = o.light_max_degree ? o.light_max_degree : kHybridLightMaxDegree;
            chunk = o.heavy_chunk ? o.heavy_chunk : kHybridHeavyChunk;
            require(light_max <= (1U << 20), "light_max_degree must be <= 2^20");
            require(chunk >= 64 && chunk <= (1U << 24), "heavy_chunk must be in [64, 2^24]");
            ensure_workspace(dg, light_max, chunk);
        }

        bool is_cached(std::uint64_t p) const { return res[p].col >= 0 && (!need_w || res[p].w >= 0); }

        void drop(std::uint64_t p)
        {
            Resident &r = res[p];
            if (r.col >= 0)
                free_units.push_back(static_cast<std::uint32_t>(r.col));
            if (r.w >= 0)
                free_units.push_back(static_cast<std::uint32_t>(r.w));
            if (r.pos >= 0)
            {
                const std::uint64_t last = resident.back();
                resident[static_cast<std::size_t>(r.pos)] = last;
                res[last].pos = r.pos;
                resident.pop_back();
            }
            r = Resident{};
        }

        // Finds units for partition p (this round's stamp protects the active
        // partitions); false when every resident partition is active now.
        bool admit(std::uint64_t p)
        {
            const std::size_t need = need_w ? 2 : 1;
            if (res[p].col >= 0)
                drop(p); // cached without the weights this run needs
            // a run that reads no weights first takes back resident weight units
            if (!need_w)
                for (std::size_t i = 0; i < resident.size() && free_units.size() < need; ++i)
                {
                    Resident &o = res[resident[i]];
                    if (o.w >= 0)
                    {
                        free_units.push_back(static_cast<std::uint32_t>(o.w));
                        o.w = -1;
                    }
                }
            while (free_units.size() < need)
            {
                std::int64_t victim = -1;
                std::uint64_t oldest = stamp;
                for (std::uint64_t v : resident)
                    if (last_active[v] < oldest)
                    {
                        oldest = last_active[v];
                        victim = static_cast<std::int64_t>(v);
                    }
                if (victim < 0)
                    return false;
                drop(static_cast<std::uint64_t>(victim));
            }
            Resident &r = res[p];
            r.col = free_units.back();
            free_units.pop_back();
            if (need_w)
            {
                r.w = free_units.back();
                free_units.pop_back();
            }
            r.pos = static_cast<std::int64_t>(resident.size());
            resident.push_back(p);
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
            // pinned H2D peak: copy up to 256 MiB of the host column array
            const std::uint64_t bytes = std::min<std::uint64_t>(256ULL << 20, E.host->E * 4);
            DeviceBuffer &dst = E.unit_mem.bytes >= bytes ? E.unit_mem : E.ring_mem;
            if (bytes >= (4ULL << 20) && dst.bytes >= bytes)
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
            require(host->memory == HYT_HOST_PINNED, "the hybrid engine needs a pinned host graph");
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
            g.off64 = host->E >= 0xFFFFFFFFULL;
            g.weighted = host->weighted;
            g.weights_f32 = f32;
            hip_check(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipStreamCreateWithFlags(&E.copy, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreate(&g.ev0), "hipEventCreate");
            hip_check(hipEventCreate(&g.ev1), "hipEventCreate");
            hip_check(hipEventCreateWithFlags(&E.ev_pack, hipEventDisableTiming), "hipEventCreate");
            g.host_ctr = std::make_unique<PinnedBuffer>(sizeof(Counters));
            void *dp = nullptr;
            hip_check(hipHostGetDevicePointer(&dp, host->col, 0), "hipHostGetDevicePointer(columns)");
            E.z_col = static_cast<const std::uint32_t *>(dp);
            if (host->weighted)
            {
                hip_check(hipHostGetDevicePointer(&dp, host->w, 0), "hipHostGetDevicePointer(weights)");
                E.z_w = static_cast<const std::uint32_t *>(dp);
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
            const std::uint64_t k = cfg.filter_combine_k ? cfg.filter_combine_k : 4;
            const std::uint64_t ring_k = std::max<std::uint64_t>(1, std::min<std::uint64_t>(k, avail / 2 / ring_unit));
            E.ring_edges = ring_k * E.unit_edges;
            E.ring_mem.alloc(ring_k * ring_unit, "hipMalloc(streaming ring)");
            for (hipEvent_t &ev : E.ring_used)
                hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
            E.units = (avail - ring_k * ring_unit) / (E.unit_edges * 4);
            if (E.units > 0)
                E.unit_mem.alloc(E.units * E.unit_edges * 4, "hipMalloc(residency units)");
            E.free_units.clear();
            for (std::uint64_t u = E.units; u-- > 0;)
                E.free_units.push_back(static_cast<std::uint32_t>(u));
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
            part_stats<OffT><<<blocks_for(F), kBlock, 0, s>>>(
                g.ws.q.as<std::uint32_t>(), F, g.off.as<OffT>(), E.part_vbegin.as<std::uint32_t>(),
                static_cast<std::uint32_t>(P), E.run_params.request_bytes, E.run_params.entry_bytes,
                E.run_params.alignment_bytes,
                E.part_stats_d.as<unsigned long long>());
            hip_check(hipGetLastError(), "launch part_stats");
            hip_check(hipMemcpyAsync(E.part_stats_h.data(), E.part_stats_d.ptr, P * 32, hipMemcpyDeviceToHost, s),
                      "read partition stats");
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
                    engine = hytm::evaluate(m, E.run_params).engine;
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

        // Op factory: Make(col_base, w_base, qd_base) -> Op
        template <class OffT, class Make>
        void execute_round(HybridEngine &E, const Plan &plan, Make make, hyt_hybrid_stats &st)
        {
            DeviceGraph &g = E.dg;
            hipStream_t s = g.stream;
            const std::uint32_t *q = g.ws.q.as<std::uint32_t>();
            const std::uint32_t *qd = g.ws.qd.as<std::uint32_t>();
            const OffT *off = g.off.as<OffT>();
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
                            hip_check(hipMemcpyAsync(E.h_qd + at, qd + plan.qlo[p0], n * 4, hipMemcpyDeviceToHost, s),
                                      "read frontier values");
                            at += n;
                        });
                hip_check(hipStreamSynchronize(s), "frontier readback");
            }

            // 2. resident partitions
            for (std::uint64_t p : plan.cached)
            {
                const hyt_partition &pt = E.parts[p];
                const Resident &r = E.res[p];
                GraphRows<OffT> rs{off, q + plan.qlo[p], pt.edge_begin, pt.edge_end};
                expand(E, rs, make(E.unit_ptr(r.col), weighted ? E.unit_ptr(r.w) : nullptr, qd + plan.qlo[p]),
                       static_cast<std::uint32_t>(plan.qn[p]));
                st.parts_cached += 1;
            }

            // 3. filter: each run of consecutive partitions streams through the
            // ring in windows of ring_edges (one copy per array per window);
            // whole partitions of a window are then promoted into the cache
            for (const Task &t : plan.tasks)
            {
                if (t.engine != HYT_ENGINE_FILTER)
                    continue;
                for_runs(t.members, [&](std::uint64_t p0, std::uint64_t p1) {
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
                        E.copy_h2d(E.ring_col(r), E.host->col + a, (b - a) * 4);
                        if (weighted)
                            E.copy_h2d(E.ring_w(r), E.host->w + a, (b - a) * 4);
                        st.bytes_filter += (b - a) * E.run_edge_bytes;
                        hip_check(hipStreamWaitEvent(s, E.copy_events[E.copy_events_used - 1].b, 0), "wait copy");
                        const std::uint64_t qa = plan.qlo[pa];
                        GraphRows<OffT> rs{off, q + qa, a, b};
                        expand(E, rs, make(E.ring_col(r), E.ring_w(r), qd + qa), rows_of(pa, pb));
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
                });
            }

            // 4. zero-copy: one window per run of consecutive partitions
            for (const Task &t : plan.tasks)
            {
                if (t.engine != HYT_ENGINE_ZERO_COPY)
                    continue;
                for_runs(t.members, [&](std::uint64_t p0, std::uint64_t p1) {
                    const std::uint64_t e0 = E.parts[p0].edge_begin, e1 = E.parts[p1].edge_end;
                    for (std::uint64_t p = p0; p <= p1; ++p) // estimated link bytes: m per modelled request
                        st.bytes_zero_copy += E.part_stats_h[4 * p + 2] * E.run_params.request_bytes;
                    GraphRows<OffT> rs{off, q + plan.qlo[p0], e0, e1};
                    expand(E, rs, make(E.z_col + e0, weighted ? E.z_w + e0 : nullptr, qd + plan.qlo[p0]),
                           rows_of(p0, p1));
                    st.parts_zero_copy += p1 - p0 + 1;
                });
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
                    const std::uint32_t u = E.h_q[r0];
                    const std::uint64_t d = E.host->off[u + 1] - E.host->off[u];
                    if (d > E.pack_cap)
                    {
                        E.h_rows[0] = u;
                        E.h_val[0] = E.h_qd[r0];
                        E.copy_h2d(E.pk_rows.ptr, E.h_rows, 4);
                        E.copy_h2d(E.pk_val.ptr, E.h_val, 4);
                        hip_check(hipStreamWaitEvent(s, E.copy_events[E.copy_events_used - 1].b, 0), "wait row copy");
                        GraphRows<OffT> rs{off, E.pk_rows.as<std::uint32_t>(), 0, ~0ULL};
                        expand(E, rs, make(E.z_col, weighted ? E.z_w : nullptr, E.pk_val.as<std::uint32_t>()), 1U);
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
                PackedRows rs{E.pk_rows.as<std::uint32_t>(), E.pk_index.as<std::uint64_t>()};
                expand(E, rs, make(E.pk_col.as<std::uint32_t>(), weighted ? E.pk_w.as<std::uint32_t>() : nullptr,
                                   E.pk_val.as<std::uint32_t>()),
                       static_cast<std::uint32_t>(nr));
                hip_check(hipEventRecord(E.ev_pack, s), "hipEventRecord");
                r0 = r1;
            }
        }

        void finish_round_stats(HybridEngine &E, hyt_hybrid_stats &st)
        {
            hip_check(hipStreamSynchronize(E.copy), "copy stream");
            st.copy_ms += E.drain_copy_time();
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
                execute_round<OffT>(E, plan, make, st);
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
                st.run.edges_total += cc.edges;
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
            launch_fill<float>(ws.dist.as<float>(), g.V, 1.0F - damping, s);
            launch_fill<float>(ws.rank.as<float>(), g.V, 0.0F, s);
            hip_check(hipMemsetAsync(d_ctr, 0, sizeof(Counters), s), "reset counters");
            PrSel<OffT> sel{ws.dist.as<float>(), ws.rank.as<float>(), ws.share.as<float>(), g.off.as<OffT>(), damping,
                            eps};
            Counters c;
            std::uint32_t F = ordered_compaction<OffT>(E, sel, c);
            auto make = [&](const std::uint32_t *col, const std::uint32_t *, const std::uint32_t *qd) {
                PrOp op;
                op.col = col;
                op.res = ws.dist.as<float>();
                op.qd = qd;
                op.damping = damping;
                return op;
            };
            const std::uint64_t max_rounds = o.max_rounds ? o.max_rounds : 100000;
            Plan plan;
            st.run.converged = F == 0 ? 1 : 0;
            for (std::uint64_t round = 0; round < max_rounds && F > 0; ++round)
            {
                plan_round<OffT>(E, F, plan, st);
                hip_check(hipMemsetAsync(d_ctr, 0, kRoundCounterBytes, s), "reset counters");
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev0, s), "hipEventRecord");
                execute_round<OffT>(E, plan, make, st);
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
                st.run.frontier_total += F;
                st.run.edges_total += cc.edges;
                st.run.updates_total += Fn;
                st.run.rounds = round + 1;
                F = Fn;
                if (F == 0)
                    st.run.converged = 1;
            }
            finalize_ranks(g, out, o.out_on_device != 0, st.run);
        }
    } // namespace
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
        E.begin_run(true);
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
        E.begin_run(false);
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

} // extern "C"
