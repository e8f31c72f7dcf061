// Device-resident CSR and the per-graph traversal workspace.
//
// HBM layout (structure of arrays, 256-byte aligned by hipMalloc):
//   off  u32[V+1] when E < 2^32, else u64[V+1]
//   col  u32[E]   destination ids, row-sorted
//   w    u32[E]   integer weights or float bit patterns (absent if unweighted)
// Workspace (single-source traversals; every [V] array has 16 spare slots):
//   dist  u32[V]  distances (PageRank: staging of the float output),
//   res, rank u64[V] PageRank residual and pushed mass, fixed-point words
//                 (mass.hpp; allocated by the first PageRank call),
//   share f32[V]  PageRank per-edge share of the last compaction's rows,
//   prev  u32[V]  distance at the end of the round a vertex last improved in,
//   touch u8[V]   set (plain store) by expand when a vertex may have improved,
//   q     u32[V]  frontier rows, qd u32[V] their snapshot values,
//   heavy HeavyChunk[cap] chunks of high-degree rows, counters.
#pragma once

#include "hip_util.hpp"
#include "splits.hpp"

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <memory>

namespace hyt
{
    enum WeightMode : int
    {
        kWU32 = 0,     // integer weights, u32 sums (overflow detected)
        kWF32 = 1,     // float weights, float sums (bit-order min on >= +0)
        kWUnitU32 = 2, // unweighted, u32 sums of 1
        kWUnitF32 = 3  // unweighted, float sums of 1.0f
    };

    // One chunk of a high-degree row, expanded by one wavefront.
    struct HeavyChunk
    {
        std::uint32_t value;  // snapshot distance / residual share (bits)
        std::uint32_t count;  // edges in this chunk
        std::uint64_t begin;  // first edge index
        std::uint32_t row;    // the row's vertex
        std::uint32_t record; // 1: the chunks keep the row's expansion record
    };

    // Edge counts of the expand kernels go to kEdgeSlots counters on separate
    // 256-byte lines (workgroup b adds to slot b % kEdgeSlots): one counter
    // taking every workgroup's atomic serialises them.
    constexpr int kEdgeSlots = 8;
    struct alignas(256) EdgeSlot
    {
        unsigned long long v; // rows walked whole or in windows
        unsigned long long h; // chunks of rows above light_max (expand_heavy)
        std::uint32_t kinv;   // max of ~dist over the rows the compaction kept (waiting hubs)
    };

    // Per-round counters. The bytes before `overflow` (kRoundCounterBytes)
    // are zeroed at the start of each round; the two flags are sticky for the
    // whole call.
    struct alignas(256) Counters
    {
        std::uint32_t next_count;          // rows written by the compaction pass
        std::uint32_t heavy_count;         // heavy chunks queued this round
        unsigned long long edges;          // edges expanded this round (besides the slots)
        std::uint32_t pending;             // improved rows held back (at/above the bucket bound)
        std::uint32_t pending_inv_max;     // max of ~dist over held-back rows (= ~min dist)
        EdgeSlot slot[kEdgeSlots];         // edges expanded this round, spread
        // Written by compactions beside next_count, which takes every
        // workgroup's atomic, so on a line of their own: the out-degree sum
        // of the compacted rows, and the rows above light_max the SSSP
        // compaction listed apart.
        alignas(256) unsigned long long frontier_edges;
        std::uint32_t next_hubs;
        std::uint32_t overflow;            // a u32 sum overflowed
        std::uint32_t heavy_overflow;      // heavy chunk queue overflowed

        __host__ __device__ unsigned long long all_edges() const
        {
            unsigned long long e = edges;
            for (int i = 0; i < kEdgeSlots; ++i)
                e += slot[i].v + slot[i].h;
            return e;
        }
        // smallest distance the compaction kept (0xFFFFFFFF: none recorded)
        __host__ __device__ std::uint32_t kept_min() const
        {
            std::uint32_t m = 0;
            for (int i = 0; i < kEdgeSlots; ++i)
                m = slot[i].kinv > m ? slot[i].kinv : m;
            return ~m;
        }
        // the part of all_edges() that the chunk kernel walked
        __host__ __device__ unsigned long long heavy_edges() const
        {
            unsigned long long e = 0;
            for (int i = 0; i < kEdgeSlots; ++i)
                e += slot[i].h;
            return e;
        }
    };
    constexpr std::size_t kRoundCounterBytes = offsetof(Counters, overflow);

    // Adds n expanded edges for workgroup blockIdx.x.
    __device__ __forceinline__ void add_edges(Counters *ctr, unsigned long long n)
    {
        atomicAdd(&ctr->slot[blockIdx.x % kEdgeSlots].v, n);
    }
    // Adds n edges walked by the chunk kernel (counted apart, same slots).
    __device__ __forceinline__ void add_heavy_edges(Counters *ctr, unsigned long long n)
    {
        atomicAdd(&ctr->slot[blockIdx.x % kEdgeSlots].h, n);
    }

    constexpr int kMaxGraphRounds = 8; // rounds per captured round graph, at most

    // Device-side control of device-driven rounds (rounds replayed from a
    // captured HIP graph without a host round trip): the queue size of the
    // current round, read by the kernels, and the run's totals.
    struct RoundCtl
    {
        std::uint32_t F;         // rows of the current round's queue (0: no work)
        std::uint32_t done;      // a round found the queue empty, or the round limit
        std::uint32_t truncated; // stopped by the round limit with work left
        std::uint32_t heavy_bound; // rows above light_max with a larger distance wait (this round)
        std::uint32_t H;           // rows of the current round's hub list (rows above light_max)
        std::uint64_t rounds, frontier_total, edges_total, updates_total;
        std::uint64_t heavy_edges_total; // of edges_total: walked by the chunk kernel
        // timed runs: device wall clock before and after the expand kernels
        // of each round of one replay (round r: t[2r], t[2r+1])
        std::uint64_t t[2 * kMaxGraphRounds];
    };

    struct Workspace
    {
        DeviceBuffer dist, prev, touch, q, qd, heavy, ctr, partial, rc;
        DeviceBuffer res, rank; // PageRank mass words u64[V] (ensure_mass_state)
        DeviceBuffer active; // frontier bitmap (bit v = row v selected by the last compaction)
        DeviceBuffer share;  // PageRank pull: d * r_u / outdeg(u) of the frontier rows, by vertex
        // device-driven ASYNC SSSP: bit j of hub_bits[t] = row 16t + j has
        // more than hub_light_max edges; the compaction lists the selected
        // ones in hub_list (hub_rows entries at most)
        DeviceBuffer hub_bits, hub_list;
        std::uint32_t hub_light_max = 0;
        std::uint64_t hub_rows = 0;
        std::uint64_t heavy_cap = 0;
        bool ready = false;
    };

    struct DeviceGraph
    {
        std::uint64_t V = 0, E = 0;
        std::uint64_t max_degree = 0;
        bool off64 = false;
        bool weighted = false;
        bool weights_f32 = false;
        int device = 0;
        DeviceBuffer off, col, w;
        // incoming CSR for pull rounds (built on first use)
        DeviceBuffer in_off, in_col, in_w, in_heavy;
        std::uint64_t in_heavy_count = 0;
        std::uint32_t in_heavy_light_max = 0, in_heavy_chunk = 0;
        hipStream_t stream = nullptr;
        Workspace ws;
        std::shared_ptr<void> kws; // workspace of the batched [V][K] calls (owned, type-erased)
        std::shared_ptr<void> round_graph; // captured single-source rounds (owned, type-erased)
        std::shared_ptr<void> pws;         // persistent-kernel queues and stamps (owned, type-erased)
        std::unique_ptr<PinnedBuffer> host_ctr;
        hipEvent_t ev0 = nullptr, ev1 = nullptr;

        ~DeviceGraph();
        std::uint64_t device_bytes() const
        {
            return off.bytes + col.bytes + w.bytes + in_off.bytes + in_col.bytes + in_w.bytes + in_heavy.bytes +
                   ws.dist.bytes + ws.res.bytes + ws.rank.bytes + ws.prev.bytes + ws.touch.bytes + ws.q.bytes + ws.qd.bytes +
                   ws.heavy.bytes + ws.ctr.bytes + ws.partial.bytes + ws.active.bytes + ws.share.bytes + ws.rc.bytes +
                   ws.hub_bits.bytes + ws.hub_list.bytes;
        }
        int weight_mode() const
        {
            if (!weighted)
                return weights_f32 ? kWUnitF32 : kWUnitU32;
            return weights_f32 ? kWF32 : kWU32;
        }
    };

    // The default work split of single-source SSSP depends on the graph's
    // size: default_sssp_split (splits.hpp).
    // PageRank (push and pull rounds), eps 1e-7 on one MI355X: with the pull
    // kernels counting no edges, rows above 32 in-edges go to the chunk
    // kernel, which gathers at 72% of the HBM peak against 22% for the
    // 64-rows-per-wavefront walk: 32 / 256 took 0.316 s on RMAT 24 against
    // 0.553 s (512 / 2048), 0.063 / 0.148 s on RMAT 22, 1.86 / 2.39 s on
    // RMAT 26, and 0.507 / 0.506 s on ER 24 (16 / 256: within 2% on RMAT,
    // 3% slower on ER). The K-lane kernels keep 256 / 1024: a K-lane
    // light row carries K queries per edge, so long light rows would leave
    // one wavefront with most of a round's work.
    constexpr std::uint32_t kPagerankLightMaxDegree = 32;
    constexpr std::uint32_t kPagerankHeavyChunk = 256;
    constexpr std::uint32_t kLaneLightMaxDegree = 256;
    constexpr std::uint32_t kLaneHeavyChunk = 1024;
    // Batched calls (hyt_sssp / hyt_ppr with K >= 2, graph in HBM): RMAT 24,
    // one MI355X, against the current kernels: 32 / 256 ran K = 8 SSSP at
    // 308 queries/s and K = 64 at 402 against 290 and 385 at 64 / 512
    // (16 / 128 and 8 / 128 within 1%, 128 / 1024 and 256 / 1024 slower:
    // 264 and 256 at K = 8); PPR within noise of every split. Lane sessions
    // keep 256 / 1024: their chunk queue is fixed per-session state, which
    // smaller splits enlarge.
    constexpr std::uint32_t kBatchedLightMaxDegree = 32;
    constexpr std::uint32_t kBatchedHeavyChunk = 256;
    // Hybrid engine expansions read windows of partitions, many of them over
    // the host link (zero-copy): short light rows and small chunks keep many
    // wavefronts with reads in flight. RMAT scale 27, 4 GiB budget, one SSSP
    // query: 64/256 1.00 s, 128/512 1.01 s, 256/1024 1.06 s, 512/2048 1.42 s,
    // 2048/4096 2.1-2.6 s.
    constexpr std::uint32_t kHybridLightMaxDegree = 128;
    constexpr std::uint32_t kHybridHeavyChunk = 512;
    // Direction choice: pull when the frontier's out-edges reach this % of E.
    // Measured on RMAT scale 24 (MI355X): PageRank pull rounds (one gather per
    // in-edge, no atomics) beat float-atomic pushes from ~25% of E; SSSP push
    // (one filtered atomic per edge) beats pull (bitmap probe + gather per
    // in-edge) even at 50%, so SSSP pulls only when asked.
    constexpr std::uint64_t kDefaultPullPercentPagerank = 25;
    // ASYNC single-source SSSP: rows above light_max wait while their
    // distance exceeds the queue's smallest by more than this (weight units).
    // A row within w_min of the smallest queued distance is final, so the
    // hubs - most of RMAT's edges - go out once instead of ~2 times. One
    // MI355X, 8 sources, u32 weights [1,255]: RMAT 22 / 24 / 26 34.2 / 114.6
    // / 486.9 ms -> 30.1 / 102.4 / 415.4 ms at 2 (relaxations per query 153
    // / 505 / 1914 M -> 108 / 345 / 1291 M, rounds 13 -> 15-19); 1 and 4
    // within 3% of 2; ER 24 (no rows above light_max) unchanged.
    constexpr std::uint32_t kDefaultHeavyDelta = 2;
    inline std::uint32_t heavy_delta_of(std::uint32_t option)
    {
        return option == 0 ? kDefaultHeavyDelta : (option == 0xFFFFFFFFU ? 0U : option);
    }
    constexpr std::uint64_t kDefaultPullPercentSssp = 101;
    // Distance bound of the waiting rows this round: the smallest distance in
    // the round's queue plus heavy_delta (float weights: added as a float);
    // 0xFFFFFFFF (no waiting) when heavy_delta is 0 or no distance was
    // recorded.
    __host__ __device__ inline std::uint32_t heavy_bound_of(std::uint32_t queue_min, std::uint32_t heavy_delta,
                                                           bool float_dist)
    {
        if (heavy_delta == 0 || queue_min == 0xFFFFFFFFU)
            return 0xFFFFFFFFU;
        if (float_dist)
        {
            float f;
            std::memcpy(&f, &queue_min, 4);
            f += static_cast<float>(heavy_delta);
            std::uint32_t b;
            std::memcpy(&b, &f, 4);
            return b;
        }
        const std::uint64_t b = static_cast<std::uint64_t>(queue_min) + heavy_delta;
        return b >= 0xFFFFFFFFULL ? 0xFFFFFFFFU : static_cast<std::uint32_t>(b);
    }
    // ASYNC single-source SSSP: light rows of more than this many edges wait
    // under the same bound as the rows above light_max (option 0: the default,
    // light_max - only chunked rows wait).
    inline std::uint32_t wait_degree_of(std::uint32_t option, std::uint32_t light_max)
    {
        return option == 0 ? light_max : option;
    }

    // Allocates the single-source workspace if needed (sized for the graph).
    void ensure_workspace(DeviceGraph &g, std::uint32_t light_max, std::uint32_t chunk);

    // Spare slots after every [V] workspace array (vector loads past the end).
    constexpr std::uint64_t kPad = 16;

    // Allocates the PageRank residual and rank words (u64[V + kPad] each).
    void ensure_mass_state(DeviceGraph &g);

    // x = rank + residual (mass words), normalised to sum 1, into out (host
    // or device); st.residual = unpushed mass / total.
    void finalize_ranks(DeviceGraph &g, float *out, bool out_on_device, hyt_run_stats &st);

    // Builds the incoming CSR and the static chunk list of its heavy rows.
    void ensure_transpose(DeviceGraph &g, std::uint32_t light_max, std::uint32_t chunk);
    // g.in_off / g.in_col only (E < 2^31; sources of each row ascending).
    void build_incoming_csr(DeviceGraph &g);

    // Device arrays of one single-source SSSP run (every [V] array has kPad
    // spare slots): the graph's workspace, or caller-provided memory. touch
    // and active are needed only by calls that may pull (null otherwise).
    struct SsspBuffers
    {
        std::uint32_t *dist, *prev, *q, *qd;
        std::uint8_t *touch;
        std::uint16_t *active;
        HeavyChunk *heavy;
        std::uint64_t heavy_cap;
        Counters *ctr;
        // non-null: push-only rounds run device-driven from a captured graph
        RoundCtl *rc = nullptr;
        // non-null (device-driven ASYNC rounds): rows above light_max by bit
        // (hub_bits, for this light_max) and the list the compaction writes
        // them to, so the chunk-queuing pass reads those rows alone
        const std::uint16_t *hub_bits = nullptr;
        std::uint32_t *hub_list = nullptr;
        std::uint64_t hub_rows = 0;
    };
    // The single-source work split (options, else the size-dependent default).
    void sssp_split(const DeviceGraph &g, const hyt_run_options &o, std::uint32_t &light_max, std::uint32_t &chunk);
    // The SSSP rounds from src over the given arrays (distances end in b.dist).
    void sssp_rounds(DeviceGraph &g, std::uint32_t src, const SsspBuffers &b, const hyt_run_options &o,
                     std::uint32_t light_max, std::uint32_t chunk, hyt_run_stats &st);
    // Every round of a single-source push run in one persistent cooperative
    // launch (data-driven queues; distances end in ws.dist). Needs the
    // workspace (ensure_workspace at the split).
    void sssp_persistent_rounds(DeviceGraph &g, std::uint32_t src, const hyt_run_options &o, std::uint32_t light_max,
                                std::uint32_t chunk, hyt_run_stats &st);
    // Round-by-round runs (steppers): `rounds` more rounds of a single-source
    // SSSP run from src (resume = continue the run the workspace holds), or
    // of a PageRank / PPR run (seed UINT32_MAX: global); the stats are this
    // call's. pagerank_read writes the current normalised ranks.
    void sssp_steps(DeviceGraph &g, std::uint32_t src, const hyt_run_options &o, bool resume, std::uint64_t rounds,
                    hyt_run_stats &st);
    void pagerank_steps(DeviceGraph &g, std::uint32_t seed, float damping, float eps, const hyt_run_options &o,
                        bool resume, std::uint64_t rounds, hyt_run_stats &st);

    // Heavy-chunk capacity of a single-source run at a split.
    inline std::uint64_t heavy_capacity(std::uint64_t E, std::uint32_t light_max, std::uint32_t chunk)
    {
        return E / chunk + E / (static_cast<std::uint64_t>(light_max) + 1) + 64;
    }
} // namespace hyt

struct hyt_graph
{
    hyt::DeviceGraph g;
};
