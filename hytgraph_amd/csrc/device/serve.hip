// Continuous batching: K-lane query sessions and the request-serving loop.
//
// A lane session holds K query slots ("lanes") of one kind (SSSP or PPR) over
// one HBM-resident graph, with its own [V][K] state, frontier, counters and
// stream. Rounds run for all occupied lanes at once (the K-lane kernels of
// lanes.hpp: each edge is read once for every lane whose row is active). A
// query is admitted into a free lane between rounds: its column of the state
// is reset, its source row is appended to the current queue with its lane
// bit, and it starts in the next round next to queries that are already
// running. A lane finishes when the compaction pass finds no frontier row
// carrying its bit; its result stays in the state until the lane is reused.
//
// Sessions of different graphs and kinds run on their own streams, so one
// step launches a round on every busy session before waiting for any of
// them: their rounds overlap on the device.
//
// The serving loop replays a request trace (arrival times, graph, kind,
// source) against the wall clock: arrivals queue per session, lanes per
// session are sized from the device budget (admission by memory: each lane
// goes to the session with the most work per lane that still fits, work being
// its requests times its graph's edges), and the
// loop polls every session: when its round has completed, finished lanes are
// read and freed, queued queries are admitted into free lanes, and the next
// round is launched - each session advances at its own pace, so a small
// graph's queries are not held back by a large graph's rounds. Each request's
// admit and completion times are recorded.
#include "hybrid_lanes.hpp"
#include "lanes.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <future>
#include <thread>
#include <type_traits>
#include <vector>

namespace hyt
{
    namespace
    {
        using namespace lanes;

        // a[v*K + k] = value for every v.
        template <class T>
        __global__ void lane_fill(T *a, std::uint64_t V, std::uint32_t K, std::uint32_t k, T value)
        {
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V;
                 v += stride)
                a[v * K + k] = value;
        }

        // Seeds lane k at source s and marks its row (sessions over the hybrid
        // engine: the next round folds the mark into its sorted queue).
        template <class T>
        __global__ void lane_seed_mark(T *val, std::uint32_t K, std::uint32_t k, std::uint32_t s, T seed,
                                       unsigned long long *mark)
        {
            if (threadIdx.x == 0 && blockIdx.x == 0)
            {
                val[static_cast<std::uint64_t>(s) * K + k] = seed;
                atomicOr(mark + s, 1ULL << k);
            }
        }

        // Seeds lane k at source s and appends (s, bit k) to the queue at F.
        template <class T>
        __global__ void lane_seed(T *val, std::uint32_t K, std::uint32_t k, std::uint32_t s, T seed,
                                  std::uint32_t *q, unsigned long long *qmask, std::uint32_t F)
        {
            if (threadIdx.x == 0 && blockIdx.x == 0)
            {
                val[static_cast<std::uint64_t>(s) * K + k] = seed;
                q[F] = s;
                qmask[F] = 1ULL << k;
            }
        }

        // out[v] = val[v][k] (SSSP distances).
        __global__ void lane_gather(const std::uint32_t *val, std::uint64_t V, std::uint32_t K, std::uint32_t k,
                                    std::uint32_t *out)
        {
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V;
                 v += stride)
                out[v] = val[v * K + k];
        }

        // PPR lane k: exact per-block sums of the mass words x = rank + res
        // (no column passes 2^62).
        __global__ __launch_bounds__(kBlock) void lane_mass_sum(const unsigned long long *res,
                                                                const unsigned long long *rank, std::uint64_t V,
                                                                std::uint32_t K, std::uint32_t k,
                                                                unsigned long long *partial)
        {
            __shared__ unsigned long long s[kBlock];
            unsigned long long acc = 0;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V;
                 v += stride)
            {
                const std::uint64_t idx = v * K + k;
                acc += rank[idx] + res[idx];
            }
            s[threadIdx.x] = acc;
            __syncthreads();
            for (int d = kBlock / 2; d > 0; d >>= 1)
            {
                if (static_cast<int>(threadIdx.x) < d)
                    s[threadIdx.x] += s[threadIdx.x + d];
                __syncthreads();
            }
            if (threadIdx.x == 0)
                partial[blockIdx.x] = s[0];
        }

        __global__ void lane_total(const unsigned long long *partial, unsigned n, unsigned long long *total)
        {
            if (threadIdx.x == 0 && blockIdx.x == 0)
            {
                unsigned long long t = 0;
                for (unsigned b = 0; b < n; ++b)
                    t += partial[b];
                total[0] = t;
            }
        }

        // out[v] = (rank + res)[v][k] / total as float bits.
        __global__ void lane_mass_out(const unsigned long long *res, const unsigned long long *rank,
                                      std::uint64_t V, std::uint32_t K, std::uint32_t k,
                                      const unsigned long long *total, std::uint32_t *out)
        {
            const double t = static_cast<double>(total[0]);
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V;
                 v += stride)
            {
                const std::uint64_t idx = v * K + k;
                const double x = static_cast<double>(rank[idx] + res[idx]);
                out[v] = __float_as_uint(static_cast<float>(t > 0.0 ? x / t : 0.0));
            }
        }

        // Entries != skip and the u64 sum of their bit patterns (result checksum).
        __global__ __launch_bounds__(kBlock) void lane_checksum(const std::uint32_t *x, std::uint64_t V,
                                                                std::uint32_t skip, unsigned long long *acc)
        {
            unsigned long long n = 0, sum = 0;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V;
                 v += stride)
                if (x[v] != skip)
                {
                    n += 1;
                    sum += x[v];
                }
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1)
            {
                n += __shfl_xor(n, d, kWave);
                sum += __shfl_xor(sum, d, kWave);
            }
            if (lane_id() == 0 && n)
            {
                atomicAdd(&acc[0], n);
                atomicAdd(&acc[1], sum);
            }
        }

        constexpr unsigned kGatherBlocks = 1024;

        unsigned grid_for(std::uint64_t n)
        {
            const std::uint64_t b = (n + kBlock - 1) / kBlock;
            return static_cast<unsigned>(b == 0 ? 1 : (b < kGatherBlocks ? b : kGatherBlocks));
        }
    } // namespace

    struct LaneSession
    {
        DeviceGraph *g = nullptr; // borrowed
        // A session over the hybrid engine (graph edges in host memory) keeps
        // its lanes in the engine's state, runs the engine's rounds on the
        // engine's stream (one step blocks the caller for a round) and admits
        // through the marks; otherwise the session owns its state.
        HybridEngine *hy = nullptr;
        hyt_hybrid_stats hst{};
        std::future<HybridLaneRound> hy_round; // the hybrid round in flight (worker thread)
        // [V][K] lane state (own or the engine's): u32 distances, or u64
        // PPR residual words with rank_p the u64 pushed-mass words (mass.hpp)
        void *val_p = nullptr;
        unsigned long long *rank_p = nullptr;
        unsigned long long *mark_p = nullptr;
        bool own_stream = true;
        int kind = HYT_QUERY_SSSP;
        std::uint32_t K = 0, KP = 0;
        float damping = 0.85F, eps = 1e-6F;
        int wm = kWU32;
        std::uint32_t inf_bits = 0xFFFFFFFFU;  // SSSP: distance of an unreached vertex
        unsigned long long seed_word = 0;      // PPR: initial residual word of the source
        unsigned long long eps_word = 0;       // PPR: threshold word
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        DeviceBuffer val, rank, snap, mark, qmask, q, heavy, ctr, lanes, colbuf, partial, acc;
        std::uint64_t heavy_cap = 0;
        std::unique_ptr<PinnedBuffer> host; // Counters | lanes u64 | checksum u64[2]
        std::uint32_t F = 0;
        std::uint32_t appended = 0; // rows appended by admissions since the last compaction
        std::uint64_t running = 0; // lanes holding a query that has not finished
        std::uint64_t occupied = 0; // lanes not free (running, or finished and unread)
        bool launched = false;
        std::uint32_t lane_rounds[64] = {};
        std::uint64_t rounds = 0, edges = 0;
        std::uint64_t frontier_rows = 0, lane_rows = 0; // lane occupancy of the rounds run
        std::uint64_t pairs = 0;                         // (row, lane) pairs of the current queue
        std::uint64_t item_rounds = 0;                   // rounds run in the item form
        std::uint32_t form = HYT_LANES_AUTO;

        ~LaneSession()
        {
            if (hy_round.valid())
                hy_round.wait();
            if (stream)
            {
                (void)hipStreamSynchronize(stream);
                if (own_stream)
                    (void)hipStreamDestroy(stream);
            }
            if (done)
                (void)hipEventDestroy(done);
        }

        // A queue holds each vertex once after a compaction, plus up to 64
        // rows appended by admissions (a vertex may then appear twice).
        static constexpr std::uint64_t kAppend = 64;

        // Heavy-chunk queue: the rows above light_max plus, per lane, the
        // item-form chunks a lane can add (a full queue walks hub rows in the
        // light kernel, one wavefront per hub).
        static std::uint64_t heavy_chunks(std::uint64_t V, std::uint64_t E)
        {
            return lanes::lane_chunk_capacity(V, E, kLaneLightMaxDegree, kLaneHeavyChunk);
        }
        static std::uint64_t heavy_chunks_per_lane(std::uint64_t E)
        {
            return lanes::lane_item_chunks_per_lane(E, kLaneHeavyChunk);
        }
        static std::uint64_t fixed_bytes(std::uint64_t V, std::uint64_t E, std::uint64_t /*max_degree*/)
        {
            const std::uint64_t Vp = V + kPad;
            return Vp * 8 + (Vp + kAppend) * 8 + (Vp + kAppend) * 4 + heavy_chunks(V, E) * sizeof(HeavyChunkK) +
                   Vp * 4 + (kGatherBlocks + 1) * sizeof(double) + sizeof(Counters) + 8 + 16;
        }
        // SSSP: u32 val [V][K]; PPR: u64 residual and rank [V][K] and
        // snapshot [V+kAppend][K] words; plus the lane's share of the chunk
        // queue
        static std::uint64_t lane_bytes(std::uint64_t V, std::uint64_t E, int kind)
        {
            const std::uint64_t Vp = V + kPad;
            return (kind == HYT_QUERY_PPR ? Vp * 8 * 2 + (Vp + kAppend) * 8 : Vp * 4) +
                   heavy_chunks_per_lane(E) * sizeof(HeavyChunkK);
        }
    };

    namespace
    {
        // Fill and seed values of a session's lanes (S.kind, S.wm, S.damping
        // and S.eps set).
        void set_values(LaneSession &S)
        {
            S.inf_bits = (S.wm == kWF32 || S.wm == kWUnitF32) ? 0x7F800000U : 0xFFFFFFFFU;
            if (S.kind == HYT_QUERY_PPR)
            {
                const MassScale m = mass_scale(1, S.damping, S.eps);
                S.seed_word = m.r0;
                S.eps_word = m.eps;
            }
        }

        void create_session(LaneSession &S, DeviceGraph *g, int kind, std::uint32_t K, float damping, float eps)
        {
            require(kind == HYT_QUERY_SSSP || kind == HYT_QUERY_PPR, "query kind must be SSSP or PPR");
            require(K >= 1 && K <= 64, "lanes must be in [1, 64]");
            require(g->V <= 0xFFFFFFF0ULL, "vertex count too large for lane sessions");
            if (kind == HYT_QUERY_PPR)
            {
                require(damping >= 0.0F && damping < 1.0F, "PageRank damping_factor must be in [0, 1)");
                require(eps > 0.0F && std::isfinite(eps), "PageRank epsilon must be finite and positive");
            }
            S.g = g;
            S.kind = kind;
            S.K = K;
            S.KP = static_cast<std::uint32_t>(pow2_at_least(K));
            S.damping = damping;
            S.eps = eps;
            S.wm = g->weight_mode();
            set_values(S);
            hip_check(hipSetDevice(g->device), "hipSetDevice");
            hip_check(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreateWithFlags(&S.done, hipEventDisableTiming), "hipEventCreate");
            const std::uint64_t V = g->V, Vp = V + kPad;
            const std::uint64_t QA = LaneSession::kAppend;
            const bool ppr = kind == HYT_QUERY_PPR;
            S.val.alloc(Vp * K * (ppr ? 8 : 4), "hipMalloc(lane state)");
            if (ppr)
            {
                S.rank.alloc(Vp * K * 8, "hipMalloc(lane ranks)");
                S.snap.alloc((Vp + QA) * K * 8, "hipMalloc(lane snapshot)"); // [F][K], F <= V + QA
            }
            S.mark.alloc(Vp * 8, "hipMalloc(lane marks)");
            S.qmask.alloc((Vp + QA) * 8, "hipMalloc(lane queue masks)");
            S.q.alloc((Vp + QA) * 4, "hipMalloc(lane queue)");
            const std::uint64_t cap = LaneSession::heavy_chunks(g->V, g->E) +
                                      static_cast<std::uint64_t>(K) * LaneSession::heavy_chunks_per_lane(g->E);
            S.heavy.alloc(cap * sizeof(HeavyChunkK), "hipMalloc(lane heavy chunks)");
            S.heavy_cap = cap;
            S.ctr.alloc(sizeof(Counters), "hipMalloc(lane counters)");
            S.lanes.alloc(8, "hipMalloc(lane mask)");
            S.colbuf.alloc(Vp * 4, "hipMalloc(lane output)");
            S.partial.alloc((kGatherBlocks + 1) * sizeof(double), "hipMalloc(lane partials)");
            S.acc.alloc(16, "hipMalloc(lane checksum)");
            S.host = std::make_unique<PinnedBuffer>(sizeof(Counters) + 8 + 16);
            S.val_p = S.val.ptr;
            S.rank_p = ppr ? S.rank.as<unsigned long long>() : nullptr;
            S.mark_p = S.mark.as<unsigned long long>();
            hip_check(hipMemsetAsync(S.mark.ptr, 0, Vp * 8, S.stream), "clear marks");
            hip_check(hipMemsetAsync(S.ctr.ptr, 0, sizeof(Counters), S.stream), "clear counters");
            // every lane starts free with a defined state
            if (ppr)
            {
                launch_fill<unsigned long long>(S.val.as<unsigned long long>(), Vp * K, 0ULL, S.stream);
                launch_fill<unsigned long long>(S.rank.as<unsigned long long>(), Vp * K, 0ULL, S.stream);
            }
            else
                launch_fill<std::uint32_t>(S.val.as<std::uint32_t>(), Vp * K, S.inf_bits, S.stream);
            hip_check(hipStreamSynchronize(S.stream), "lane session init");
        }

        void create_hybrid_session(LaneSession &S, HybridEngine &E, int kind, std::uint32_t K, float damping,
                                   float eps)
        {
            require(kind == HYT_QUERY_SSSP || kind == HYT_QUERY_PPR, "query kind must be SSSP or PPR");
            require(K >= 1 && K <= 64, "lanes must be in [1, 64]");
            if (kind == HYT_QUERY_PPR)
            {
                require(damping >= 0.0F && damping < 1.0F, "PageRank damping_factor must be in [0, 1)");
                require(eps > 0.0F && std::isfinite(eps), "PageRank epsilon must be finite and positive");
            }
            DeviceGraph *g = &hybrid_device_graph(E);
            S.g = g;
            S.hy = &E;
            S.kind = kind;
            S.K = K;
            S.KP = static_cast<std::uint32_t>(pow2_at_least(K));
            S.damping = damping;
            S.eps = eps;
            S.wm = g->weight_mode();
            set_values(S);
            const HybridLaneView v = hybrid_lanes_open(E, kind, K);
            S.val_p = v.val;
            S.rank_p = v.rank;
            S.mark_p = v.mark;
            S.stream = g->stream;
            S.own_stream = false;
            hip_check(hipEventCreateWithFlags(&S.done, hipEventDisableTiming), "hipEventCreate");
            const std::uint64_t Vp = g->V + kPad;
            S.colbuf.alloc(Vp * 4, "hipMalloc(lane output)");
            S.partial.alloc((kGatherBlocks + 1) * sizeof(double), "hipMalloc(lane partials)");
            S.acc.alloc(16, "hipMalloc(lane checksum)");
            S.host = std::make_unique<PinnedBuffer>(sizeof(Counters) + 8 + 16);
        }

        void admit(LaneSession &S, std::uint32_t lane, std::uint32_t source)
        {
            require(lane < S.K, "lane out of range");
            require(((S.occupied >> lane) & 1ULL) == 0, "lane is occupied");
            require(source < S.g->V, "source vertex is out of range");
            require(!S.launched, "a round is in flight");
            require(S.appended < LaneSession::kAppend, "too many admissions since the last round");
            const std::uint64_t V = S.g->V;
            auto seed = [&](auto *val, auto fill, auto seed_value) {
                using T = std::remove_pointer_t<decltype(val)>;
                lane_fill<T><<<grid_for(V), kBlock, 0, S.stream>>>(val, V, S.K, lane, fill);
                hip_check(hipGetLastError(), "launch lane_fill");
                if (S.kind == HYT_QUERY_PPR)
                {
                    lane_fill<unsigned long long><<<grid_for(V), kBlock, 0, S.stream>>>(S.rank_p, V, S.K, lane, 0ULL);
                    hip_check(hipGetLastError(), "launch lane_fill");
                }
                if (S.hy)
                {
                    lane_seed_mark<T><<<1, 64, 0, S.stream>>>(val, S.K, lane, source, seed_value, S.mark_p);
                    hip_check(hipGetLastError(), "launch lane_seed_mark");
                }
                else
                {
                    lane_seed<T><<<1, 64, 0, S.stream>>>(val, S.K, lane, source, seed_value, S.q.as<std::uint32_t>(),
                                                         S.qmask.as<unsigned long long>(), S.F);
                    hip_check(hipGetLastError(), "launch lane_seed");
                    S.F += 1;
                    S.pairs += 1;
                }
            };
            if (S.kind == HYT_QUERY_PPR)
                seed(static_cast<unsigned long long *>(S.val_p), 0ULL, S.seed_word);
            else
                seed(static_cast<std::uint32_t *>(S.val_p), S.inf_bits, 0U);
            S.appended += 1;
            S.running |= 1ULL << lane;
            S.occupied |= 1ULL << lane;
            S.lane_rounds[lane] = 0;
        }

        template <class OffT, int KP, class Op>
        void launch_round_t(LaneSession &S, const Op &op)
        {
            DeviceGraph &g = *S.g;
            hipStream_t s = S.stream;
            Counters *d_ctr = S.ctr.as<Counters>();
            hip_check(hipMemsetAsync(d_ctr, 0, kRoundCounterBytes, s), "reset lane counters");
            hip_check(hipMemsetAsync(S.lanes.ptr, 0, 8, s), "reset lane mask");
            KArgs a;
            a.rs = nullptr;
            a.F = S.F;
            a.ctr = d_ctr;
            a.heavy = S.heavy.as<HeavyChunkK>();
            a.heavy_cap = S.heavy_cap;
            a.light_max = kLaneLightMaxDegree;
            a.chunk = kLaneHeavyChunk;
            a.K = S.K;
            a.mark = S.mark.as<unsigned long long>();
            const QueueRowsK<OffT> rows{g.off.as<OffT>(), S.q.as<std::uint32_t>(), S.qmask.as<unsigned long long>()};
            if (S.kind == HYT_QUERY_PPR)
            {
                const std::uint64_t n = static_cast<std::uint64_t>(S.F) * S.K;
                pprk_snapshot<OffT><<<snapshot_blocks(n), kBlock, 0, s>>>(
                    rows.q, S.F, S.K, rows.qmask, S.val.as<unsigned long long>(), S.rank.as<unsigned long long>(),
                    S.snap.as<unsigned long long>(), g.off.as<OffT>(), op.damping);
                hip_check(hipGetLastError(), "launch pprk_snapshot");
            }
            if (use_items(S.form, S.pairs, S.F, S.KP))
            {
                expandk_light_items<QueueRowsK<OffT>, Op><<<blocks_for(S.F), kBlock, 0, s>>>(rows, op, a);
                hip_check(hipGetLastError(), "launch expandk_light_items");
                expandk_heavy_items<Op><<<kHeavyGrid, kBlock, 0, s>>>(op, a);
                hip_check(hipGetLastError(), "launch expandk_heavy_items");
                S.item_rounds += 1;
            }
            else
            {
                expandk_light<QueueRowsK<OffT>, KP, Op><<<blocks_for(S.F), kBlock, 0, s>>>(rows, op, a);
                hip_check(hipGetLastError(), "launch expandk_light");
                expandk_heavy<KP, Op><<<kHeavyGrid, kBlock, 0, s>>>(op, a);
                hip_check(hipGetLastError(), "launch expandk_heavy");
            }
            MarkSel sel{S.mark.as<unsigned long long>(), S.lanes.as<unsigned long long>()};
            compact_rows<unsigned long long, std::uint32_t, MarkSel><<<compact_blocks(g.V), kBlock, 0, s>>>(
                g.V, sel, S.q.as<std::uint32_t>(), S.qmask.as<unsigned long long>(), d_ctr, nullptr, nullptr);
            hip_check(hipGetLastError(), "launch lane compaction");
            unsigned char *h = S.host->as<unsigned char>();
            hip_check(hipMemcpyAsync(h, d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, s), "read lane counters");
            hip_check(hipMemcpyAsync(h + sizeof(Counters), S.lanes.ptr, 8, hipMemcpyDeviceToHost, s),
                      "read lane mask");
            hip_check(hipEventRecord(S.done, s), "hipEventRecord");
        }

        template <class OffT, int KP>
        void launch_kind(LaneSession &S)
        {
            DeviceGraph &g = *S.g;
            if (S.kind == HYT_QUERY_PPR)
            {
                PprKOp op;
                op.col = g.col.as<std::uint32_t>();
                op.res = S.val.as<unsigned long long>();
                op.snap = S.snap.as<unsigned long long>();
                op.K = S.K;
                op.damping = S.damping;
                op.eps = S.eps_word;
                launch_round_t<OffT, KP>(S, op);
                return;
            }
            auto run = [&](auto op) {
                op.col = g.col.as<std::uint32_t>();
                op.w = g.w.as<std::uint32_t>();
                op.dist = S.val.as<std::uint32_t>();
                op.K = S.K;
                op.overflow = &S.ctr.as<Counters>()->overflow;
                launch_round_t<OffT, KP>(S, op);
            };
            switch (S.wm)
            {
            case kWU32: run(SsspKOp<kWU32>{}); break;
            case kWF32: run(SsspKOp<kWF32>{}); break;
            case kWUnitU32: run(SsspKOp<kWUnitU32>{}); break;
            default: run(SsspKOp<kWUnitF32>{}); break;
            }
        }

        template <class OffT>
        void launch_kp(LaneSession &S)
        {
            switch (S.KP)
            {
            case 2: launch_kind<OffT, 2>(S); break;
            case 4: launch_kind<OffT, 4>(S); break;
            case 8: launch_kind<OffT, 8>(S); break;
            case 16: launch_kind<OffT, 16>(S); break;
            case 32: launch_kind<OffT, 32>(S); break;
            default: launch_kind<OffT, 64>(S); break;
            }
        }

        // A hybrid session's round runs on a worker thread (it plans on the
        // host between its kernels and copies); nothing else touches the
        // session until the round is collected by finish_round.
        void run_hybrid_round(LaneSession &S)
        {
            HybridEngine *E = S.hy;
            const int kind = S.kind;
            const std::uint32_t K = S.K, F = S.F, form = S.form;
            const std::uint64_t pairs = S.pairs;
            const bool merge = S.appended > 0;
            const float damping = S.damping, eps = S.eps;
            hyt_hybrid_stats *st = &S.hst;
            const int device = S.g->device;
            S.hy_round = std::async(std::launch::async, [=] {
                hip_check(hipSetDevice(device), "hipSetDevice");
                return hybrid_lanes_round(*E, kind, K, F, pairs, merge, damping, eps, form, *st);
            });
            S.launched = true;
        }

        // Collects the worker's round: its counters and lane mask land where
        // finish_round reads them.
        void collect_hybrid_round(LaneSession &S)
        {
            const HybridLaneRound r = S.hy_round.get();
            S.F = r.F;
            S.pairs = r.pairs;
            unsigned char *h = S.host->as<unsigned char>();
            std::memcpy(h, &r.c, sizeof(Counters));
            std::memcpy(h + sizeof(Counters), &r.lanes, 8);
            if (r.items)
                S.item_rounds += 1;
        }

        void launch_round(LaneSession &S)
        {
            S.launched = false;
            if (S.hy)
            {
                if (S.F > 0 || S.appended > 0)
                    run_hybrid_round(S);
                return;
            }
            if (S.F == 0)
                return;
            hip_check(hipSetDevice(S.g->device), "hipSetDevice");
            if (S.g->off64)
                launch_kp<std::uint64_t>(S);
            else
                launch_kp<std::uint32_t>(S);
            S.launched = true;
        }

        // Waits for the round, updates the frontier size and the lane masks;
        // returns the lanes that finished in it.
        std::uint64_t finish_round(LaneSession &S)
        {
            if (!S.launched)
                return 0;
            S.launched = false;
            if (S.hy)
                collect_hybrid_round(S);
            else
                hip_check(hipEventSynchronize(S.done), "lane round");
            const unsigned char *h = S.host->as<unsigned char>();
            Counters c;
            std::memcpy(&c, h, sizeof(Counters));
            std::uint64_t active;
            std::memcpy(&active, h + sizeof(Counters), 8);
            if (c.heavy_overflow)
                fail(HYT_ERR_INTERNAL, "heavy chunk queue overflow");
            if (c.overflow)
                fail(HYT_ERR_OVERFLOW, "u32 distance sum exceeded 2^32-2");
            for (std::uint32_t k = 0; k < S.K; ++k)
                if ((S.running >> k) & 1ULL)
                    S.lane_rounds[k] += 1;
            S.rounds += 1;
            S.edges += c.all_edges();
            S.frontier_rows += S.F;
            S.lane_rows += S.pairs;
            S.pairs = c.frontier_edges; // (row, lane) pairs of the next frontier
            S.F = c.next_count;
            S.appended = 0;
            const std::uint64_t finished = S.running & ~active;
            S.running &= active;
            return finished;
        }

        // True when the session's round in flight has completed.
        bool round_ready(LaneSession &S)
        {
            if (S.hy)
                return S.hy_round.wait_for(std::chrono::seconds(0)) == std::future_status::ready;
            const hipError_t e = hipEventQuery(S.done);
            if (e == hipErrorNotReady)
                return false;
            hip_check(e, "lane round");
            return true;
        }

        // Result of a finished lane into S.colbuf (SSSP distances, or PPR
        // ranks normalised to sum 1); optional checksum (entries != skip,
        // u64 sum of their bits) and copy to out.
        void read_lane(LaneSession &S, std::uint32_t lane, void *out, bool out_on_device, std::uint64_t *reached,
                       std::uint64_t *checksum)
        {
            require(lane < S.K, "lane out of range");
            require(((S.occupied >> lane) & 1ULL) != 0, "lane holds no query");
            require(((S.running >> lane) & 1ULL) == 0, "lane is still running");
            require(!S.launched, "a round is in flight");
            const std::uint64_t V = S.g->V;
            hipStream_t s = S.stream;
            const bool ppr = S.kind == HYT_QUERY_PPR;
            unsigned long long *partial = S.partial.as<unsigned long long>();
            const unsigned nb = grid_for(V);
            if (ppr)
            {
                const unsigned long long *res = static_cast<const unsigned long long *>(S.val_p);
                lane_mass_sum<<<nb, kBlock, 0, s>>>(res, S.rank_p, V, S.K, lane, partial);
                hip_check(hipGetLastError(), "launch lane_mass_sum");
                lane_total<<<1, 64, 0, s>>>(partial, nb, partial + kGatherBlocks);
                hip_check(hipGetLastError(), "launch lane_total");
                lane_mass_out<<<nb, kBlock, 0, s>>>(res, S.rank_p, V, S.K, lane, partial + kGatherBlocks,
                                                    S.colbuf.as<std::uint32_t>());
                hip_check(hipGetLastError(), "launch lane_mass_out");
            }
            else
            {
                lane_gather<<<nb, kBlock, 0, s>>>(static_cast<const std::uint32_t *>(S.val_p), V, S.K, lane,
                                                  S.colbuf.as<std::uint32_t>());
                hip_check(hipGetLastError(), "launch lane_gather");
            }
            std::uint64_t *hacc = reinterpret_cast<std::uint64_t *>(S.host->as<unsigned char>() + sizeof(Counters) + 8);
            if (reached || checksum)
            {
                hip_check(hipMemsetAsync(S.acc.ptr, 0, 16, s), "clear checksum");
                lane_checksum<<<nb, kBlock, 0, s>>>(S.colbuf.as<std::uint32_t>(), V, ppr ? 0U : S.inf_bits,
                                                    S.acc.as<unsigned long long>());
                hip_check(hipGetLastError(), "launch lane_checksum");
                hip_check(hipMemcpyAsync(hacc, S.acc.ptr, 16, hipMemcpyDeviceToHost, s), "read checksum");
            }
            if (out)
                hip_check(hipMemcpyAsync(out, S.colbuf.ptr, V * 4,
                                         out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s),
                          "copy lane result");
            hip_check(hipStreamSynchronize(s), "lane read");
            if (reached)
                *reached = hacc[0];
            if (checksum)
                *checksum = hacc[1];
        }

        void step(LaneSession *const *ss, std::uint32_t n, std::uint32_t rounds, std::uint64_t *finished)
        {
            for (std::uint32_t i = 0; i < n; ++i)
                finished[i] = 0;
            for (std::uint32_t r = 0; r < rounds; ++r)
            {
                bool any = false;
                // sessions in HBM launch first; a hybrid session's round then
                // runs while theirs are in flight
                for (int pass = 0; pass < 2; ++pass)
                    for (std::uint32_t i = 0; i < n; ++i)
                        if ((ss[i]->hy != nullptr) == (pass == 1))
                        {
                            launch_round(*ss[i]);
                            any = any || ss[i]->launched;
                        }
                if (!any)
                    break;
                for (std::uint32_t i = 0; i < n; ++i)
                    finished[i] |= finish_round(*ss[i]);
            }
        }
    } // namespace
} // namespace hyt

struct hyt_lanes
{
    hyt::LaneSession s;
};

namespace hyt
{
    namespace
    {
        using Clock = std::chrono::steady_clock;

        double secs(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

        // Lanes per session from the budget: every session gets one lane,
        // then each further lane goes to the session with the most queued
        // work per lane that still fits (work = its requests in the trace x
        // its graph's edge count, the cost of one query to first order), until
        // nothing fits or every session has min(demand, max_lanes) lanes. A
        // burst is done when the slowest session is, so lanes go where the
        // work is: a large graph's queries each stream far more edges.
        std::vector<std::uint32_t> plan_lanes(const std::vector<std::uint64_t> &fixed,
                                              const std::vector<std::uint64_t> &per_lane,
                                              const std::vector<std::uint64_t> &demand,
                                              const std::vector<double> &cost, std::uint64_t avail,
                                              std::uint32_t max_lanes)
        {
            const std::size_t n = fixed.size();
            std::vector<std::uint32_t> K(n, 1);
            std::uint64_t used = 0;
            for (std::size_t i = 0; i < n; ++i)
                used += fixed[i] + per_lane[i];
            if (used > avail)
                fail(HYT_ERR_OUT_OF_MEMORY, "device budget too small for one lane per session (" +
                                                std::to_string(used) + " bytes needed, " + std::to_string(avail) +
                                                " available)");
            while (true)
            {
                std::size_t best = n;
                double best_ratio = 0.0;
                for (std::size_t i = 0; i < n; ++i)
                {
                    if (K[i] >= max_lanes || K[i] >= demand[i] || used + per_lane[i] > avail)
                        continue;
                    const double ratio = static_cast<double>(demand[i]) * cost[i] / K[i];
                    if (best == n || ratio > best_ratio)
                    {
                        best = i;
                        best_ratio = ratio;
                    }
                }
                if (best == n)
                    break;
                K[best] += 1;
                used += per_lane[best];
            }
            return K;
        }

        // A served graph: resident in HBM, or in host memory behind hybrid
        // engines (one per query kind).
        struct ServeGraph
        {
            DeviceGraph *resident = nullptr;
            HybridEngine *engine[2] = {nullptr, nullptr};
            DeviceGraph *shape() const
            {
                if (resident)
                    return resident;
                for (HybridEngine *e : engine)
                    if (e)
                        return &hybrid_device_graph(*e);
                return nullptr;
            }
        };

        void serve_trace(const std::vector<ServeGraph> &graphs, const hyt_serve_config &cfg, const hyt_request *reqs,
                         std::uint64_t n, void *const *results, hyt_completion *out, hyt_serve_stats &st)
        {
            const std::uint32_t ngraphs = static_cast<std::uint32_t>(graphs.size());
            require(ngraphs > 0, "no graphs");
            for (const ServeGraph &sg : graphs)
                require(sg.shape() != nullptr, "null graph");
            require(n == 0 || (reqs != nullptr && out != nullptr), "null requests or completions");
            require(cfg.max_lanes >= 1 && cfg.max_lanes <= 64, "max_lanes must be in [1, 64]");
            require(cfg.lane_form <= HYT_LANES_ITEMS, "lane_form must be 0, 1 or 2");
            for (std::uint64_t i = 0; i < n; ++i)
            {
                require(reqs[i].graph < ngraphs, "request graph index out of range");
                require(reqs[i].kind == HYT_QUERY_SSSP || reqs[i].kind == HYT_QUERY_PPR, "request kind");
                const ServeGraph &sg = graphs[reqs[i].graph];
                require(sg.resident != nullptr || sg.engine[reqs[i].kind] != nullptr,
                        "no engine for this query kind on a host-resident graph");
                require(reqs[i].source < sg.shape()->V, "request source out of range");
                require(std::isfinite(reqs[i].arrival_s) && reqs[i].arrival_s >= 0.0, "arrival time");
                require(i == 0 || reqs[i].arrival_s >= reqs[i - 1].arrival_s, "requests must be sorted by arrival");
            }
            // sessions: one per (graph, kind) that appears
            std::vector<int> sid(static_cast<std::size_t>(ngraphs) * 2, -1);
            std::vector<std::uint64_t> fixed, per_lane, demand;
            std::vector<double> cost;
            std::vector<std::pair<std::uint32_t, int>> skey;
            for (std::uint64_t i = 0; i < n; ++i)
            {
                const std::size_t key = static_cast<std::size_t>(reqs[i].graph) * 2 + reqs[i].kind;
                if (sid[key] < 0)
                {
                    sid[key] = static_cast<int>(skey.size());
                    skey.emplace_back(reqs[i].graph, static_cast<int>(reqs[i].kind));
                    const ServeGraph &sg = graphs[reqs[i].graph];
                    const DeviceGraph &g = *sg.shape();
                    const int kind = static_cast<int>(reqs[i].kind);
                    if (sg.resident)
                    {
                        fixed.push_back(LaneSession::fixed_bytes(g.V, g.E, g.max_degree));
                        per_lane.push_back(LaneSession::lane_bytes(g.V, g.E, kind));
                    }
                    else
                    {
                        HybridEngine &E = *sg.engine[kind];
                        const std::uint64_t one = hybrid_lane_state_bytes(E, kind, 1);
                        const std::uint64_t lane = hybrid_lane_state_bytes(E, kind, 2) - one;
                        // + the session's own output column, partials and checksum
                        fixed.push_back(one - lane + (g.V + kPad) * 4 + (kGatherBlocks + 1) * sizeof(double) + 64);
                        per_lane.push_back(lane);
                    }
                    demand.push_back(0);
                    cost.push_back(static_cast<double>(g.E) + 1.0);
                }
                demand[static_cast<std::size_t>(sid[key])] += 1;
            }
            std::uint64_t graph_bytes = 0;
            std::vector<HybridEngine *> engines;
            for (const ServeGraph &sg : graphs)
            {
                if (sg.resident)
                    graph_bytes += sg.resident->device_bytes();
                else
                    for (HybridEngine *e : sg.engine)
                        if (e && std::find(engines.begin(), engines.end(), e) == engines.end())
                        {
                            engines.push_back(e);
                            graph_bytes += hybrid_engine_bytes(*e);
                        }
            }
            require(cfg.device_budget_bytes > graph_bytes, "device budget smaller than the resident graphs");
            st = hyt_serve_stats{};
            st.graph_bytes = graph_bytes;
            const std::size_t S = skey.size();
            std::vector<std::unique_ptr<LaneSession>> ses;
            if (S > 0)
            {
                const std::vector<std::uint32_t> K =
                    plan_lanes(fixed, per_lane, demand, cost, cfg.device_budget_bytes - graph_bytes, cfg.max_lanes);
                for (std::size_t i = 0; i < S; ++i)
                {
                    ses.push_back(std::make_unique<LaneSession>());
                    const ServeGraph &sg = graphs[skey[i].first];
                    if (sg.resident)
                        create_session(*ses.back(), sg.resident, skey[i].second, K[i], cfg.damping, cfg.epsilon);
                    else
                        create_hybrid_session(*ses.back(), *sg.engine[skey[i].second], skey[i].second, K[i],
                                              cfg.damping, cfg.epsilon);
                    ses.back()->form = cfg.lane_form;
                    st.state_bytes += fixed[i] + K[i] * per_lane[i];
                    if (i < 16)
                        st.lanes[i] = K[i];
                }
            }
            st.sessions = static_cast<std::uint32_t>(S);

            // sessions in HBM first: a hybrid session's round blocks the loop
            // while the others' rounds are in flight
            std::vector<std::size_t> order;
            for (int pass = 0; pass < 2; ++pass)
                for (std::size_t i = 0; i < S; ++i)
                    if ((ses[i]->hy != nullptr) == (pass == 1))
                        order.push_back(i);
            std::vector<std::deque<std::uint64_t>> waiting(S);
            std::vector<std::vector<std::int64_t>> lane_req(S);
            for (std::size_t i = 0; i < S; ++i)
                lane_req[i].assign(ses[i]->K, -1);
            std::uint64_t next = 0, completed = 0;
            // progress lines on stderr (HYT_SERVE_PROGRESS set): every
            // twentieth of the requests
            const bool progress = std::getenv("HYT_SERVE_PROGRESS") != nullptr;
            const std::uint64_t progress_every = n >= 20 ? n / 20 : 1;
            const auto t0 = Clock::now();
            double busy_from = -1.0; // start of the current stretch with a round in flight
            while (completed < n)
            {
                const double now = secs(t0);
                while (next < n && reqs[next].arrival_s <= now)
                {
                    const std::size_t key = static_cast<std::size_t>(reqs[next].graph) * 2 + reqs[next].kind;
                    waiting[static_cast<std::size_t>(sid[key])].push_back(next);
                    ++next;
                }
                bool in_flight = false, progress = false;
                for (const std::size_t i : order)
                {
                    LaneSession &L = *ses[i];
                    if (L.launched)
                    {
                        if (!round_ready(L))
                        {
                            in_flight = true;
                            continue;
                        }
                        const std::uint64_t fin = finish_round(L);
                        st.steps += 1;
                        progress = true;
                        for (std::uint32_t lane = 0; lane < L.K; ++lane)
                        {
                            if (((fin >> lane) & 1ULL) == 0)
                                continue;
                            const std::uint64_t r = static_cast<std::uint64_t>(lane_req[i][lane]);
                            read_lane(L, lane, results ? results[r] : nullptr, false, &out[r].reached,
                                      &out[r].checksum);
                            out[r].rounds = L.lane_rounds[lane];
                            out[r].done_s = secs(t0);
                            L.occupied &= ~(1ULL << lane);
                            lane_req[i][lane] = -1;
                            ++completed;
                            if (progress && (completed % progress_every == 0 || completed == n))
                            {
                                std::fprintf(stderr, "[serve] %llu / %llu requests done at %.1f s\n",
                                             static_cast<unsigned long long>(completed),
                                             static_cast<unsigned long long>(n), secs(t0));
                                std::fflush(stderr);
                            }
                        }
                    }
                    // admission into free lanes, then the session's next round
                    const std::uint64_t all = L.K == 64 ? ~0ULL : ((1ULL << L.K) - 1ULL);
                    while (!waiting[i].empty() && L.occupied != all && L.appended < LaneSession::kAppend)
                    {
                        std::uint32_t lane = 0;
                        while ((L.occupied >> lane) & 1ULL)
                            ++lane;
                        const std::uint64_t r = waiting[i].front();
                        waiting[i].pop_front();
                        admit(L, lane, reqs[r].source);
                        lane_req[i][lane] = static_cast<std::int64_t>(r);
                        out[r] = hyt_completion{};
                        out[r].admit_s = secs(t0);
                        out[r].session = static_cast<std::uint32_t>(i);
                        out[r].lane = lane;
                        st.admitted += 1;
                        progress = true;
                    }
                    if (L.running)
                    {
                        launch_round(L);
                        in_flight = in_flight || L.launched;
                    }
                }
                if (in_flight && busy_from < 0)
                    busy_from = now;
                if (!in_flight && busy_from >= 0)
                {
                    st.busy_s += secs(t0) - busy_from;
                    busy_from = -1.0;
                }
                if (!in_flight)
                {
                    if (next < n)
                    {
                        const double wait = reqs[next].arrival_s - secs(t0);
                        if (wait > 0)
                            std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 2e-4)));
                    }
                }
                else if (!progress)
                    std::this_thread::yield();
            }
            if (busy_from >= 0)
                st.busy_s += secs(t0) - busy_from;
            st.wall_s = secs(t0);
            for (auto &L : ses)
            {
                st.rounds += L->rounds;
                st.edges += L->edges;
                st.frontier_rows += L->frontier_rows;
                st.lane_rows += L->lane_rows;
                st.item_rounds += L->item_rounds;
                if (L->hy)
                {
                    st.host_bytes += L->hst.bytes_filter + L->hst.bytes_compaction;
                    st.host_zero_copy_bytes += L->hst.bytes_zero_copy;
                    st.parts_cached += L->hst.parts_cached;
                    st.parts_moved += L->hst.parts_filter + L->hst.parts_compaction + L->hst.parts_zero_copy;
                }
            }
        }
    } // namespace
} // namespace hyt

extern "C" {

int hyt_lanes_state_bytes(uint64_t num_vertices, uint64_t num_edges, uint64_t max_degree, uint32_t kind,
                          uint32_t lanes, uint64_t *bytes)
{
    return hyt::guarded("hyt_lanes_state_bytes", [&] {
        hyt::require(bytes != nullptr, "null output");
        hyt::require(kind == HYT_QUERY_SSSP || kind == HYT_QUERY_PPR, "query kind must be SSSP or PPR");
        *bytes = hyt::LaneSession::fixed_bytes(num_vertices, num_edges, max_degree) +
                 static_cast<std::uint64_t>(lanes) *
                     hyt::LaneSession::lane_bytes(num_vertices, num_edges, static_cast<int>(kind));
    });
}

int hyt_lanes_create(hyt_graph_t g, uint32_t kind, uint32_t lanes, float damping, float epsilon, hyt_lanes_t *out)
{
    return hyt::guarded("hyt_lanes_create", [&] {
        hyt::require(g != nullptr && out != nullptr, "null pointer");
        *out = nullptr;
        auto h = std::make_unique<hyt_lanes>();
        hyt::create_session(h->s, &g->g, static_cast<int>(kind), lanes, damping, epsilon);
        *out = h.release();
    });
}

int hyt_lanes_free(hyt_lanes_t s)
{
    return hyt::guarded("hyt_lanes_free", [&] { delete s; });
}

int hyt_lanes_admit(hyt_lanes_t s, uint32_t lane, uint32_t source)
{
    return hyt::guarded("hyt_lanes_admit", [&] {
        hyt::require(s != nullptr, "null session");
        hyt::admit(s->s, lane, source);
    });
}

int hyt_lanes_step(hyt_lanes_t *sessions, uint32_t count, uint32_t rounds, uint64_t *finished, uint64_t *running)
{
    return hyt::guarded("hyt_lanes_step", [&] {
        hyt::require(count == 0 || (sessions != nullptr && finished != nullptr), "null pointer");
        std::vector<hyt::LaneSession *> ss(count);
        for (uint32_t i = 0; i < count; ++i)
        {
            hyt::require(sessions[i] != nullptr, "null session");
            ss[i] = &sessions[i]->s;
            for (uint32_t j = 0; j < i; ++j)
                hyt::require(ss[j] != ss[i], "a session appears twice");
        }
        hyt::step(ss.data(), count, rounds, finished);
        if (running)
            for (uint32_t i = 0; i < count; ++i)
                running[i] = ss[i]->running;
    });
}

int hyt_lanes_read(hyt_lanes_t s, uint32_t lane, void *out, uint32_t out_on_device, uint64_t *reached,
                   uint64_t *checksum, uint32_t *rounds)
{
    return hyt::guarded("hyt_lanes_read", [&] {
        hyt::require(s != nullptr, "null session");
        hyt::read_lane(s->s, lane, out, out_on_device != 0, reached, checksum);
        if (rounds)
            *rounds = lane < s->s.K ? s->s.lane_rounds[lane] : 0;
    });
}

int hyt_lanes_release(hyt_lanes_t s, uint32_t lane)
{
    return hyt::guarded("hyt_lanes_release", [&] {
        hyt::require(s != nullptr, "null session");
        hyt::require(lane < s->s.K, "lane out of range");
        hyt::require(((s->s.running >> lane) & 1ULL) == 0, "lane is still running");
        s->s.occupied &= ~(1ULL << lane);
    });
}

int hyt_lanes_set_form(hyt_lanes_t s, uint32_t form)
{
    return hyt::guarded("hyt_lanes_set_form", [&] {
        hyt::require(s != nullptr, "null session");
        hyt::require(form <= HYT_LANES_ITEMS, "lane form must be 0, 1 or 2");
        s->s.form = form;
    });
}

int hyt_lanes_info(hyt_lanes_t s, uint32_t *lanes, uint64_t *occupied, uint64_t *running, uint32_t *frontier,
                   uint64_t *state_bytes)
{
    return hyt::guarded("hyt_lanes_info", [&] {
        hyt::require(s != nullptr, "null session");
        const hyt::LaneSession &L = s->s;
        if (lanes)
            *lanes = L.K;
        if (occupied)
            *occupied = L.occupied;
        if (running)
            *running = L.running;
        if (frontier)
            *frontier = L.F;
        if (state_bytes)
            *state_bytes = (L.hy ? hyt::hybrid_lane_state_bytes(*L.hy, L.kind, L.K) : 0) + L.val.bytes +
                           L.rank.bytes + L.snap.bytes + L.mark.bytes + L.qmask.bytes + L.q.bytes + L.heavy.bytes +
                           L.ctr.bytes + L.lanes.bytes + L.colbuf.bytes + L.partial.bytes + L.acc.bytes;
    });
}

int hyt_serve_config_default(hyt_serve_config *out)
{
    return hyt::guarded("hyt_serve_config_default", [&] {
        hyt::require(out != nullptr, "null output");
        hyt_serve_config c{};
        c.device_budget_bytes = 64ULL << 30;
        c.max_lanes = 64;
        c.damping = 0.85F;
        c.epsilon = 1e-6F;
        *out = c;
    });
}

int hyt_serve_trace(const hyt_graph_t *graphs, uint32_t num_graphs, const hyt_serve_config *config,
                    const hyt_request *requests, uint64_t num_requests, void *const *results,
                    hyt_completion *completions, hyt_serve_stats *stats)
{
    return hyt::guarded("hyt_serve_trace", [&] {
        hyt::require(config != nullptr, "null config");
        hyt::require(graphs != nullptr || num_graphs == 0, "null graph array");
        std::vector<hyt::ServeGraph> gs(num_graphs);
        for (uint32_t i = 0; i < num_graphs; ++i)
        {
            hyt::require(graphs[i] != nullptr, "null graph");
            gs[i].resident = &graphs[i]->g;
        }
        hyt_serve_stats st{};
        hyt::serve_trace(gs, *config, requests, num_requests, results, completions, st);
        if (stats)
            *stats = st;
    });
}

int hyt_serve_trace_mixed(const hyt_serve_graph *graphs, uint32_t num_graphs, const hyt_serve_config *config,
                          const hyt_request *requests, uint64_t num_requests, void *const *results,
                          hyt_completion *completions, hyt_serve_stats *stats)
{
    return hyt::guarded("hyt_serve_trace_mixed", [&] {
        hyt::require(config != nullptr, "null config");
        hyt::require(graphs != nullptr || num_graphs == 0, "null graph array");
        std::vector<hyt::ServeGraph> gs(num_graphs);
        for (uint32_t i = 0; i < num_graphs; ++i)
        {
            if (graphs[i].resident)
                gs[i].resident = &graphs[i].resident->g;
            else
            {
                gs[i].engine[HYT_QUERY_SSSP] =
                    graphs[i].sssp_engine ? &hyt::hybrid_engine(graphs[i].sssp_engine) : nullptr;
                gs[i].engine[HYT_QUERY_PPR] = graphs[i].ppr_engine ? &hyt::hybrid_engine(graphs[i].ppr_engine) : nullptr;
            }
        }
        // a session keeps its lanes in its engine's state: one session per engine
        std::vector<const void *> seen;
        for (const auto &g : gs)
            for (const auto *e : g.engine)
                if (e)
                {
                    hyt::require(std::find(seen.begin(), seen.end(), e) == seen.end(),
                                 "each hybrid engine may serve one graph and one query kind");
                    seen.push_back(e);
                }
        hyt_serve_stats st{};
        hyt::serve_trace(gs, *config, requests, num_requests, results, completions, st);
        if (stats)
            *stats = st;
    });
}

int hyt_lanes_create_hybrid(hyt_hybrid_t h, uint32_t kind, uint32_t lanes, float damping, float epsilon,
                            hyt_lanes_t *out)
{
    return hyt::guarded("hyt_lanes_create_hybrid", [&] {
        hyt::require(h != nullptr && out != nullptr, "null pointer");
        *out = nullptr;
        auto s = std::make_unique<hyt_lanes>();
        hyt::create_hybrid_session(s->s, hyt::hybrid_engine(h), static_cast<int>(kind), lanes, damping, epsilon);
        *out = s.release();
    });
}

} // extern "C"
