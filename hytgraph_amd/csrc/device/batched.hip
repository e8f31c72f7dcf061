// Multi-source [V][K] SSSP and personalized PageRank (2 <= K <= 64) over a
// graph resident in HBM, with the K-lane kernels of lanes.hpp.
//
// SSSP-K reads live distances (async rounds; final distances equal the
// single-source results): a lane marks v when its candidate beats a plain load
// of dist[v][k], which cannot miss an improvement (the first lowering atomic
// of a round comes from a lane whose load passed). PPR-K needs the exact
// eps-crossing, so it uses returning integer atomics on 64-bit fixed-point
// residual words (mass.hpp), and snapshots the active residuals of the
// frontier before each round like the single-source form.
#include "lanes.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace hyt
{
    namespace
    {
        using namespace lanes;

        // Workspace of the batched forms (grown to the largest K used).
        struct KWorkspace
        {
            DeviceBuffer val, rank, snap, mark, qmask, heavy, seeds, partial, lanes, round;
            std::unique_ptr<PinnedBuffer> host_round;
            std::uint64_t heavy_cap = 0;
        };

        // The graph owns its batched workspace (released with it or by
        // hyt_graph_trim).
        KWorkspace &kws_of(DeviceGraph &g)
        {
            if (!g.kws)
                g.kws = std::make_shared<KWorkspace>();
            return *static_cast<KWorkspace *>(g.kws.get());
        }


        void prepare(DeviceGraph &g, KWorkspace &kw, std::uint32_t K, bool ppr, std::uint32_t light_max,
                     std::uint32_t chunk)
        {
            const std::uint64_t VK = g.V * K;
            // SSSP: u32 distances; PPR: u64 residual, rank and snapshot words
            kw.val.ensure(VK * (ppr ? 8 : 4), "hipMalloc([V][K] state)");
            if (ppr)
            {
                kw.rank.ensure(VK * 8, "hipMalloc([V][K] rank)");
                kw.snap.ensure(VK * 8, "hipMalloc([V][K] snapshot)");
            }
            if (kw.mark.bytes < (g.V + kPad) * 8)
            {
                // mark is kept all-zero between calls (compaction clears what it reads)
                kw.mark.alloc((g.V + kPad) * 8, "hipMalloc(lane marks)");
                kw.qmask.alloc((g.V + kPad) * 8, "hipMalloc(queue masks)");
                hip_check(hipMemsetAsync(kw.mark.ptr, 0, (g.V + kPad) * 8, g.stream), "clear marks");
            }
            const std::uint64_t cap = lane_chunk_capacity_k(g.V, g.E, light_max, chunk, K);
            if (kw.heavy.bytes < cap * sizeof(HeavyChunkK))
                kw.heavy.alloc(cap * sizeof(HeavyChunkK), "hipMalloc(heavy chunks)");
            kw.heavy_cap = kw.heavy.bytes / sizeof(HeavyChunkK);
            kw.seeds.ensure(64 * 4 * 4, "hipMalloc(seeds)");
            kw.lanes.ensure(8, "hipMalloc(lane mask)");
            kw.round.ensure(sizeof(RoundState), "hipMalloc(round state)");
            if (!kw.host_round)
                kw.host_round = std::make_unique<PinnedBuffer>(sizeof(RoundState));
            kw.partial.ensure((1024 + 1) * 64 * sizeof(double), "hipMalloc(partials)");
            ensure_workspace(g, light_max, chunk);
        }

        Counters compact_marks(DeviceGraph &g, KWorkspace &kw)
        {
            Counters *d_ctr = g.ws.ctr.as<Counters>();
            MarkSel sel{kw.mark.as<unsigned long long>(), kw.lanes.as<unsigned long long>()};
            compact_rows<unsigned long long, std::uint32_t, MarkSel><<<compact_blocks(g.V), kBlock, 0, g.stream>>>(
                g.V, sel, g.ws.q.as<std::uint32_t>(), kw.qmask.as<unsigned long long>(), d_ctr, nullptr, nullptr);
            hip_check(hipGetLastError(), "launch mark compaction");
            Counters *h_ctr = g.host_ctr->as<Counters>();
            hip_check(hipMemcpyAsync(h_ctr, d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, g.stream), "read counters");
            hip_check(hipStreamSynchronize(g.stream), "batched round");
            return *h_ctr;
        }

        // Rounds per captured HIP graph in device-driven mode.
        constexpr int kRoundsPerGraph = 4;

        // Starts a device-driven round: the queue the last compaction wrote
        // becomes this round's (or the round limit ends the run), the form is
        // chosen, and the round counters are cleared (a kernel, not a memset,
        // so the sequence can be captured).
        __global__ void round_begin(Counters *ctr, RoundState *rs, std::uint32_t KP, std::uint32_t forced,
                                    std::uint64_t max_rounds)
        {
            if (threadIdx.x != 0 || blockIdx.x != 0)
                return;
            std::uint32_t F = ctr->next_count;
            const std::uint64_t pairs = ctr->frontier_edges;
            if (F != 0 && rs->rounds >= max_rounds)
            {
                rs->truncated = 1;
                F = 0;
            }
            rs->F = F;
            rs->pairs = pairs;
            rs->form = forced == HYT_LANES_AUTO
                           ? (pairs * 2 < static_cast<std::uint64_t>(F) * KP ? HYT_LANES_ITEMS : HYT_LANES_GROUP)
                           : forced;
            if (F != 0)
            {
                rs->rounds += 1;
                rs->frontier_total += F;
            }
            else
                rs->done = 1;
            ctr->next_count = 0;
            ctr->heavy_count = 0;
            ctr->edges = 0;
            for (int i = 0; i < kEdgeSlots; ++i)
                ctr->slot[i].v = ctr->slot[i].h = 0;
            ctr->frontier_edges = 0;
        }

        __global__ void round_end(const Counters *ctr, RoundState *rs)
        {
            if (threadIdx.x != 0 || blockIdx.x != 0 || rs->F == 0)
                return;
            rs->edges_total += ctr->all_edges();
            rs->updates_total += ctr->next_count;
        }

        // Device-driven rounds: kRoundsPerGraph rounds (round_begin, PPR
        // snapshot, both forms' light and heavy kernels - the form not chosen
        // exits at once -, the mark compaction, round_end) are captured into
        // one HIP graph, which is replayed until a round finds the queue
        // empty; the host reads the round state once per replay.
        template <class OffT, int KP, class Op>
        void run_k(DeviceGraph &g, KWorkspace &kw, Op op, std::uint32_t K, const hyt_run_options &o,
                   hyt_run_stats &st, std::uint32_t light_max, std::uint32_t chunk, bool ppr)
        {
            hipStream_t s = g.stream;
            Counters *d_ctr = g.ws.ctr.as<Counters>();
            const std::uint64_t max_rounds = o.max_rounds ? o.max_rounds : (ppr ? 100000 : g.V);
            RoundState *rs = kw.round.as<RoundState>();
            KArgs a;
            a.rs = rs;
            a.F = 0;
            a.ctr = d_ctr;
            a.heavy = kw.heavy.as<HeavyChunkK>();
            a.heavy_cap = kw.heavy_cap;
            a.light_max = light_max;
            a.chunk = chunk;
            a.K = K;
            a.mark = kw.mark.as<unsigned long long>();
            hip_check(hipMemsetAsync(d_ctr, 0, sizeof(Counters), s), "reset counters");
            hip_check(hipMemsetAsync(rs, 0, sizeof(RoundState), s), "reset round state");
            const Counters c0 = compact_marks(g, kw); // seeds -> first queue (eager)
            st.converged = c0.next_count == 0 ? 1 : 0;
            if (c0.next_count == 0)
                return;

            // one 256-row tile per workgroup at the largest queue (V rows):
            // workgroups past this round's rows exit at once, and no
            // workgroup waits at a barrier for another tile's slow wavefront
            const unsigned light_grid = blocks_for(g.V);
            const QueueRowsK<OffT> rows{g.off.as<OffT>(), g.ws.q.as<std::uint32_t>(), kw.qmask.as<unsigned long long>()};
            const unsigned nb = compact_blocks(g.V);
            MarkSel sel{kw.mark.as<unsigned long long>(), kw.lanes.as<unsigned long long>()};
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin round capture");
            for (int r = 0; r < kRoundsPerGraph; ++r)
            {
                round_begin<<<1, 64, 0, s>>>(d_ctr, rs, KP, o.lane_form, max_rounds);
                if (ppr)
                    pprk_snapshot<OffT><<<snapshot_blocks(g.V * static_cast<std::uint64_t>(K)), kBlock, 0, s>>>(
                        rows.q, 0, K, rows.qmask, kw.val.as<unsigned long long>(), kw.rank.as<unsigned long long>(),
                        kw.snap.as<unsigned long long>(), g.off.as<OffT>(), op.damping, rs);
                expandk_light<QueueRowsK<OffT>, KP, Op><<<light_grid, kBlock, 0, s>>>(rows, op, a);
                expandk_light_items<QueueRowsK<OffT>, Op><<<light_grid, kBlock, 0, s>>>(rows, op, a);
                expandk_heavy<KP, Op><<<kHeavyGrid, kBlock, 0, s>>>(op, a);
                expandk_heavy_items<Op><<<kHeavyGrid, kBlock, 0, s>>>(op, a);
                compact_rows<unsigned long long, std::uint32_t, MarkSel><<<nb, kBlock, 0, s>>>(
                    g.V, sel, g.ws.q.as<std::uint32_t>(), kw.qmask.as<unsigned long long>(), d_ctr, nullptr, nullptr);
                round_end<<<1, 64, 0, s>>>(d_ctr, rs);
            }
            const hipError_t launch_err = hipGetLastError();
            const hipError_t cap_err = hipStreamEndCapture(s, &graph);
            hip_check(launch_err, "launch captured round");
            hip_check(cap_err, "end round capture");
            struct GraphGuard
            {
                hipGraph_t g;
                hipGraphExec_t e;
                ~GraphGuard()
                {
                    if (e)
                        (void)hipGraphExecDestroy(e);
                    if (g)
                        (void)hipGraphDestroy(g);
                }
            } guard{graph, nullptr};
            hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate round graph");
            guard.e = exec;

            RoundState *h_rs = reinterpret_cast<RoundState *>(kw.host_round->ptr);
            Counters *h_ctr = g.host_ctr->as<Counters>();
            while (true)
            {
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev0, s), "hipEventRecord");
                hip_check(hipGraphLaunch(exec, s), "launch round graph");
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev1, s), "hipEventRecord");
                hip_check(hipMemcpyAsync(h_rs, rs, sizeof(RoundState), hipMemcpyDeviceToHost, s), "read round state");
                hip_check(hipMemcpyAsync(h_ctr, d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, s), "read counters");
                hip_check(hipStreamSynchronize(s), "round graph");
                if (o.time_kernels)
                {
                    float ms = 0;
                    hip_check(hipEventElapsedTime(&ms, g.ev0, g.ev1), "hipEventElapsedTime");
                    st.expand_ms += ms;
                    st.expand_launches += 1;
                }
                if (h_ctr->heavy_overflow)
                    fail(HYT_ERR_INTERNAL, "heavy chunk queue overflow");
                if (h_ctr->overflow)
                {
                    st.overflow = 1;
                    fail(HYT_ERR_OVERFLOW, "u32 distance sum exceeded 2^32-2");
                }
                if (h_rs->done)
                    break;
            }
            st.rounds = h_rs->rounds;
            st.frontier_total = h_rs->frontier_total;
            st.edges_total = h_rs->edges_total;
            st.updates_total = h_rs->updates_total;
            st.converged = h_rs->truncated ? 0 : 1;
        }

        template <class OffT, class Op>
        void dispatch_kp(DeviceGraph &g, KWorkspace &kw, Op op, std::uint32_t K, const hyt_run_options &o,
                         hyt_run_stats &st, std::uint32_t lm, std::uint32_t ch, bool ppr)
        {
            switch (pow2_at_least(K))
            {
            case 2: run_k<OffT, 2>(g, kw, op, K, o, st, lm, ch, ppr); break;
            case 4: run_k<OffT, 4>(g, kw, op, K, o, st, lm, ch, ppr); break;
            case 8: run_k<OffT, 8>(g, kw, op, K, o, st, lm, ch, ppr); break;
            case 16: run_k<OffT, 16>(g, kw, op, K, o, st, lm, ch, ppr); break;
            case 32: run_k<OffT, 32>(g, kw, op, K, o, st, lm, ch, ppr); break;
            default: run_k<OffT, 64>(g, kw, op, K, o, st, lm, ch, ppr); break;
            }
        }

        template <class Op>
        void dispatch_off(DeviceGraph &g, KWorkspace &kw, Op op, std::uint32_t K, const hyt_run_options &o,
                          hyt_run_stats &st, std::uint32_t lm, std::uint32_t ch, bool ppr)
        {
            if (g.off64)
                dispatch_kp<std::uint64_t>(g, kw, op, K, o, st, lm, ch, ppr);
            else
                dispatch_kp<std::uint32_t>(g, kw, op, K, o, st, lm, ch, ppr);
        }

        void split_params(const hyt_run_options &o, std::uint32_t &lm, std::uint32_t &ch)
        {
            lm = o.light_max_degree ? o.light_max_degree : kBatchedLightMaxDegree;
            ch = o.heavy_chunk ? o.heavy_chunk : kBatchedHeavyChunk;
            require(lm <= (1U << 20), "light_max_degree must be <= 2^20");
            require(ch >= 64 && ch <= (1U << 24), "heavy_chunk must be in [64, 2^24]");
            require(o.lane_form <= HYT_LANES_ITEMS, "lane_form must be 0, 1 or 2");
        }
    } // namespace

    void sssp_batched(DeviceGraph &g, const std::uint32_t *sources, std::uint32_t K, std::uint32_t *dist_out,
                      const hyt_run_options &o, hyt_run_stats &st)
    {
        std::uint32_t lm, ch;
        split_params(o, lm, ch);
        KWorkspace &kw = kws_of(g);
        prepare(g, kw, K, false, lm, ch);
        const int wm = g.weight_mode();
        const std::uint32_t inf = (wm == kWF32 || wm == kWUnitF32) ? 0x7F800000U : 0xFFFFFFFFU;
        launch_fill<std::uint32_t>(kw.val.as<std::uint32_t>(), g.V * K, inf, g.stream);
        const SeedPlan p = plan_seeds(sources, K);
        upload_seeds<std::uint32_t>(g.stream, kw.seeds, kw.mark.as<unsigned long long>(), p, sources, K,
                                    kw.val.as<std::uint32_t>(), 0U);
        auto run = [&](auto op) {
            op.col = g.col.as<std::uint32_t>();
            op.w = g.w.as<std::uint32_t>();
            op.dist = kw.val.as<std::uint32_t>();
            op.K = K;
            op.overflow = &g.ws.ctr.as<Counters>()->overflow;
            dispatch_off(g, kw, op, K, o, st, lm, ch, false);
        };
        switch (wm)
        {
        case kWU32: run(SsspKOp<kWU32>{}); break;
        case kWF32: run(SsspKOp<kWF32>{}); break;
        case kWUnitU32: run(SsspKOp<kWUnitU32>{}); break;
        default: run(SsspKOp<kWUnitF32>{}); break;
        }
        hip_check(hipMemcpyAsync(dist_out, kw.val.ptr, g.V * K * 4,
                                 o.out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, g.stream),
                  "copy distances");
        hip_check(hipStreamSynchronize(g.stream), "sssp output");
    }

    void ppr_batched(DeviceGraph &g, const std::uint32_t *seeds, std::uint32_t K, float damping, float eps,
                     float *out, const hyt_run_options &o, hyt_run_stats &st)
    {
        require(damping >= 0.0F && damping < 1.0F, "PageRank damping_factor must be in [0, 1)");
        require(eps > 0.0F && std::isfinite(eps), "PageRank epsilon must be finite and positive");
        std::uint32_t lm, ch;
        split_params(o, lm, ch);
        KWorkspace &kw = kws_of(g);
        prepare(g, kw, K, true, lm, ch);
        const std::uint64_t VK = g.V * K;
        const MassScale m = mass_scale(1, damping, eps);
        launch_fill<unsigned long long>(kw.val.as<unsigned long long>(), VK, 0ULL, g.stream);
        launch_fill<unsigned long long>(kw.rank.as<unsigned long long>(), VK, 0ULL, g.stream);
        const SeedPlan p = plan_seeds(seeds, K);
        upload_seeds<unsigned long long>(g.stream, kw.seeds, kw.mark.as<unsigned long long>(), p, seeds, K,
                                         kw.val.as<unsigned long long>(), m.r0);
        PprKOp op;
        op.col = g.col.as<std::uint32_t>();
        op.res = kw.val.as<unsigned long long>();
        op.snap = kw.snap.as<unsigned long long>();
        op.K = K;
        op.damping = damping;
        op.eps = m.eps;
        dispatch_off(g, kw, op, K, o, st, lm, ch, true);

        float *dst = o.out_on_device ? out : kw.snap.as<float>();
        pprk_finish(g.stream, kw.rank.as<unsigned long long>(), op.res, g.V, K, kw.partial.as<unsigned long long>(),
                    dst);
        if (!o.out_on_device)
            hip_check(hipMemcpyAsync(out, dst, VK * 4, hipMemcpyDeviceToHost, g.stream), "copy ppr");
        hip_check(hipStreamSynchronize(g.stream), "ppr output");
    }
} // namespace hyt
