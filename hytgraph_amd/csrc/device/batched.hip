// Multi-source [V][K] SSSP and personalized PageRank (2 <= K <= 64) over a
// graph resident in HBM, with the K-lane kernels of lanes.hpp.
//
// SSSP-K reads live distances (async rounds; final distances equal the
// single-source results): a lane marks v when its candidate beats a plain load
// of dist[v][k], which cannot miss an improvement (the first lowering atomic
// of a round comes from a lane whose load passed). PPR-K needs the exact
// eps-crossing, so it uses returning float atomics, and snapshots the active
// residuals of the frontier before each round like the single-source form.
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

        __global__ void seed_rows(std::uint32_t nrows, const std::uint32_t *rows, const unsigned long long *masks,
                                  const std::uint32_t *lane_src, std::uint32_t K, unsigned long long *mark,
                                  std::uint32_t *val, std::uint32_t seed_value)
        {
            const std::uint32_t t = threadIdx.x;
            if (t < nrows)
                mark[rows[t]] = masks[t];
            if (t < K)
                val[static_cast<std::uint64_t>(lane_src[t]) * K + t] = seed_value;
        }

        // Column sums of x = rank + res ([V][K]) into partial[block][K] (double).
        __global__ __launch_bounds__(kBlock) void pprk_fold(float *rank, const float *res, std::uint64_t V,
                                                           std::uint32_t K, std::uint32_t KP, double *partial)
        {
            __shared__ double s[kBlock];
            const std::uint32_t rows_per = kBlock / KP;
            const std::uint32_t k = threadIdx.x % KP;
            const std::uint32_t r0 = threadIdx.x / KP;
            double acc = 0.0;
            if (k < K)
            {
                for (std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * rows_per + r0; v < V;
                     v += static_cast<std::uint64_t>(gridDim.x) * rows_per)
                {
                    const std::uint64_t idx = v * K + k;
                    const float x = rank[idx] + res[idx];
                    rank[idx] = x;
                    acc += x;
                }
            }
            s[threadIdx.x] = acc;
            __syncthreads();
            if (threadIdx.x < KP && k < K)
            {
                double t = 0.0;
                for (std::uint32_t r = 0; r < rows_per; ++r)
                    t += s[r * KP + threadIdx.x];
                partial[static_cast<std::uint64_t>(blockIdx.x) * 64 + threadIdx.x] = t;
            }
        }

        __global__ void pprk_total(double *partial, unsigned nblocks, std::uint32_t K, double *total)
        {
            const std::uint32_t k = threadIdx.x;
            if (k >= K)
                return;
            double t = 0.0;
            for (unsigned b = 0; b < nblocks; ++b)
                t += partial[static_cast<std::uint64_t>(b) * 64 + k];
            total[k] = t;
        }

        __global__ void pprk_normalize(const float *x, std::uint64_t n, std::uint32_t K, const double *total,
                                       float *out)
        {
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; i < n; i += stride)
            {
                const double t = total[i % K];
                out[i] = static_cast<float>(t > 0.0 ? static_cast<double>(x[i]) / t : 0.0);
            }
        }

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


        struct SeedPlan
        {
            std::vector<std::uint32_t> rows;
            std::vector<unsigned long long> masks;
        };

        SeedPlan plan_seeds(const std::uint32_t *src, std::uint32_t K)
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

        void prepare(DeviceGraph &g, KWorkspace &kw, std::uint32_t K, bool ppr, std::uint32_t light_max,
                     std::uint32_t chunk)
        {
            const std::uint64_t VK = g.V * K;
            kw.val.ensure(VK * 4, "hipMalloc([V][K] state)");
            if (ppr)
            {
                kw.rank.ensure(VK * 4, "hipMalloc([V][K] rank)");
                kw.snap.ensure(VK * 4, "hipMalloc([V][K] snapshot)");
            }
            if (kw.mark.bytes < (g.V + kPad) * 8)
            {
                // mark is kept all-zero between calls (compaction clears what it reads)
                kw.mark.alloc((g.V + kPad) * 8, "hipMalloc(lane marks)");
                kw.qmask.alloc((g.V + kPad) * 8, "hipMalloc(queue masks)");
                hip_check(hipMemsetAsync(kw.mark.ptr, 0, (g.V + kPad) * 8, g.stream), "clear marks");
            }
            const std::uint64_t cap = lane_chunk_capacity(g.V, g.E, light_max, chunk);
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

        void upload_seeds(DeviceGraph &g, KWorkspace &kw, const SeedPlan &p, const std::uint32_t *src,
                          std::uint32_t K, std::uint32_t *val, std::uint32_t seed_value)
        {
            // seeds buffer: rows u32[64] | masks u64[64] | lane sources u32[64]
            std::vector<unsigned char> blob(64 * 4 + 64 * 8 + 64 * 4, 0);
            std::memcpy(blob.data(), p.rows.data(), p.rows.size() * 4);
            std::memcpy(blob.data() + 256, p.masks.data(), p.masks.size() * 8);
            std::memcpy(blob.data() + 768, src, K * 4);
            unsigned char *d = kw.seeds.as<unsigned char>();
            hip_check(hipMemcpyAsync(d, blob.data(), blob.size(), hipMemcpyHostToDevice, g.stream), "upload seeds");
            seed_rows<<<1, 64, 0, g.stream>>>(static_cast<std::uint32_t>(p.rows.size()),
                                              reinterpret_cast<const std::uint32_t *>(d),
                                              reinterpret_cast<const unsigned long long *>(d + 256),
                                              reinterpret_cast<const std::uint32_t *>(d + 768), K,
                                              kw.mark.as<unsigned long long>(), val, seed_value);
            hip_check(hipGetLastError(), "launch seed_rows");
            hip_check(hipStreamSynchronize(g.stream), "seed upload");
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
            ctr->frontier_edges = 0;
        }

        __global__ void round_end(const Counters *ctr, RoundState *rs)
        {
            if (threadIdx.x != 0 || blockIdx.x != 0 || rs->F == 0)
                return;
            rs->edges_total += ctr->edges;
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
            a.q = g.ws.q.as<std::uint32_t>();
            a.qmask = kw.qmask.as<unsigned long long>();
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
            const unsigned nb = compact_blocks(g.V);
            MarkSel sel{kw.mark.as<unsigned long long>(), kw.lanes.as<unsigned long long>()};
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin round capture");
            for (int r = 0; r < kRoundsPerGraph; ++r)
            {
                round_begin<<<1, 64, 0, s>>>(d_ctr, rs, KP, o.lane_form, max_rounds);
                if (ppr)
                    pprk_snapshot<<<snapshot_blocks(g.V * static_cast<std::uint64_t>(K)), kBlock, 0, s>>>(
                        a.q, 0, K, a.qmask, reinterpret_cast<float *>(kw.val.ptr), kw.rank.as<float>(),
                        kw.snap.as<float>(), rs);
                expandk_light<OffT, KP, Op><<<light_grid, kBlock, 0, s>>>(g.off.as<OffT>(), op, a);
                expandk_light_items<OffT, Op><<<light_grid, kBlock, 0, s>>>(g.off.as<OffT>(), op, a);
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
            lm = o.light_max_degree ? o.light_max_degree : kLaneLightMaxDegree;
            ch = o.heavy_chunk ? o.heavy_chunk : kLaneHeavyChunk;
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
        upload_seeds(g, kw, p, sources, K, kw.val.as<std::uint32_t>(), 0U);
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
        launch_fill<float>(reinterpret_cast<float *>(kw.val.ptr), VK, 0.0F, g.stream);
        launch_fill<float>(kw.rank.as<float>(), VK, 0.0F, g.stream);
        const SeedPlan p = plan_seeds(seeds, K);
        const float r0 = 1.0F - damping;
        std::uint32_t r0_bits;
        std::memcpy(&r0_bits, &r0, 4);
        upload_seeds(g, kw, p, seeds, K, kw.val.as<std::uint32_t>(), r0_bits);
        PprKOp op;
        op.col = g.col.as<std::uint32_t>();
        op.res = reinterpret_cast<float *>(kw.val.ptr);
        op.snap = kw.snap.as<float>();
        op.K = K;
        op.damping = damping;
        op.eps = eps;
        dispatch_off(g, kw, op, K, o, st, lm, ch, true);

        // finalize: x = rank + res, column sums, normalise
        const std::uint32_t KP = static_cast<std::uint32_t>(pow2_at_least(K));
        const std::uint64_t rows_per = kBlock / KP;
        std::uint64_t nb64 = (g.V + rows_per - 1) / rows_per;
        const unsigned nb = static_cast<unsigned>(nb64 < 1024 ? (nb64 ? nb64 : 1) : 1024);
        double *partial = kw.partial.as<double>();
        pprk_fold<<<nb, kBlock, 0, g.stream>>>(kw.rank.as<float>(), op.res, g.V, K, KP, partial);
        hip_check(hipGetLastError(), "launch pprk_fold");
        double *total = partial + static_cast<std::uint64_t>(1024) * 64;
        pprk_total<<<1, 64, 0, g.stream>>>(partial, nb, K, total);
        hip_check(hipGetLastError(), "launch pprk_total");
        float *dst = o.out_on_device ? out : kw.snap.as<float>();
        std::uint64_t b = (VK + kBlock - 1) / kBlock;
        pprk_normalize<<<static_cast<unsigned>(b < 65536 ? b : 65536), kBlock, 0, g.stream>>>(
            kw.rank.as<float>(), VK, K, total, dst);
        hip_check(hipGetLastError(), "launch pprk_normalize");
        if (!o.out_on_device)
            hip_check(hipMemcpyAsync(out, dst, VK * 4, hipMemcpyDeviceToHost, g.stream), "copy ppr");
        hip_check(hipStreamSynchronize(g.stream), "ppr output");
    }
} // namespace hyt
