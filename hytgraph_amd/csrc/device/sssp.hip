// Frontier-relaxation SSSP (single source) on gfx950.
//
// Semantics follow the original synchronous Bellman-Ford
// (src/algorithms/sssp.cpp:109-234, cuda/sssp.cu:100-154): in round k every
// vertex improved in round k-1 offers dist_{k-1}(u) + w(u,v) to each
// out-neighbour v, v keeps the minimum, and v joins the next frontier when its
// distance dropped. The original pulls over a transposed CSR and scans every
// edge each round; here the frontier pushes over the forward CSR, so a round
// reads only the frontier's out-edges:
//   * dist is u32 (integer weights, or float bit patterns: for values >= +0
//     the IEEE order equals the unsigned order, so one atomicMin serves both);
//   * expand: a plain load of dist[v] filters candidates, survivors issue a
//     non-returning atomicMin and a plain byte store touch[v] = 1 (no
//     returning atomics, no shared counters on the edge path);
//   * compaction: a dense scan of touch[] keeps v when dist[v] < prev[v]
//     (prev = distance at the end of the round v last improved in), so the
//     next frontier is exactly the improved set, and records dist[v] as the
//     snapshot value of the next round; calls that never pull skip touch[]
//     and compare dist with prev on every row instead;
//   * SYNC mode expands with the snapshot values (Jacobi), which reproduces
//     the original iteration count; ASYNC mode reads live distances;
//   * direction: a round whose frontier out-edges are a large share of E runs
//     as a pull over the transpose instead (each destination takes the
//     minimum over in-edges from frontier sources, prev[] being exactly the
//     end-of-previous-round distances); same result, no global atomics.
// Why touch[] cannot miss an improvement: the first atomicMin that lowers
// dist[v] in a round came from a thread whose filter load passed, and that
// thread stored touch[v].
//
// Tolerance > 0 (float weights, SYNC): the original's rule - v takes a
// candidate c only when c + tol < best, best starting at v's distance and
// moving as the in-edges are read in incoming-CSR order - depends on that
// order, so these rounds run as ordered pulls: one thread per destination
// walks its in-edges in the transpose's (source, edge position) order, the
// order of build_incoming_csr (sssp.cpp:21-80), and applies the rule
// sequentially; v joins the next frontier when best + tol < its distance.
#include "ops.hpp"
#include "pull.hpp"

#include <chrono>
#include <cmath>
#include <cstring>
#include <algorithm>
#include <string>
#include <utility>
#include <vector>

namespace hyt
{
    namespace
    {
        template <int WM, bool SYNC>
        struct SsspPullOp
        {
            const std::uint32_t *__restrict__ in_col;
            const std::uint32_t *__restrict__ in_w;
            const std::uint32_t *active;
            const std::uint32_t *prev;
            std::uint32_t *dist;
            std::uint8_t *touch;
            std::uint32_t *overflow;

            __device__ std::uint32_t identity() const { return 0xFFFFFFFFU; }
            __device__ bool edge(std::uint64_t pos, std::uint32_t &c) const
            {
                const std::uint32_t u = in_col[pos];
                if (!bit_set(active, u))
                    return false;
                const std::uint32_t du = SYNC ? prev[u] : dist[u];
                if (Weights<WM>::add(du, Weights<WM>::load(in_w, pos), c))
                    return true;
                atomicOr(overflow, 1U);
                return false;
            }
            __device__ std::uint32_t fold(std::uint32_t a, std::uint32_t b) const { return a < b ? a : b; }
            __device__ void finish(std::uint32_t v, std::uint32_t acc) const
            {
                if (acc < dist[v])
                {
                    dist[v] = acc;
                    touch[v] = 1;
                }
            }
            __device__ void finish_atomic(std::uint32_t v, std::uint32_t acc) const
            {
                if (acc < dist[v])
                {
                    atomicMin(&dist[v], acc);
                    touch[v] = 1;
                }
            }
        };

        __global__ void seed_kernel(std::uint32_t *dist, std::uint8_t *touch, std::uint32_t src, std::uint32_t zero)
        {
            if (threadIdx.x == 0 && blockIdx.x == 0)
            {
                dist[src] = zero;
                if (touch != nullptr)
                    touch[src] = 1;
            }
        }

        // One ordered pull round of the tolerance rule (cuda/sssp.cu:100-154):
        // cur/next are float bit patterns, act/next_act one byte per vertex;
        // ctr->next_count counts the vertices of the next frontier.
        template <class OffT, bool WEIGHTED>
        __global__ __launch_bounds__(kBlock) void sssp_tol_round(const OffT *__restrict__ in_off,
                                                                 const std::uint32_t *__restrict__ in_col,
                                                                 const std::uint32_t *__restrict__ in_w,
                                                                 const std::uint32_t *cur, const std::uint8_t *act,
                                                                 std::uint32_t *next, std::uint8_t *next_act,
                                                                 std::uint64_t V, float tol, Counters *ctr)
        {
            __shared__ std::uint32_t s_n[kWavesPerBlock];
            std::uint32_t n = 0;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V;
                 v += stride)
            {
                const float dv = __uint_as_float(cur[v]);
                float best = dv;
                const std::uint64_t end = in_off[v + 1];
                for (std::uint64_t pos = in_off[v]; pos < end; ++pos)
                {
                    const std::uint32_t u = in_col[pos];
                    if (act[u] == 0)
                        continue;
                    const std::uint32_t du = cur[u];
                    if ((du & 0x7F800000U) == 0x7F800000U) // not finite
                        continue;
                    const float c = __fadd_rn(__uint_as_float(du), WEIGHTED ? __uint_as_float(in_w[pos]) : 1.0F);
                    if (__fadd_rn(c, tol) < best)
                        best = c;
                }
                next[v] = __float_as_uint(best);
                const bool improved = __fadd_rn(best, tol) < dv;
                next_act[v] = improved ? 1 : 0;
                n += improved ? 1U : 0U;
            }
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1)
                n += __shfl_xor(n, d, kWave);
            if (lane_id() == 0)
                s_n[threadIdx.x / kWave] = n;
            __syncthreads();
            if (threadIdx.x == 0)
            {
                std::uint32_t t = 0;
#pragma unroll
                for (int w = 0; w < kWavesPerBlock; ++w)
                    t += s_n[w];
                if (t)
                    atomicAdd(&ctr->next_count, t);
            }
        }

        __global__ void tol_seed(std::uint32_t *dist, std::uint8_t *act, std::uint32_t src)
        {
            if (threadIdx.x == 0 && blockIdx.x == 0)
            {
                dist[src] = 0U; // +0.0f
                act[src] = 1;
            }
        }

        // The original's SYNC rounds with tolerance > 0 (sssp.cpp:140-232):
        // iterations = rounds run, converged on the first round that changes
        // nothing; max_rounds 0 means V.
        template <class OffT>
        void run_tolerance(DeviceGraph &g, std::uint32_t src, const hyt_run_options &o, std::uint32_t light_max,
                           std::uint32_t chunk, hyt_run_stats &st)
        {
            Workspace &ws = g.ws;
            hipStream_t s = g.stream;
            Counters *d_ctr = ws.ctr.as<Counters>();
            const std::uint64_t Vp = g.V + kPad;
            DeviceBuffer act_a(Vp, "hipMalloc(tolerance frontier)"), act_b(Vp, "hipMalloc(tolerance frontier)");
            std::uint32_t *cur = ws.dist.as<std::uint32_t>(), *next = ws.prev.as<std::uint32_t>();
            std::uint8_t *act = act_a.as<std::uint8_t>(), *next_act = act_b.as<std::uint8_t>();
            launch_fill<std::uint32_t>(cur, g.V, 0x7F800000U, s);
            launch_fill<std::uint8_t>(act, g.V, 0, s);
            tol_seed<<<1, 64, 0, s>>>(cur, act, src);
            hip_check(hipGetLastError(), "launch tol_seed");
            if (g.E > 0)
                ensure_transpose(g, light_max, chunk);
            const std::uint64_t max_rounds = o.max_rounds ? o.max_rounds : g.V;
            const std::uint64_t b = (g.V + kBlock - 1) / kBlock;
            const unsigned grid = static_cast<unsigned>(b < 8192 ? (b ? b : 1) : 8192);
            Counters *h_ctr = g.host_ctr->as<Counters>();
            std::uint64_t frontier = 1;
            st.converged = 0;
            for (std::uint64_t round = 0; round < max_rounds; ++round)
            {
                if (g.E == 0)
                {
                    // no in-edges: the round keeps every distance and changes nothing
                    st.rounds = round + 1;
                    st.converged = 1;
                    break;
                }
                clear_round_counters(d_ctr, s);
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev0, s), "hipEventRecord");
                if (g.weighted)
                    sssp_tol_round<OffT, true><<<grid, kBlock, 0, s>>>(g.in_off.as<OffT>(), g.in_col.as<std::uint32_t>(),
                                                                       g.in_w.as<std::uint32_t>(), cur, act, next,
                                                                       next_act, g.V, o.tolerance, d_ctr);
                else
                    sssp_tol_round<OffT, false><<<grid, kBlock, 0, s>>>(g.in_off.as<OffT>(), g.in_col.as<std::uint32_t>(),
                                                                        nullptr, cur, act, next, next_act, g.V,
                                                                        o.tolerance, d_ctr);
                hip_check(hipGetLastError(), "launch sssp_tol_round");
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev1, s), "hipEventRecord");
                hip_check(hipMemcpyAsync(h_ctr, d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, s), "read counters");
                hip_check(hipStreamSynchronize(s), "sssp tolerance round");
                if (o.time_kernels)
                {
                    float ms = 0;
                    hip_check(hipEventElapsedTime(&ms, g.ev0, g.ev1), "hipEventElapsedTime");
                    st.expand_ms += ms;
                    st.expand_launches += 1;
                }
                std::swap(cur, next);
                std::swap(act, next_act);
                st.rounds = round + 1;
                st.frontier_total += frontier;
                st.edges_total += g.E; // every in-edge is read
                st.pull_rounds += 1;
                st.pull_edges += g.E;
                frontier = h_ctr->next_count;
                st.updates_total += frontier;
                if (frontier == 0)
                {
                    st.converged = 1;
                    break;
                }
            }
            // the distances of the last round: callers read ws.dist
            if (cur != ws.dist.as<std::uint32_t>())
                hip_check(hipMemcpyAsync(ws.dist.ptr, cur, g.V * 4, hipMemcpyDeviceToDevice, s), "copy distances");
        }

        // ------------------------------------------------ device-driven rounds
        // Push-only rounds (SYNC or ASYNC, no buckets) run from a captured HIP
        // graph of kGraphRounds rounds, replayed until a round finds the queue
        // empty: round_begin moves the last compaction's count into rc->F and
        // clears the round counters (a kernel, no memset node), the expand
        // kernels read F on the device and loop over tiles, the compaction is
        // gated on F, round_end adds the round to the totals. The host reads
        // rc once per replay instead of the counters once per round.
        constexpr int kGraphRounds = 4;
        static_assert(kGraphRounds <= kMaxGraphRounds, "round stamps");
        constexpr unsigned kDeviceLightGrid = 16384; // tiles of 256 rows per sweep

        // hub_bits[t]: bit j = row 16t + j has more than light_max edges;
        // *count += the rows marked.
        template <class OffT>
        __global__ __launch_bounds__(kBlock) void hub_bits_kernel(const OffT *off, std::uint64_t V,
                                                                  std::uint32_t light_max, std::uint16_t *bits,
                                                                  unsigned long long *count)
        {
            const std::uint64_t t = static_cast<std::uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
            const std::uint64_t r0 = t * 16;
            std::uint32_t m = 0;
            if (r0 < V)
            {
                for (std::uint32_t j = 0; j < 16 && r0 + j < V; ++j)
                    if (static_cast<std::uint64_t>(off[r0 + j + 1] - off[r0 + j]) > light_max)
                        m |= 1U << j;
                bits[t] = static_cast<std::uint16_t>(m);
            }
            unsigned long long n = static_cast<unsigned long long>(__popc(m));
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1)
                n += __shfl_xor(n, d, kWave);
            if (lane_id() == 0 && n != 0ULL)
                atomicAdd(count, n);
        }

        // The hub list of a queue the compaction did not list (the first
        // round of a call): the marked rows among q[0, F), appended at *count.
        __global__ __launch_bounds__(kBlock) void hub_filter(const std::uint32_t *q, std::uint32_t F,
                                                             const std::uint16_t *bits, std::uint32_t *list,
                                                             std::uint32_t *count)
        {
            const std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
            const int lane = lane_id();
            std::uint32_t u = 0;
            bool hub = false;
            if (i < F)
            {
                u = q[i];
                hub = ((bits[u >> 4] >> (u & 15U)) & 1U) != 0U;
            }
            const unsigned long long m = __ballot(hub);
            if (m == 0ULL)
                return; // uniform across the wavefront
            std::uint32_t base = 0;
            if (lane == 0)
                base = atomicAdd(count, static_cast<std::uint32_t>(__popcll(m)));
            base = __shfl(base, 0, kWave);
            if (hub)
                list[base + static_cast<std::uint32_t>(__popcll(m & ((1ULL << lane) - 1ULL)))] = u;
        }

        __global__ void rc_reset(RoundCtl *rc)
        {
            if (threadIdx.x == 0 && blockIdx.x == 0)
                *rc = RoundCtl{};
        }

        // t_start (timed runs): the device wall clock when the round's
        // control is done, the start of its expand kernels
        __global__ void round_begin(Counters *ctr, RoundCtl *rc, std::uint64_t max_rounds, std::uint32_t heavy_delta,
                                    bool float_dist, std::uint64_t *t_start)
        {
            if (threadIdx.x != 0 || blockIdx.x != 0)
                return;
            std::uint32_t F = ctr->next_count;
            if (F != 0 && rc->rounds >= max_rounds)
            {
                // the round limit: the counters keep the queue for a resumed run
                rc->truncated = 1;
                rc->F = 0;
                rc->H = 0;
                rc->done = 1;
                return;
            }
            rc->F = F;
            // the rows above light_max among them, listed by the compaction
            rc->H = F != 0 ? ctr->next_hubs : 0U;
            ctr->next_hubs = 0;
            rc->heavy_bound = heavy_bound_of(ctr->kept_min(), heavy_delta, float_dist);
            if (F != 0)
            {
                rc->rounds += 1;
                rc->frontier_total += F;
            }
            else
                rc->done = 1;
            ctr->next_count = 0;
            ctr->heavy_count = 0;
            ctr->edges = 0;
            ctr->frontier_edges = 0;
            ctr->pending = 0;
            ctr->pending_inv_max = 0;
            for (int i = 0; i < kEdgeSlots; ++i)
            {
                ctr->slot[i].v = ctr->slot[i].h = 0;
                ctr->slot[i].kinv = 0;
            }
            if (t_start != nullptr)
                *t_start = static_cast<std::uint64_t>(wall_clock64());
        }

        __global__ void round_end(const Counters *ctr, RoundCtl *rc)
        {
            if (threadIdx.x != 0 || blockIdx.x != 0 || rc->F == 0)
                return;
            rc->edges_total += ctr->all_edges();
            rc->heavy_edges_total += ctr->heavy_edges();
            rc->updates_total += ctr->next_count;
        }

        // The instantiated round graph of a device graph, kept across calls
        // (its nodes name the workspace arrays and the split, not the source);
        // rebuilt when any of them changes.
        struct RoundGraph
        {
            std::vector<std::uint64_t> key;
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            ~RoundGraph()
            {
                if (exec)
                    (void)hipGraphExecDestroy(exec);
                if (graph)
                    (void)hipGraphDestroy(graph);
            }
        };

        struct Launch
        {
            DeviceGraph &g;
            const hyt_run_options &o;
            std::uint32_t light_max, chunk;
            SsspBuffers b;
            // resume = continue a run that stopped at its round limit (the
            // arrays and counters hold its state); limit = rounds of this call
            // (0: the options' max_rounds, else V)
            bool resume = false;
            std::uint64_t limit = 0;
        };

        // Compaction into the queue; SYNC rounds also take the snapshot values,
        // and rounds that may pull the frontier bitmap and out-edge sum (ASYNC
        // push-only rounds read none of them). Returns the counters after it.
        template <class OffT>
        Counters compact_and_read(DeviceGraph &g, const SsspBuffers &b, const SsspSel &sel, bool values,
                                  bool direction)
        {
            Counters *d_ctr = b.ctr;
            compact_rows<std::uint32_t, OffT, SsspSel><<<compact_blocks(g.V), kBlock, 0, g.stream>>>(
                g.V, sel, b.q, values ? b.qd : nullptr, d_ctr, direction ? g.off.as<OffT>() : nullptr,
                direction ? b.active : nullptr);
            hip_check(hipGetLastError(), "launch sssp compaction");
            Counters *h_ctr = g.host_ctr->as<Counters>();
            hip_check(hipMemcpyAsync(h_ctr, d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, g.stream), "read counters");
            hip_check(hipStreamSynchronize(g.stream), "sssp round");
            return *h_ctr;
        }

        template <class OffT, int WM, bool SYNC>
        void run_device_driven(Launch &L, const SsspSel &sel, const SsspOp<WM, SYNC> &op, ExpandArgs a,
                               std::uint64_t max_rounds, std::uint32_t F0, hyt_run_stats &st)
        {
            DeviceGraph &g = L.g;
            const SsspBuffers &B = L.b;
            hipStream_t s = g.stream;
            const bool timed = L.o.time_kernels != 0;
            const std::uint32_t heavy_delta = SYNC ? 0U : heavy_delta_of(L.o.heavy_delta);
            const std::uint32_t wait_degree = wait_degree_of(L.o.wait_degree, L.light_max);
            constexpr bool kFloatDist = WM == kWF32 || WM == kWUnitF32;
            a.F_dev = &B.rc->F;
            const unsigned light_grid = std::min(blocks_for(g.V), kDeviceLightGrid);
            const unsigned nb = compact_blocks(g.V);
            const GraphRows<OffT> rows = whole_graph(g.off.as<OffT>(), B.q);
            // ASYNC with the hub bitmap: the chunk-queuing pass reads the
            // compaction's list of rows above light_max instead of the whole
            // queue (the light pass still reads the queue and skips them)
            const bool hub_list = !SYNC && B.hub_bits != nullptr;
            const GraphRows<OffT> hub_rows = whole_graph(g.off.as<OffT>(), B.hub_list);
            const unsigned hub_grid = std::min(blocks_for(B.hub_rows), kDeviceLightGrid);
            const HubOut hubs = hub_list ? HubOut{B.hub_bits, B.hub_list, &B.ctr->next_hubs} : HubOut{};
            const std::vector<std::uint64_t> key{
                static_cast<std::uint64_t>(sizeof(OffT)), static_cast<std::uint64_t>(WM), SYNC ? 1ULL : 0ULL,
                timed ? 1ULL : 0ULL, max_rounds, L.light_max, L.chunk, B.heavy_cap, heavy_delta, wait_degree,
                reinterpret_cast<std::uint64_t>(B.dist), reinterpret_cast<std::uint64_t>(B.prev),
                reinterpret_cast<std::uint64_t>(B.q), reinterpret_cast<std::uint64_t>(B.qd),
                reinterpret_cast<std::uint64_t>(B.heavy), reinterpret_cast<std::uint64_t>(B.ctr),
                reinterpret_cast<std::uint64_t>(B.rc), g.off.ptr ? reinterpret_cast<std::uint64_t>(g.off.ptr) : 0,
                reinterpret_cast<std::uint64_t>(B.hub_bits), reinterpret_cast<std::uint64_t>(B.hub_list), B.hub_rows,
                reinterpret_cast<std::uint64_t>(g.col.ptr), reinterpret_cast<std::uint64_t>(g.w.ptr)};
            auto *cached = static_cast<RoundGraph *>(g.round_graph.get());
            if (cached == nullptr || cached->key != key)
            {
                g.round_graph.reset();
                auto rg = std::make_shared<RoundGraph>();
                rg->key = key;
                hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin round capture");
                for (int r = 0; r < kGraphRounds; ++r)
                {
                    round_begin<<<1, 64, 0, s>>>(B.ctr, B.rc, max_rounds, heavy_delta, kFloatDist,
                                                 timed ? &B.rc->t[2 * r] : nullptr);
                    if (SYNC)
                    {
                        expand_light<GraphRows<OffT>, SsspOp<WM, SYNC>><<<light_grid, kBlock, 0, s>>>(rows, op, a);
                        expand_heavy<SsspOp<WM, SYNC>><<<kSingleHeavyGrid, kBlock, 0, s>>>(op, a);
                    }
                    else
                    {
                        ExpandArgs a1 = a, a2 = a;
                        a1.pass = 1;
                        a1.heavy_bound_dev = &B.rc->heavy_bound;
                        a2.pass = 2;
                        a2.heavy_bound_dev = &B.rc->heavy_bound;
                        a2.wait_degree = wait_degree;
                        if (hub_list)
                        {
                            ExpandArgs ah = a1;
                            ah.F_dev = &B.rc->H;
                            expand_light<GraphRows<OffT>, SsspOp<WM, SYNC>><<<hub_grid, kBlock, 0, s>>>(hub_rows, op, ah);
                        }
                        else
                            expand_light<GraphRows<OffT>, SsspOp<WM, SYNC>><<<light_grid, kBlock, 0, s>>>(rows, op, a1);
                        expand_heavy<SsspOp<WM, SYNC>><<<kSingleHeavyGrid, kBlock, 0, s>>>(op, a1);
                        expand_light<GraphRows<OffT>, SsspOp<WM, SYNC>><<<light_grid, kBlock, 0, s>>>(rows, op, a2);
                    }
                    // the compaction's first workgroup stamps the end of the expand kernels
                    compact_rows<std::uint32_t, OffT, SsspSel><<<nb, kBlock, 0, s>>>(
                        g.V, sel, B.q, SYNC ? B.qd : nullptr, B.ctr, static_cast<const OffT *>(nullptr),
                        static_cast<std::uint16_t *>(nullptr), static_cast<std::uint32_t *>(nullptr), &B.rc->F, hubs,
                        timed ? &B.rc->t[2 * r + 1] : nullptr);
                    round_end<<<1, 64, 0, s>>>(B.ctr, B.rc);
                }
                const hipError_t launch_err = hipGetLastError();
                const hipError_t cap_err = hipStreamEndCapture(s, &rg->graph);
                hip_check(launch_err, "launch captured round");
                hip_check(cap_err, "end round capture");
                hip_check(hipGraphInstantiate(&rg->exec, rg->graph, nullptr, nullptr, 0), "instantiate round graph");
                g.round_graph = rg;
                cached = rg.get();
            }
            rc_reset<<<1, 64, 0, s>>>(B.rc);
            hip_check(hipGetLastError(), "launch rc_reset");
            if (hub_list)
            {
                // the queue of the first round came from a compaction that
                // listed no rows apart (or from a call that stopped at its
                // round limit): list its rows above light_max again
                launch_fill<std::uint32_t>(&B.ctr->next_hubs, 1, 0U, s);
                hub_filter<<<blocks_for(F0), kBlock, 0, s>>>(B.q, F0, B.hub_bits, B.hub_list, &B.ctr->next_hubs);
                hip_check(hipGetLastError(), "launch hub_filter");
            }
            double clock_khz = 1.0;
            if (timed)
            {
                int khz = 0;
                hip_check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g.device), "wall clock rate");
                require(khz > 0, "device wall clock rate unknown");
                clock_khz = static_cast<double>(khz);
            }
            RoundCtl *h_rc = reinterpret_cast<RoundCtl *>(g.host_ctr->as<unsigned char>() + sizeof(Counters));
            Counters *h_ctr = g.host_ctr->as<Counters>();
            std::uint64_t seen_rounds = 0;
            while (true)
            {
                hip_check(hipGraphLaunch(cached->exec, s), "launch round graph");
                hip_check(hipMemcpyAsync(h_rc, B.rc, sizeof(RoundCtl), hipMemcpyDeviceToHost, s), "read rounds");
                hip_check(hipMemcpyAsync(h_ctr, B.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s), "read counters");
                hip_check(hipStreamSynchronize(s), "sssp rounds");
                if (timed)
                {
                    // rounds of this replay that had work (the rest exit at
                    // once); the stamps bracket their expand kernels
                    const std::uint64_t ran = std::min<std::uint64_t>(h_rc->rounds - seen_rounds, kGraphRounds);
                    for (std::uint64_t r = 0; r < ran; ++r)
                    {
                        st.expand_ms += static_cast<double>(h_rc->t[2 * r + 1] - h_rc->t[2 * r]) / clock_khz;
                        st.expand_launches += 1;
                    }
                }
                seen_rounds = h_rc->rounds;
                if (h_ctr->heavy_overflow)
                    fail(HYT_ERR_INTERNAL, "heavy chunk queue overflow");
                if (h_ctr->overflow)
                {
                    st.overflow = 1;
                    fail(HYT_ERR_OVERFLOW, "u32 distance sum exceeded 2^32-2");
                }
                if (h_rc->done)
                    break;
            }
            st.rounds = h_rc->rounds;
            st.frontier_total += h_rc->frontier_total;
            st.edges_total += h_rc->edges_total;
            st.heavy_edges += h_rc->heavy_edges_total;
            st.updates_total += h_rc->updates_total;
            st.converged = h_rc->truncated ? 0 : 1;
        }

        template <class OffT, int WM, bool SYNC>
        void run_rounds(Launch &L, std::uint32_t src, hyt_run_stats &st)
        {
            DeviceGraph &g = L.g;
            const SsspBuffers &B = L.b;
            hipStream_t s = g.stream;
            Counters *d_ctr = B.ctr;

            if (!L.resume)
            {
                launch_fill<std::uint32_t>(B.dist, g.V, Weights<WM>::kInf, s);
                launch_fill<std::uint32_t>(B.prev, g.V, Weights<WM>::kInf, s);
                clear_counters(d_ctr, s);
            }

            const std::uint32_t dir = L.o.direction;
            const std::uint64_t pct = L.o.pull_percent ? L.o.pull_percent : kDefaultPullPercentSssp;
            const bool may_pull = dir == HYT_DIR_PULL || (dir == HYT_DIR_AUTO && pct <= 100);
            // Push-only calls find improved rows by comparing dist with prev in
            // the compaction instead of a byte store per improvement (random
            // stores that each move a cache line); pull rounds need touch[].
            require(!may_pull || B.touch != nullptr, "pull rounds need the traversal workspace");
            std::uint8_t *touch = may_pull ? B.touch : nullptr;
            if (!L.resume)
            {
                seed_kernel<<<1, 64, 0, s>>>(B.dist, touch, src, 0U);
                hip_check(hipGetLastError(), "launch seed");
            }

            // delta-stepping buckets (ASYNC only): admit improved rows below
            // `bound`; when none is left, move the bound past the nearest
            // held-back distance
            const bool bucketed = !SYNC && L.o.delta > 0;
            auto bound_after = [&](std::uint32_t d) -> std::uint32_t {
                if (WM == kWF32 || WM == kWUnitF32)
                {
                    float f;
                    std::memcpy(&f, &d, 4);
                    f += static_cast<float>(L.o.delta);
                    std::uint32_t b;
                    std::memcpy(&b, &f, 4);
                    return b;
                }
                const std::uint64_t b = static_cast<std::uint64_t>(d) + L.o.delta;
                return b >= 0xFFFFFFFFULL ? 0xFFFFFFFFU : static_cast<std::uint32_t>(b);
            };
            SsspSel sel{touch, B.dist, B.prev,
                        bucketed ? bound_after(0U) : 0xFFFFFFFFU};
            // rows above light_max wait while far above the queue's smallest
            // distance (ASYNC): the compaction records that distance
            const std::uint32_t heavy_delta = SYNC ? 0U : heavy_delta_of(L.o.heavy_delta);
            const std::uint32_t wait_degree = wait_degree_of(L.o.wait_degree, L.light_max);
            constexpr bool kFloatDist = WM == kWF32 || WM == kWUnitF32;
            sel.track_min = heavy_delta > 0;

            if (may_pull && g.E > 0 && g.E < (1ULL << 31))
                ensure_transpose(g, L.light_max, L.chunk);
            // a transpose left by an earlier call is only used when this call may pull
            const bool can_pull = may_pull && g.in_col.ptr != nullptr;
            require(dir != HYT_DIR_PULL || can_pull || g.E == 0, "pull rounds need fewer than 2^31 edges");

            require(!(L.resume && bucketed), "a resumed run has no delta-stepping buckets");
            Counters c0{};
            if (L.resume)
            {
                // the queue and counters of the last compaction of the run
                hip_check(hipMemcpyAsync(g.host_ctr->ptr, d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, s),
                          "read counters");
                hip_check(hipStreamSynchronize(s), "sssp resume");
                c0 = *g.host_ctr->as<Counters>();
            }
            else
                c0 = compact_and_read<OffT>(g, B, sel, SYNC, can_pull);
            std::uint32_t F = c0.next_count;
            std::uint32_t kept_min = c0.kept_min();
            unsigned long long Fe = c0.frontier_edges;
            std::uint32_t pending = c0.pending;
            std::uint32_t pending_min = ~c0.pending_inv_max;
            SsspPullOp<WM, SYNC> pop;
            pop.in_col = g.in_col.as<std::uint32_t>();
            pop.in_w = g.in_w.as<std::uint32_t>();
            pop.active = reinterpret_cast<const std::uint32_t *>(B.active);
            pop.prev = B.prev;
            pop.dist = B.dist;
            pop.touch = B.touch;
            pop.overflow = &d_ctr->overflow;

            SsspOp<WM, SYNC> op;
            op.col = g.col.as<std::uint32_t>();
            op.w = g.w.as<std::uint32_t>();
            op.dist = B.dist;
            op.touch = touch;
            op.qd = B.qd;
            op.overflow = &d_ctr->overflow;
            op.expanded_at = SYNC ? nullptr : B.prev;

            ExpandArgs a;
            a.q = B.q;
            a.ctr = d_ctr;
            a.heavy = B.heavy;
            a.heavy_cap = B.heavy_cap;
            a.light_max = L.light_max;
            a.chunk = L.chunk;

            const std::uint64_t max_rounds = L.limit ? L.limit : (L.o.max_rounds ? L.o.max_rounds : g.V);
            st.converged = 0;
            if (B.rc != nullptr && !bucketed && !can_pull && F > 0)
            {
                run_device_driven<OffT, WM, SYNC>(L, sel, op, a, max_rounds, F, st);
                return;
            }
            for (std::uint64_t round = 0; round < max_rounds && (F > 0 || pending > 0);)
            {
                if (F == 0)
                {
                    // bucket exhausted: advance to the nearest held-back row
                    sel.bound = bound_after(pending_min);
                    clear_round_counters(d_ctr, s);
                    const Counters cb = compact_and_read<OffT>(g, B, sel, SYNC, can_pull);
                    F = cb.next_count;
                    kept_min = cb.kept_min();
                    Fe = cb.frontier_edges;
                    pending = cb.pending;
                    pending_min = ~cb.pending_inv_max;
                    continue;
                }
                ++round;
                a.F = F;
                clear_round_counters(d_ctr, s);
                const bool pull = can_pull && (dir == HYT_DIR_PULL ||
                                               (dir == HYT_DIR_AUTO && Fe * 100ULL >= pct * g.E));
                if (L.o.time_kernels)
                    hip_check(hipEventRecord(g.ev0, s), "hipEventRecord");
                if (pull)
                {
                    pull_light<OffT, SsspPullOp<WM, SYNC>><<<blocks_for(g.V), kBlock, 0, s>>>(
                        g.in_off.as<OffT>(), pop, g.V, L.light_max);
                    hip_check(hipGetLastError(), "launch sssp_pull_light");
                    pull_heavy<SsspPullOp<WM, SYNC>><<<kSingleHeavyGrid, kBlock, 0, s>>>(pop, g.in_heavy.as<HeavyChunk>(),
                                                                                  g.in_heavy_count);
                    hip_check(hipGetLastError(), "launch sssp_pull_heavy");
                    st.pull_rounds += 1;
                    st.pull_active_edges += Fe;
                }
                else if (!SYNC)
                {
                    // ASYNC: the chunks of rows above light_max (the hubs) go
                    // first - pass 1 only queues them, the chunk kernel walks
                    // them - and the light rows follow in pass 2, reading the
                    // distances the hubs just lowered (RMAT 24: 590 M -> 518 M
                    // relaxations per query, 0.565 -> 0.548 ms per round)
                    ExpandArgs a1 = a, a2 = a;
                    a1.pass = 1;
                    a1.heavy_bound = heavy_bound_of(kept_min, heavy_delta, kFloatDist);
                    a2.pass = 2;
                    a2.heavy_bound = a1.heavy_bound;
                    a2.wait_degree = wait_degree;
                    expand_light<GraphRows<OffT>, SsspOp<WM, SYNC>><<<blocks_for(F), kBlock, 0, s>>>(
                        whole_graph(g.off.as<OffT>(), a.q), op, a1);
                    hip_check(hipGetLastError(), "launch sssp_expand_light");
                    expand_heavy<SsspOp<WM, SYNC>><<<kSingleHeavyGrid, kBlock, 0, s>>>(op, a1);
                    hip_check(hipGetLastError(), "launch sssp_expand_heavy");
                    expand_light<GraphRows<OffT>, SsspOp<WM, SYNC>><<<blocks_for(F), kBlock, 0, s>>>(
                        whole_graph(g.off.as<OffT>(), a.q), op, a2);
                    hip_check(hipGetLastError(), "launch sssp_expand_light");
                }
                else
                {
                    expand_light<GraphRows<OffT>, SsspOp<WM, SYNC>><<<blocks_for(F), kBlock, 0, s>>>(
                        whole_graph(g.off.as<OffT>(), a.q), op, a);
                    hip_check(hipGetLastError(), "launch sssp_expand_light");
                    expand_heavy<SsspOp<WM, SYNC>><<<kSingleHeavyGrid, kBlock, 0, s>>>(op, a);
                    hip_check(hipGetLastError(), "launch sssp_expand_heavy");
                }
                if (L.o.time_kernels)
                    hip_check(hipEventRecord(g.ev1, s), "hipEventRecord");
                const Counters c = compact_and_read<OffT>(g, B, sel, SYNC, can_pull);
                if (L.o.time_kernels)
                {
                    float ms = 0;
                    hip_check(hipEventElapsedTime(&ms, g.ev0, g.ev1), "hipEventElapsedTime");
                    st.expand_ms += ms;
                    st.expand_launches += 1;
                }
                if (c.heavy_overflow)
                    fail(HYT_ERR_INTERNAL, "heavy chunk queue overflow");
                st.frontier_total += F;
                // a pull round scans every in-edge (its kernels count none)
                const std::uint64_t round_edges = pull ? g.E : c.all_edges();
                st.edges_total += round_edges;
                if (!pull)
                    st.heavy_edges += c.heavy_edges();
                if (pull)
                {
                    st.pull_edges += round_edges;
                    st.pull_frontier += F;
                }
                st.updates_total += c.next_count;
                st.rounds = round;
                if (c.overflow)
                {
                    st.overflow = 1;
                    fail(HYT_ERR_OVERFLOW, "u32 distance sum exceeded 2^32-2");
                }
                F = c.next_count;
                kept_min = c.kept_min();
                Fe = c.frontier_edges;
                pending = c.pending;
                pending_min = ~c.pending_inv_max;
                if (F == 0 && pending == 0)
                    st.converged = 1;
            }
            if (pending > 0 && touch != nullptr)
            {
                // stopped by max_rounds with held-back rows: touch[] must be
                // all-zero for the next call (touch-free calls keep none)
                hip_check(hipMemsetAsync(touch, 0, g.V + kPad, s), "clear touch");
            }
        }

        template <class OffT, int WM>
        void dispatch_mode(Launch &L, std::uint32_t src, hyt_run_stats &st)
        {
            if (L.o.mode == HYT_SSSP_ASYNC)
                run_rounds<OffT, WM, false>(L, src, st);
            else
                run_rounds<OffT, WM, true>(L, src, st);
        }

        template <class OffT>
        void dispatch_weights(Launch &L, std::uint32_t src, hyt_run_stats &st)
        {
            switch (L.g.weight_mode())
            {
            case kWU32: dispatch_mode<OffT, kWU32>(L, src, st); break;
            case kWF32: dispatch_mode<OffT, kWF32>(L, src, st); break;
            case kWUnitU32: dispatch_mode<OffT, kWUnitU32>(L, src, st); break;
            default: dispatch_mode<OffT, kWUnitF32>(L, src, st); break;
            }
        }
    } // namespace

    void sssp_split(const DeviceGraph &g, const hyt_run_options &o, std::uint32_t &light_max, std::uint32_t &chunk)
    {
        std::uint32_t dl = 0, dc = 0;
        default_sssp_split(g.E, dl, dc);
        light_max = o.light_max_degree ? o.light_max_degree : dl;
        chunk = o.heavy_chunk ? o.heavy_chunk : dc;
        require(light_max <= (1U << 20), "light_max_degree must be <= 2^20");
        require(chunk >= 64 && chunk <= (1U << 24), "heavy_chunk must be in [64, 2^24]");
    }

    namespace
    {
        // The hub bitmap of the workspace at a split, and a list long
        // enough for every marked row (built once per graph and light_max).
        template <class OffT>
        void build_hub_rows(DeviceGraph &g, std::uint32_t light_max)
        {
            Workspace &ws = g.ws;
            ws.hub_list.release();
            ws.hub_bits.release();
            ws.hub_rows = 0;
            const std::uint64_t words = (g.V + 15) / 16;
            ws.hub_bits.alloc((words + 1) * sizeof(std::uint16_t), "hipMalloc(hub bitmap)");
            DeviceBuffer count(sizeof(unsigned long long), "hipMalloc(hub count)");
            launch_fill<unsigned long long>(count.as<unsigned long long>(), 1, 0ULL, g.stream);
            if (words > 0)
            {
                hub_bits_kernel<OffT><<<blocks_for(words), kBlock, 0, g.stream>>>(
                    g.off.as<OffT>(), g.V, light_max, ws.hub_bits.as<std::uint16_t>(), count.as<unsigned long long>());
                hip_check(hipGetLastError(), "launch hub_bits");
            }
            unsigned long long n = 0;
            hip_check(hipMemcpyAsync(&n, count.ptr, sizeof(n), hipMemcpyDeviceToHost, g.stream), "read hub count");
            hip_check(hipStreamSynchronize(g.stream), "hub bitmap");
            ws.hub_list.alloc((n + 1) * sizeof(std::uint32_t), "hipMalloc(hub list)");
            ws.hub_rows = n;
            ws.hub_light_max = light_max;
        }

        // Device-driven ASYNC calls list the rows above light_max apart
        // (SsspBuffers::hub_bits); other calls leave the fields null.
        void attach_hub_rows(DeviceGraph &g, const hyt_run_options &o, std::uint32_t light_max, SsspBuffers &B)
        {
            if (o.mode == HYT_SSSP_SYNC || B.rc == nullptr)
                return;
            Workspace &ws = g.ws;
            if (ws.hub_bits.ptr == nullptr || ws.hub_list.ptr == nullptr || ws.hub_light_max != light_max)
            {
                if (g.off64)
                    build_hub_rows<std::uint64_t>(g, light_max);
                else
                    build_hub_rows<std::uint32_t>(g, light_max);
            }
            B.hub_bits = ws.hub_bits.as<std::uint16_t>();
            B.hub_list = ws.hub_list.as<std::uint32_t>();
            B.hub_rows = ws.hub_rows;
        }
    } // namespace

    void sssp_rounds(DeviceGraph &g, std::uint32_t src, const SsspBuffers &b, const hyt_run_options &o,
                     std::uint32_t light_max, std::uint32_t chunk, hyt_run_stats &st)
    {
        Launch L{g, o, light_max, chunk, b};
        if (g.off64)
            dispatch_weights<std::uint64_t>(L, src, st);
        else
            dispatch_weights<std::uint32_t>(L, src, st);
    }

    void sssp_steps(DeviceGraph &g, std::uint32_t src, const hyt_run_options &o, bool resume, std::uint64_t rounds,
                    hyt_run_stats &st)
    {
        require(o.tolerance == 0.0F && o.delta == 0, "round-by-round SSSP runs have no tolerance or buckets");
        std::uint32_t light_max = 0, chunk = 0;
        sssp_split(g, o, light_max, chunk);
        ensure_workspace(g, light_max, chunk);
        Workspace &ws = g.ws;
        SsspBuffers B{ws.dist.as<std::uint32_t>(), ws.prev.as<std::uint32_t>(), ws.q.as<std::uint32_t>(),
                      ws.qd.as<std::uint32_t>(),   ws.touch.as<std::uint8_t>(), ws.active.as<std::uint16_t>(),
                      ws.heavy.as<HeavyChunk>(),   ws.heavy_cap,                ws.ctr.as<Counters>(),
                      o.host_rounds ? nullptr : ws.rc.as<RoundCtl>()};
        attach_hub_rows(g, o, light_max, B);
        Launch L{g, o, light_max, chunk, B, resume, rounds};
        if (g.off64)
            dispatch_weights<std::uint64_t>(L, src, st);
        else
            dispatch_weights<std::uint32_t>(L, src, st);
    }

    void sssp_single(DeviceGraph &g, std::uint32_t src, std::uint32_t *dist_out, const hyt_run_options &o,
                     hyt_run_stats &st)
    {
        std::uint32_t light_max = 0, chunk = 0;
        sssp_split(g, o, light_max, chunk);
        ensure_workspace(g, light_max, chunk);
        Workspace &ws = g.ws;
        SsspBuffers B{ws.dist.as<std::uint32_t>(), ws.prev.as<std::uint32_t>(), ws.q.as<std::uint32_t>(),
                      ws.qd.as<std::uint32_t>(),   ws.touch.as<std::uint8_t>(), ws.active.as<std::uint16_t>(),
                      ws.heavy.as<HeavyChunk>(),   ws.heavy_cap,                ws.ctr.as<Counters>(),
                      o.host_rounds ? nullptr : ws.rc.as<RoundCtl>()};
        if (o.tolerance != 0.0F)
        {
            require(o.tolerance > 0.0F && std::isfinite(o.tolerance), "SSSP tolerance must be finite and non-negative");
            require(g.weights_f32, "SSSP tolerance > 0 needs float weights (HYT_GRAPH_WEIGHTS_F32)");
            require(o.mode == HYT_SSSP_SYNC && o.delta == 0 && o.direction != HYT_DIR_PUSH,
                    "SSSP tolerance > 0 runs SYNC pull rounds (no ASYNC mode, buckets or push direction)");
            require(g.E < (1ULL << 31), "SSSP tolerance > 0 needs fewer than 2^31 edges (transpose)");
            if (g.off64)
                run_tolerance<std::uint64_t>(g, src, o, light_max, chunk, st);
            else
                run_tolerance<std::uint32_t>(g, src, o, light_max, chunk, st);
        }
        else if (o.persistent)
        {
            require(o.delta == 0 && o.direction != HYT_DIR_PULL,
                    "persistent SSSP runs push rounds without buckets (direction PUSH or AUTO, delta 0)");
            sssp_persistent_rounds(g, src, o, light_max, chunk, st);
        }
        else
        {
            try
            {
                attach_hub_rows(g, o, light_max, B);
                sssp_rounds(g, src, B, o, light_max, chunk, st);
            }
            catch (...)
            {
                // a failed round may leave touch[] marks behind: restore all-zero
                (void)hipMemsetAsync(g.ws.touch.ptr, 0, g.V + kPad, g.stream);
                (void)hipStreamSynchronize(g.stream);
                throw;
            }
        }
        hip_check(hipMemcpyAsync(dist_out, g.ws.dist.ptr, g.V * 4,
                                 o.out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, g.stream),
                  "copy distances");
        hip_check(hipStreamSynchronize(g.stream), "sssp output");
    }

    void sssp_batched(DeviceGraph &g, const std::uint32_t *sources, std::uint32_t K, std::uint32_t *dist_out,
                      const hyt_run_options &o, hyt_run_stats &st);
} // namespace hyt

extern "C" int hyt_sssp(hyt_graph_t g, const uint32_t *sources, uint32_t num_sources, uint32_t *dist,
                        const hyt_run_options *opts, hyt_run_stats *stats)
{
    return hyt::guarded("hyt_sssp", [&] {
        using namespace hyt;
        require(g != nullptr, "null graph");
        require(sources != nullptr && dist != nullptr, "null sources or dist");
        require(num_sources >= 1 && num_sources <= 64, "num_sources must be in [1, 64]");
        for (std::uint32_t k = 0; k < num_sources; ++k)
            require(sources[k] < g->g.V, "SSSP source vertex is out of range");
        hyt_run_options o{};
        if (opts)
            o = *opts;
        require(o.mode == HYT_SSSP_SYNC || o.mode == HYT_SSSP_ASYNC, "unknown SSSP mode");
        require(o.tolerance == 0.0F || num_sources == 1, "SSSP tolerance > 0 runs one source per call");
        hyt_run_stats st{};
        hip_check(hipSetDevice(g->g.device), "hipSetDevice");
        const auto t0 = std::chrono::steady_clock::now();
        try
        {
            if (num_sources == 1)
                sssp_single(g->g, sources[0], dist, o, st);
            else
                sssp_batched(g->g, sources, num_sources, dist, o, st);
        }
        catch (...)
        {
            if (stats)
                *stats = st;
            throw;
        }
        st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (stats)
            *stats = st;
    });
}
