// Delta-push PageRank / personalized PageRank on gfx950.
//
// The original (src/algorithms/pagerank.cpp:89-163, cuda/pagerank.cu:116-153)
// runs Jacobi power iteration x' = (1-d)/n + d*M/n + d*P^T x with the dangling
// mass M spread uniformly. That fixed point is x = y / |y|_1 where
// y = (1-d) 1 + d P~^T y and P~ drops dangling rows, because uniform dangling
// redistribution only rescales the teleport term. Delta-push computes y
// directly:
//   residual r = (1-d) everywhere (PPR: at the seed), rank y = 0;
//   round: every vertex u with r_u > eps moves r_u into y_u and adds
//          d * r_u / outdeg(u) to the residual of each out-neighbour;
//   stop when no residual exceeds eps; result (y + r) / |y + r|_1.
// The frontier of a round is exactly {u : r_u > eps}: the compaction pass
// scans the residuals, moves the selected ones into rank and records each
// row's share, so which vertices push in a round does not depend on thread
// timing. Residuals and ranks are 64-bit fixed-point words (mass.hpp), so the
// sums themselves do not depend on it either. Push rounds add shares with
// non-returning integer atomics; rounds whose frontier out-edges are a large
// share of E pull instead: each destination sums the shares of its in-edges
// from frontier sources in registers and LDS and adds the total to its
// residual with one store.
#include "ops.hpp"
#include "pull.hpp"

#include <chrono>
#include <cmath>

namespace hyt
{
    namespace
    {
        template <bool WORDS>
        struct PrPullOp
        {
            const std::uint32_t *__restrict__ in_col;
            const ShareT<WORDS> *share;
            unsigned long long *res;

            __device__ unsigned long long identity() const { return 0ULL; }
            __device__ bool edge(std::uint64_t pos, unsigned long long &c) const
            {
                c = word_of(share[in_col[pos]]);
                return c != 0ULL;
            }
            __device__ unsigned long long fold(unsigned long long a, unsigned long long b) const { return a + b; }
            __device__ void finish(std::uint32_t v, unsigned long long acc) const
            {
                if (acc != 0ULL)
                    res[v] += acc;
            }
            __device__ void finish_atomic(std::uint32_t v, unsigned long long acc) const { atomicAdd(&res[v], acc); }
        };

        __global__ void pr_init_all(unsigned long long *res, unsigned long long *rank, std::uint64_t V,
                                    unsigned long long r0)
        {
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; i < V; i += stride)
            {
                res[i] = r0;
                rank[i] = 0ULL;
            }
        }

        __global__ void pr_seed(unsigned long long *res, std::uint32_t s, unsigned long long r0)
        {
            if (threadIdx.x == 0 && blockIdx.x == 0)
                res[s] = r0;
        }

        // Exact per-block sums of x = rank + res and of res (no word sum
        // passes 2^62, mass.hpp); the state is left as it is, so a run can
        // be read between its rounds.
        __global__ __launch_bounds__(kBlock) void pr_fold(const unsigned long long *rank, const unsigned long long *res,
                                                          std::uint64_t V, unsigned long long *partial)
        {
            __shared__ unsigned long long sx[kBlock], sr[kBlock];
            unsigned long long ax = 0, ar = 0;
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; i < V; i += stride)
            {
                const unsigned long long r = res[i];
                ax += rank[i] + r;
                ar += r;
            }
            sx[threadIdx.x] = ax;
            sr[threadIdx.x] = ar;
            __syncthreads();
            for (int s = kBlock / 2; s > 0; s >>= 1)
            {
                if (static_cast<int>(threadIdx.x) < s)
                {
                    sx[threadIdx.x] += sx[threadIdx.x + s];
                    sr[threadIdx.x] += sr[threadIdx.x + s];
                }
                __syncthreads();
            }
            if (threadIdx.x == 0)
            {
                partial[2 * blockIdx.x] = sx[0];
                partial[2 * blockIdx.x + 1] = sr[0];
            }
        }

        // Single workgroup: totals of the partials into partial[2*nblocks .. +1].
        __global__ __launch_bounds__(kBlock) void pr_total(unsigned long long *partial, unsigned nblocks)
        {
            __shared__ unsigned long long sx[kBlock], sr[kBlock];
            unsigned long long ax = 0, ar = 0;
            for (unsigned b = threadIdx.x; b < nblocks; b += blockDim.x)
            {
                ax += partial[2 * b];
                ar += partial[2 * b + 1];
            }
            sx[threadIdx.x] = ax;
            sr[threadIdx.x] = ar;
            __syncthreads();
            for (int s = kBlock / 2; s > 0; s >>= 1)
            {
                if (static_cast<int>(threadIdx.x) < s)
                {
                    sx[threadIdx.x] += sx[threadIdx.x + s];
                    sr[threadIdx.x] += sr[threadIdx.x + s];
                }
                __syncthreads();
            }
            if (threadIdx.x == 0)
            {
                partial[2 * nblocks] = sx[0];
                partial[2 * nblocks + 1] = sr[0];
            }
        }

        __global__ void pr_normalize(const unsigned long long *rank, const unsigned long long *res, std::uint64_t V,
                                     const unsigned long long *total, float *out)
        {
            const double t = static_cast<double>(total[0]);
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; i < V; i += stride)
                out[i] = static_cast<float>(t > 0.0 ? static_cast<double>(rank[i] + res[i]) / t : 0.0);
        }

        constexpr unsigned kFoldBlocks = 1024; // partial holds 2*kFoldBlocks+2 words

        template <class OffT, bool WORDS>
        Counters compact_and_read(DeviceGraph &g, const PrSel<OffT, WORDS> &sel)
        {
            Workspace &ws = g.ws;
            Counters *d_ctr = ws.ctr.as<Counters>();
            compact_rows<std::uint32_t, OffT, PrSel<OffT, WORDS>><<<compact_blocks(g.V), kBlock, 0, g.stream>>>(
                g.V, sel, ws.q.as<std::uint32_t>(), ws.qd.as<std::uint32_t>(), d_ctr, g.off.as<OffT>(),
                static_cast<std::uint16_t *>(nullptr));
            hip_check(hipGetLastError(), "launch pagerank compaction");
            Counters *h_ctr = g.host_ctr->as<Counters>();
            hip_check(hipMemcpyAsync(h_ctr, d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, g.stream), "read counters");
            hip_check(hipStreamSynchronize(g.stream), "pagerank round");
            return *h_ctr;
        }

        // WORDS: PPR (share words by vertex); global PageRank: float shares
        template <class OffT, bool WORDS>
        void run_push(DeviceGraph &g, const hyt_run_options &o, float damping, const MassScale &m,
                      hyt_run_stats &st, std::uint32_t light_max, std::uint32_t chunk, bool resume = false,
                      std::uint64_t limit = 0)
        {
            Workspace &ws = g.ws;
            hipStream_t s = g.stream;
            Counters *d_ctr = ws.ctr.as<Counters>();
            const std::uint64_t max_rounds = limit ? limit : (o.max_rounds ? o.max_rounds : 100000);

            PrSel<OffT, WORDS> sel{ws.res.as<unsigned long long>(), ws.rank.as<unsigned long long>(),
                                   ws.share.as<ShareT<WORDS>>(), g.off.as<OffT>(), damping, m.eps};
            PrOp<WORDS> op;
            op.col = g.col.as<std::uint32_t>();
            op.res = ws.res.as<unsigned long long>();
            op.share = ws.share.as<unsigned long long>();
            op.qd = ws.qd.as<std::uint32_t>();

            ExpandArgs a;
            a.q = ws.q.as<std::uint32_t>();
            a.ctr = d_ctr;
            a.heavy = ws.heavy.as<HeavyChunk>();
            a.heavy_cap = ws.heavy_cap;
            a.light_max = light_max;
            a.chunk = chunk;

            const std::uint32_t dir = o.direction;
            const std::uint64_t pct = o.pull_percent ? o.pull_percent : kDefaultPullPercentPagerank;
            if (dir != HYT_DIR_PUSH && g.E > 0 && g.E < (1ULL << 31))
                ensure_transpose(g, light_max, chunk);
            const bool can_pull = g.in_col.ptr != nullptr;
            require(dir != HYT_DIR_PULL || can_pull || g.E == 0, "pull rounds need fewer than 2^31 edges");
            PrPullOp<WORDS> pop{g.in_col.as<std::uint32_t>(), ws.share.as<ShareT<WORDS>>(),
                                ws.res.as<unsigned long long>()};

            Counters c0{};
            if (resume)
            {
                // the queue and counters of the last compaction of the run
                hip_check(hipMemcpyAsync(g.host_ctr->ptr, d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, s),
                          "read counters");
                hip_check(hipStreamSynchronize(s), "pagerank resume");
                c0 = *g.host_ctr->as<Counters>();
            }
            else
            {
                clear_counters(d_ctr, s);
                c0 = compact_and_read(g, sel);
            }
            std::uint32_t F = c0.next_count;
            unsigned long long Fe = c0.frontier_edges;
            st.converged = F == 0 ? 1 : 0;
            for (std::uint64_t round = 0; round < max_rounds && F > 0; ++round)
            {
                a.F = F;
                hip_check(hipMemsetAsync(d_ctr, 0, kRoundCounterBytes, s), "reset counters");
                const bool pull = can_pull && (dir == HYT_DIR_PULL ||
                                               (dir == HYT_DIR_AUTO && Fe * 100ULL >= pct * g.E));
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev0, s), "hipEventRecord");
                if (pull)
                {
                    pull_light<OffT, PrPullOp<WORDS>><<<blocks_for(g.V), kBlock, 0, s>>>(g.in_off.as<OffT>(), pop, g.V,
                                                                                   light_max);
                    hip_check(hipGetLastError(), "launch pr_pull_light");
                    pull_heavy<PrPullOp<WORDS>><<<kSingleHeavyGrid, kBlock, 0, s>>>(pop, g.in_heavy.as<HeavyChunk>(),
                                                                      g.in_heavy_count);
                    hip_check(hipGetLastError(), "launch pr_pull_heavy");
                    st.pull_rounds += 1;
                    st.pull_active_edges += Fe;
                }
                else
                {
                    expand_light<GraphRows<OffT>, PrOp<WORDS>><<<blocks_for(F), kBlock, 0, s>>>(whole_graph(g.off.as<OffT>(), a.q),
                                                                                     op, a);
                    hip_check(hipGetLastError(), "launch pr_expand_light");
                    expand_heavy<PrOp<WORDS>><<<kSingleHeavyGrid, kBlock, 0, s>>>(op, a);
                    hip_check(hipGetLastError(), "launch pr_expand_heavy");
                }
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev1, s), "hipEventRecord");
                const Counters c = compact_and_read(g, sel);
                if (o.time_kernels)
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
                st.rounds = round + 1;
                F = c.next_count;
                Fe = c.frontier_edges;
                if (F == 0)
                    st.converged = 1;
            }
        }

        void check_params(float damping, float eps)
        {
            require(damping >= 0.0F && damping < 1.0F, "PageRank damping_factor must be in [0, 1)");
            require(eps > 0.0F && std::isfinite(eps), "PageRank epsilon must be finite and positive");
        }
    } // namespace

    void finalize_ranks(DeviceGraph &g, float *out, bool out_on_device, hyt_run_stats &st)
    {
        Workspace &ws = g.ws;
        hipStream_t s = g.stream;
        unsigned long long *partial = ws.partial.as<unsigned long long>();
        std::uint64_t nb64 = (g.V + kBlock - 1) / kBlock;
        const unsigned nb = static_cast<unsigned>(nb64 < kFoldBlocks ? (nb64 ? nb64 : 1) : kFoldBlocks);
        pr_fold<<<nb, kBlock, 0, s>>>(ws.rank.as<unsigned long long>(), ws.res.as<unsigned long long>(), g.V, partial);
        hip_check(hipGetLastError(), "launch pr_fold");
        pr_total<<<1, kBlock, 0, s>>>(partial, nb);
        hip_check(hipGetLastError(), "launch pr_total");
        unsigned long long totals[2] = {0, 0};
        hip_check(hipMemcpyAsync(totals, partial + 2 * nb, sizeof(totals), hipMemcpyDeviceToHost, s),
                  "read totals");
        float *dst = out_on_device ? out : ws.dist.as<float>();
        pr_normalize<<<nb, kBlock, 0, s>>>(ws.rank.as<unsigned long long>(), ws.res.as<unsigned long long>(), g.V,
                                           partial + 2 * nb, dst);
        hip_check(hipGetLastError(), "launch pr_normalize");
        if (!out_on_device)
            hip_check(hipMemcpyAsync(out, dst, g.V * 4, hipMemcpyDeviceToHost, s), "copy ranks");
        hip_check(hipStreamSynchronize(s), "pagerank output");
        st.residual = totals[0] > 0 ? static_cast<double>(totals[1]) / static_cast<double>(totals[0]) : 0.0;
    }


    // One delta-push solve; seed == UINT32_MAX means global PageRank.
    void pagerank_single(DeviceGraph &g, std::uint32_t seed, float damping, float eps, float *out,
                         const hyt_run_options &o, hyt_run_stats &st)
    {
        pagerank_steps(g, seed, damping, eps, o, false, 0, st);
        finalize_ranks(g, out, o.out_on_device != 0, st);
    }

    void pagerank_steps(DeviceGraph &g, std::uint32_t seed, float damping, float eps, const hyt_run_options &o,
                        bool resume, std::uint64_t rounds, hyt_run_stats &st)
    {
        check_params(damping, eps);
        const std::uint32_t light_max = o.light_max_degree ? o.light_max_degree : kPagerankLightMaxDegree;
        const std::uint32_t chunk = o.heavy_chunk ? o.heavy_chunk : kPagerankHeavyChunk;
        require(light_max <= (1U << 20), "light_max_degree must be <= 2^20");
        require(chunk >= 64 && chunk <= (1U << 24), "heavy_chunk must be in [64, 2^24]");
        ensure_workspace(g, light_max, chunk);
        ensure_mass_state(g);
        hipStream_t s = g.stream;
        const bool global = seed == 0xFFFFFFFFU;
        const MassScale m = mass_scale(global ? g.V : 1, damping, eps);
        unsigned long long *res = g.ws.res.as<unsigned long long>();
        unsigned long long *rank = g.ws.rank.as<unsigned long long>();
        if (!resume && global)
        {
            std::uint64_t b = (g.V + kBlock - 1) / kBlock;
            pr_init_all<<<static_cast<unsigned>(b < 65536 ? b : 65536), kBlock, 0, s>>>(res, rank, g.V, m.r0);
            hip_check(hipGetLastError(), "launch pr_init");
        }
        else if (!resume)
        {
            launch_fill<unsigned long long>(res, g.V, 0ULL, s);
            launch_fill<unsigned long long>(rank, g.V, 0ULL, s);
            pr_seed<<<1, 64, 0, s>>>(res, seed, m.r0);
            hip_check(hipGetLastError(), "launch pr_seed");
        }
        // PPR keeps share words (its normalised mass is one seed's, so a row
        // flipping at eps would move eps of it); global PageRank float shares
        // (a flip moves eps of V, and pull rounds gather 4 bytes)
        if (g.off64)
        {
            if (global)
                run_push<std::uint64_t, false>(g, o, damping, m, st, light_max, chunk, resume, rounds);
            else
                run_push<std::uint64_t, true>(g, o, damping, m, st, light_max, chunk, resume, rounds);
        }
        else if (global)
            run_push<std::uint32_t, false>(g, o, damping, m, st, light_max, chunk, resume, rounds);
        else
            run_push<std::uint32_t, true>(g, o, damping, m, st, light_max, chunk, resume, rounds);
    }

    void ppr_batched(DeviceGraph &g, const std::uint32_t *seeds, std::uint32_t K, float damping, float eps,
                     float *out, const hyt_run_options &o, hyt_run_stats &st);
} // namespace hyt

extern "C" int hyt_pagerank(hyt_graph_t g, float damping, float epsilon, float *ranks,
                            const hyt_run_options *opts, hyt_run_stats *stats)
{
    return hyt::guarded("hyt_pagerank", [&] {
        using namespace hyt;
        require(g != nullptr && ranks != nullptr, "null graph or ranks");
        hyt_run_options o{};
        if (opts)
            o = *opts;
        hyt_run_stats st{};
        hip_check(hipSetDevice(g->g.device), "hipSetDevice");
        const auto t0 = std::chrono::steady_clock::now();
        pagerank_single(g->g, 0xFFFFFFFFU, damping, epsilon, ranks, o, st);
        st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (stats)
            *stats = st;
    });
}

extern "C" int hyt_ppr(hyt_graph_t g, const uint32_t *seeds, uint32_t num_seeds, float damping, float epsilon,
                       float *out, const hyt_run_options *opts, hyt_run_stats *stats)
{
    return hyt::guarded("hyt_ppr", [&] {
        using namespace hyt;
        require(g != nullptr && seeds != nullptr && out != nullptr, "null graph, seeds or out");
        require(num_seeds >= 1 && num_seeds <= 64, "num_seeds must be in [1, 64]");
        for (std::uint32_t k = 0; k < num_seeds; ++k)
            require(seeds[k] < g->g.V, "PPR seed vertex is out of range");
        hyt_run_options o{};
        if (opts)
            o = *opts;
        hyt_run_stats st{};
        hip_check(hipSetDevice(g->g.device), "hipSetDevice");
        const auto t0 = std::chrono::steady_clock::now();
        if (num_seeds == 1)
            pagerank_single(g->g, seeds[0], damping, epsilon, out, o, st);
        else
            ppr_batched(g->g, seeds, num_seeds, damping, epsilon, out, o, st);
        st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (stats)
            *stats = st;
    });
}
