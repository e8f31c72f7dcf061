// Delta-push PageRank / personalized PageRank (FP32) on gfx950.
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
// A round snapshots (and zeroes) the residuals of its frontier first, so the
// set of vertices pushed in each round does not depend on thread timing; a
// vertex joins the next frontier exactly when an atomicAdd carries its
// residual across eps, which happens at most once per round.
#include "expand.hpp"

#include <chrono>
#include <cmath>

namespace hyt
{
    namespace
    {
        struct PrOp
        {
            const std::uint32_t *__restrict__ col;
            float *res;
            const std::uint32_t *qd; // snapshot residuals (bits)
            float damping;
            float eps;

            __device__ std::uint32_t row_value(std::uint32_t i, std::uint32_t, std::uint64_t deg) const
            {
                if (deg == 0)
                    return 0U;
                return __float_as_uint(damping * __uint_as_float(qd[i]) / static_cast<float>(deg));
            }
            __device__ bool load(std::uint64_t pos, std::uint32_t share, std::uint32_t &v, std::uint32_t &c) const
            {
                v = col[pos];
                c = share;
                return true;
            }
            __device__ bool apply(std::uint32_t v, std::uint32_t c) const
            {
                const float add = __uint_as_float(c);
                const float old = unsafeAtomicAdd(&res[v], add);
                return old <= eps && old + add > eps;
            }
        };

        // Moves the frontier's residuals into rank and records them.
        __global__ void pr_snapshot(const std::uint32_t *q, std::uint32_t F, float *res, float *rank,
                                    std::uint32_t *qd)
        {
            const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i < F)
            {
                const std::uint32_t u = q[i];
                const float r = res[u];
                res[u] = 0.0F;
                rank[u] += r;
                qd[i] = __float_as_uint(r);
            }
        }

        __global__ void pr_init_all(float *res, float *rank, std::uint32_t *q, std::uint64_t V, float r0)
        {
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; i < V; i += stride)
            {
                res[i] = r0;
                rank[i] = 0.0F;
                q[i] = static_cast<std::uint32_t>(i);
            }
        }

        __global__ void pr_seed(float *res, std::uint32_t *q, std::uint32_t s, float r0)
        {
            if (threadIdx.x == 0 && blockIdx.x == 0)
            {
                res[s] = r0;
                q[0] = s;
            }
        }

        // x = rank + res (written into rank); partial sums of x and res.
        __global__ __launch_bounds__(kBlock) void pr_fold(float *rank, const float *res, std::uint64_t V,
                                                          double *partial)
        {
            __shared__ double sx[kBlock], sr[kBlock];
            double ax = 0.0, ar = 0.0;
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; i < V; i += stride)
            {
                const float r = res[i];
                const float x = rank[i] + r;
                rank[i] = x;
                ax += x;
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

        // Single workgroup: totals of the partials into partial[0..1] of the
        // tail slot (index 2*nblocks).
        __global__ __launch_bounds__(kBlock) void pr_total(double *partial, unsigned nblocks)
        {
            __shared__ double sx[kBlock], sr[kBlock];
            double ax = 0.0, ar = 0.0;
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

        __global__ void pr_normalize(const float *x, std::uint64_t V, const double *total, float *out,
                                     std::uint32_t out_stride)
        {
            const double t = total[0];
            const double inv = t > 0.0 ? 1.0 / t : 0.0;
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; i < V; i += stride)
                out[i * out_stride] = static_cast<float>(static_cast<double>(x[i]) * inv);
        }

        constexpr unsigned kFoldBlocks = 1024; // partial holds 2*kFoldBlocks+2 doubles

        template <class OffT>
        void run_push(DeviceGraph &g, const hyt_run_options &o, float damping, float eps, hyt_run_stats &st,
                      std::uint32_t F, std::uint32_t light_max, std::uint32_t chunk)
        {
            Workspace &ws = g.ws;
            hipStream_t s = g.stream;
            Counters *h_ctr = g.host_ctr->as<Counters>();
            Counters *d_ctr = ws.ctr.as<Counters>();
            std::uint32_t *qc = ws.q0.as<std::uint32_t>();
            std::uint32_t *qn = ws.q1.as<std::uint32_t>();
            const std::uint64_t max_rounds = o.max_rounds ? o.max_rounds : 100000;

            PrOp op;
            op.col = g.col.as<std::uint32_t>();
            op.res = ws.dist.as<float>();
            op.qd = ws.qd.as<std::uint32_t>();
            op.damping = damping;
            op.eps = eps;

            ExpandArgs a;
            a.ctr = d_ctr;
            a.heavy = ws.heavy.as<HeavyChunk>();
            a.heavy_cap = ws.heavy_cap;
            a.light_max = light_max;
            a.chunk = chunk;

            hip_check(hipMemsetAsync(d_ctr, 0, sizeof(Counters), s), "reset counters");
            st.converged = 0;
            for (std::uint64_t round = 0; round < max_rounds; ++round)
            {
                pr_snapshot<<<blocks_for(F), kBlock, 0, s>>>(qc, F, op.res, ws.rank.as<float>(), ws.qd.as<std::uint32_t>());
                hip_check(hipGetLastError(), "launch pr_snapshot");
                a.q = qc;
                a.F = F;
                a.qn = qn;
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev0, s), "hipEventRecord");
                expand_light<OffT, PrOp><<<blocks_for(F), kBlock, 0, s>>>(g.off.as<OffT>(), op, a);
                hip_check(hipGetLastError(), "launch pr_expand_light");
                expand_heavy<PrOp><<<kHeavyGrid, kBlock, 0, s>>>(op, a);
                hip_check(hipGetLastError(), "launch pr_expand_heavy");
                if (o.time_kernels)
                    hip_check(hipEventRecord(g.ev1, s), "hipEventRecord");
                hip_check(hipMemcpyAsync(h_ctr, d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, s), "read counters");
                hip_check(hipStreamSynchronize(s), "pagerank round");
                if (o.time_kernels)
                {
                    float ms = 0;
                    hip_check(hipEventElapsedTime(&ms, g.ev0, g.ev1), "hipEventElapsedTime");
                    st.expand_ms += ms;
                    st.expand_launches += 1;
                }
                const Counters c = *h_ctr;
                if (c.heavy_overflow)
                    fail(HYT_ERR_INTERNAL, "heavy chunk queue overflow");
                st.frontier_total += F;
                st.edges_total += c.edges;
                st.updates_total += c.next_count;
                st.rounds = round + 1;
                if (c.next_count == 0)
                {
                    st.converged = 1;
                    break;
                }
                F = c.next_count;
                std::swap(qc, qn);
                hip_check(hipMemsetAsync(d_ctr, 0, offsetof(Counters, overflow), s), "reset counters");
                hip_check(hipMemsetAsync(&d_ctr->edges, 0, sizeof(unsigned long long), s), "reset counters");
            }
        }

        void finalize(DeviceGraph &g, float *out, bool out_on_device, std::uint32_t out_stride, hyt_run_stats &st)
        {
            Workspace &ws = g.ws;
            hipStream_t s = g.stream;
            double *partial = ws.partial.as<double>();
            std::uint64_t nb64 = (g.V + kBlock - 1) / kBlock;
            const unsigned nb = static_cast<unsigned>(nb64 < kFoldBlocks ? (nb64 ? nb64 : 1) : kFoldBlocks);
            pr_fold<<<nb, kBlock, 0, s>>>(ws.rank.as<float>(), ws.dist.as<float>(), g.V, partial);
            hip_check(hipGetLastError(), "launch pr_fold");
            pr_total<<<1, kBlock, 0, s>>>(partial, nb);
            hip_check(hipGetLastError(), "launch pr_total");
            double totals[2] = {0, 0};
            hip_check(hipMemcpyAsync(totals, partial + 2 * nb, sizeof(totals), hipMemcpyDeviceToHost, s),
                      "read totals");
            float *dst = out_on_device ? out : ws.dist.as<float>();
            const std::uint32_t stride = out_on_device ? out_stride : 1U;
            pr_normalize<<<nb, kBlock, 0, s>>>(ws.rank.as<float>(), g.V, partial + 2 * nb, dst, stride);
            hip_check(hipGetLastError(), "launch pr_normalize");
            if (!out_on_device)
            {
                if (out_stride == 1)
                    hip_check(hipMemcpyAsync(out, dst, g.V * 4, hipMemcpyDeviceToHost, s), "copy ranks");
                else
                    hip_check(hipMemcpy2DAsync(out, out_stride * 4, dst, 4, 4, g.V, hipMemcpyDeviceToHost, s),
                              "copy ranks");
            }
            hip_check(hipStreamSynchronize(s), "pagerank output");
            st.residual = totals[0] > 0 ? totals[1] / totals[0] : 0.0;
        }

        void check_params(float damping, float eps)
        {
            require(damping >= 0.0F && damping < 1.0F, "PageRank damping_factor must be in [0, 1)");
            require(eps > 0.0F && std::isfinite(eps), "PageRank epsilon must be finite and positive");
        }

        struct Split
        {
            std::uint32_t light_max, chunk;
        };
        Split split_of(const hyt_run_options &o)
        {
            Split sp{o.light_max_degree ? o.light_max_degree : kDefaultLightMaxDegree,
                     o.heavy_chunk ? o.heavy_chunk : kDefaultHeavyChunk};
            require(sp.light_max <= (1U << 20), "light_max_degree must be <= 2^20");
            require(sp.chunk >= 64 && sp.chunk <= (1U << 24), "heavy_chunk must be in [64, 2^24]");
            return sp;
        }
    } // namespace

    // One delta-push solve; seed == UINT32_MAX means global PageRank.
    void pagerank_single(DeviceGraph &g, std::uint32_t seed, float damping, float eps, float *out,
                         std::uint32_t out_stride, const hyt_run_options &o, hyt_run_stats &st)
    {
        check_params(damping, eps);
        const Split sp = split_of(o);
        ensure_workspace(g, sp.light_max, sp.chunk);
        hipStream_t s = g.stream;
        const float r0 = 1.0F - damping;
        std::uint32_t F;
        if (seed == 0xFFFFFFFFU)
        {
            std::uint64_t b = (g.V + kBlock - 1) / kBlock;
            pr_init_all<<<static_cast<unsigned>(b < 65536 ? b : 65536), kBlock, 0, s>>>(
                g.ws.dist.as<float>(), g.ws.rank.as<float>(), g.ws.q0.as<std::uint32_t>(), g.V, r0);
            hip_check(hipGetLastError(), "launch pr_init");
            F = static_cast<std::uint32_t>(g.V);
        }
        else
        {
            launch_fill<float>(g.ws.dist.as<float>(), g.V, 0.0F, s);
            launch_fill<float>(g.ws.rank.as<float>(), g.V, 0.0F, s);
            pr_seed<<<1, 64, 0, s>>>(g.ws.dist.as<float>(), g.ws.q0.as<std::uint32_t>(), seed, r0);
            hip_check(hipGetLastError(), "launch pr_seed");
            F = 1;
        }
        if (g.off64)
            run_push<std::uint64_t>(g, o, damping, eps, st, F, sp.light_max, sp.chunk);
        else
            run_push<std::uint32_t>(g, o, damping, eps, st, F, sp.light_max, sp.chunk);
        finalize(g, out, o.out_on_device != 0, out_stride, st);
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
        pagerank_single(g->g, 0xFFFFFFFFU, damping, epsilon, ranks, 1, o, st);
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
            pagerank_single(g->g, seeds[0], damping, epsilon, out, 1, o, st);
        else
            ppr_batched(g->g, seeds, num_seeds, damping, epsilon, out, o, st);
        st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (stats)
            *stats = st;
    });
}
