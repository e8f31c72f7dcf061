// Drop-in GPU algorithms over the C ABI: sssp_gpu / pagerank_gpu with the
// signatures of the original include/algorithms/{sssp,pagerank}.hpp, plus the
// batched additions and the host contribution functions.
//
// This translation unit only uses CSRGraph's public accessors, so it builds
// against this repository's include/ or the original project's include/
// unchanged (it replaces cuda/sssp.cu and cuda/pagerank.cu there).
//
// Error mapping (original cuda/sssp.cu:18-25,81-98, cuda/pagerank.cu:76-98):
// invalid graphs and options -> std::invalid_argument with the original's
// messages; no device or a HIP failure -> std::runtime_error.
#include "algorithms/pagerank.hpp"
#include "algorithms/sssp.hpp"
#include "hytgraph_amd.h"

#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace hytgraph::algorithms
{
    namespace
    {
        void check(int status, const char *what)
        {
            if (status == HYT_OK)
                return;
            std::string msg = std::string(what) + ": " + hyt_last_error();
            if (status == HYT_ERR_INVALID_ARGUMENT)
                throw std::invalid_argument(msg);
            throw std::runtime_error(msg);
        }

        void require_device(const char *what)
        {
            int n = 0;
            check(hyt_device_count(&n), what);
            if (n == 0)
                throw std::runtime_error(std::string(what) + ": no HIP device is available");
        }

        // Owns a device graph for the duration of one call (the original
        // allocates and frees device memory inside each call as well).
        struct ScopedGraph
        {
            hyt_graph_t g = nullptr;
            ~ScopedGraph()
            {
                if (g)
                    hyt_graph_free(g);
            }
        };

        // Uploads the CSR; float weights travel as IEEE bit patterns.
        void upload(const graph::CSRGraph &graph, bool with_weights, ScopedGraph &sg, const char *what)
        {
            const auto &off = graph.row_offsets();
            const auto &col = graph.column_indices();
            const std::uint32_t *w = nullptr;
            std::vector<std::uint32_t> bits;
            if (with_weights && graph.has_weights())
            {
                const auto &wf = graph.edge_weights();
                bits.resize(wf.size());
                if (!wf.empty())
                    std::memcpy(bits.data(), wf.data(), wf.size() * sizeof(float));
                w = bits.data();
            }
            check(hyt_graph_upload(graph.num_vertices(), graph.num_edges(), off.data(), col.data(), w,
                                   HYT_GRAPH_WEIGHTS_F32, &sg.g),
                  what);
        }

        void validate_sssp(const graph::CSRGraph &graph, const SSSPOptions &options)
        {
            if (graph.num_vertices() == 0U)
                throw std::invalid_argument("SSSP requires at least one vertex");
            if (static_cast<graph::CSRGraph::offset_type>(options.source) >= graph.num_vertices())
                throw std::invalid_argument("SSSP source vertex is out of range");
            if (!(options.tolerance >= 0.0F) || !std::isfinite(options.tolerance))
                throw std::invalid_argument("SSSP tolerance must be finite and non-negative");
            for (const float w : graph.edge_weights())
                if (!std::isfinite(w) || w < 0.0F)
                    throw std::invalid_argument("SSSP requires finite non-negative edge weights");
        }

        void validate_pagerank(const graph::CSRGraph &graph, const PageRankOptions &options)
        {
            if (graph.num_vertices() == 0U)
                throw std::invalid_argument("PageRank requires at least one vertex");
            if (!(options.damping_factor >= 0.0F && options.damping_factor < 1.0F))
                throw std::invalid_argument("PageRank damping_factor must be in [0, 1)");
            if (!(options.tolerance > 0.0F) || !std::isfinite(options.tolerance))
                throw std::invalid_argument("PageRank tolerance must be finite and positive");
            if (options.max_iterations == 0U)
                throw std::invalid_argument("PageRank max_iterations must be greater than zero");
        }

        float bits_to_float(std::uint32_t b)
        {
            float f;
            std::memcpy(&f, &b, 4);
            return f;
        }
    } // namespace

    SSSPResult sssp_gpu(const graph::CSRGraph &graph, const SSSPOptions &options)
    {
        validate_sssp(graph, options);
        require_device("sssp_gpu");
        ScopedGraph sg;
        upload(graph, true, sg, "sssp_gpu");
        const auto n = static_cast<std::size_t>(graph.num_vertices());
        std::vector<std::uint32_t> bits(n);
        hyt_run_options o{};
        o.mode = HYT_SSSP_SYNC;
        o.max_rounds = options.max_iterations;
        hyt_run_stats st{};
        const std::uint32_t src = options.source;
        if (options.tolerance > 0.0F)
        {
            // The original's tolerance rule (candidate + tol < best, applied
            // in incoming-edge order) is order-dependent; the frontier kernels
            // implement the tolerance-0 rule only.
            throw std::invalid_argument("sssp_gpu: tolerance > 0 is not supported by the HIP path");
        }
        check(hyt_sssp(sg.g, &src, 1, bits.data(), &o, &st), "sssp_gpu");
        SSSPResult r;
        r.distances.resize(n);
        for (std::size_t v = 0; v < n; ++v)
            r.distances[v] = bits_to_float(bits[v]);
        r.iterations = static_cast<std::size_t>(st.rounds);
        r.converged = st.converged != 0;
        return r;
    }

    SSSPBatchResult sssp_gpu_batch(const graph::CSRGraph &graph,
                                   const std::vector<graph::CSRGraph::vertex_id> &sources,
                                   std::size_t max_iterations)
    {
        if (sources.empty() || sources.size() > 64)
            throw std::invalid_argument("sssp_gpu_batch: between 1 and 64 sources are supported");
        SSSPOptions probe;
        for (auto s : sources)
        {
            probe.source = s;
            validate_sssp(graph, probe);
        }
        require_device("sssp_gpu_batch");
        ScopedGraph sg;
        upload(graph, true, sg, "sssp_gpu_batch");
        const auto n = static_cast<std::size_t>(graph.num_vertices());
        const std::size_t K = sources.size();
        std::vector<std::uint32_t> bits(n * K);
        hyt_run_options o{};
        o.mode = HYT_SSSP_ASYNC;
        o.max_rounds = max_iterations;
        hyt_run_stats st{};
        check(hyt_sssp(sg.g, sources.data(), static_cast<std::uint32_t>(K), bits.data(), &o, &st),
              "sssp_gpu_batch");
        SSSPBatchResult r;
        r.distances.assign(K, std::vector<float>(n));
        for (std::size_t v = 0; v < n; ++v)
            for (std::size_t k = 0; k < K; ++k)
                r.distances[k][v] = bits_to_float(bits[v * K + k]);
        r.rounds = static_cast<std::size_t>(st.rounds);
        r.converged = st.converged != 0;
        return r;
    }

    PageRankResult pagerank_gpu(const graph::CSRGraph &graph, const PageRankOptions &options)
    {
        validate_pagerank(graph, options);
        require_device("pagerank_gpu");
        ScopedGraph sg;
        upload(graph, false, sg, "pagerank_gpu");
        const float d = options.damping_factor;
        // Normalised L1 error of delta-push <= 2 d eps / (1-d)^2; this eps
        // makes it <= d * tol / (1-d), the bound of the original's own
        // stopping rule.
        const float eps = options.tolerance * (1.0F - d) * 0.5F;
        hyt_run_options o{};
        o.max_rounds = options.max_iterations * 4;
        hyt_run_stats st{};
        PageRankResult r;
        r.ranks.resize(static_cast<std::size_t>(graph.num_vertices()));
        check(hyt_pagerank(sg.g, d, eps > 0.0F ? eps : std::numeric_limits<float>::min(), r.ranks.data(), &o, &st),
              "pagerank_gpu");
        r.iterations = static_cast<std::size_t>(st.rounds);
        r.residual = static_cast<float>(st.residual);
        r.converged = st.converged != 0;
        return r;
    }

    PPRResult ppr_gpu(const graph::CSRGraph &graph, const std::vector<graph::CSRGraph::vertex_id> &seeds,
                      float damping_factor, float epsilon)
    {
        if (seeds.empty() || seeds.size() > 64)
            throw std::invalid_argument("ppr_gpu: between 1 and 64 seeds are supported");
        PageRankOptions probe;
        probe.damping_factor = damping_factor;
        validate_pagerank(graph, probe);
        for (auto s : seeds)
            if (static_cast<graph::CSRGraph::offset_type>(s) >= graph.num_vertices())
                throw std::invalid_argument("PPR seed vertex is out of range");
        require_device("ppr_gpu");
        ScopedGraph sg;
        upload(graph, false, sg, "ppr_gpu");
        const auto n = static_cast<std::size_t>(graph.num_vertices());
        const std::size_t K = seeds.size();
        std::vector<float> out(n * K);
        hyt_run_options o{};
        hyt_run_stats st{};
        check(hyt_ppr(sg.g, seeds.data(), static_cast<std::uint32_t>(K), damping_factor, epsilon, out.data(), &o,
                      &st),
              "ppr_gpu");
        PPRResult r;
        r.ranks.assign(K, std::vector<float>(n));
        for (std::size_t v = 0; v < n; ++v)
            for (std::size_t k = 0; k < K; ++k)
                r.ranks[k][v] = out[v * K + k];
        r.rounds = static_cast<std::size_t>(st.rounds);
        r.converged = st.converged != 0;
        return r;
    }

    std::vector<SSSPContribution> sssp_contributions(const std::vector<float> &current,
                                                     const std::vector<float> &candidate, float tolerance)
    {
        if (!(tolerance >= 0.0F) || !std::isfinite(tolerance))
            throw std::invalid_argument("SSSP contribution tolerance must be finite and non-negative");
        if (current.size() != candidate.size())
            throw std::invalid_argument("SSSP contribution distance vectors must have equal size");
        std::vector<SSSPContribution> out(current.size());
        for (std::size_t v = 0; v < current.size(); ++v)
        {
            const float c = current[v], n = candidate[v];
            double value = 0.0;
            if (std::isfinite(c) && std::isfinite(n))
            {
                const float gain = c - n;
                if (gain > tolerance)
                    value = static_cast<double>(gain);
            }
            else if (!std::isfinite(c) && std::isfinite(n))
            {
                value = 1.0; // newly reachable: unit priority
            }
            out[v] = SSSPContribution{v, value};
        }
        return out;
    }

    std::vector<PageRankContribution> pagerank_contributions(const graph::CSRGraph &graph,
                                                             const std::vector<float> &current,
                                                             const PageRankOptions &options)
    {
        validate_pagerank(graph, options);
        const auto n = static_cast<std::size_t>(graph.num_vertices());
        if (current.size() != n)
            throw std::invalid_argument("PageRank current_ranks size must match graph vertex count");
        const float d = options.damping_factor;
        const float base = (1.0F - d) / static_cast<float>(n);
        const auto &off = graph.row_offsets();
        const auto &col = graph.column_indices();
        float dangling = 0.0F;
        for (std::size_t u = 0; u < n; ++u)
            if (off[u + 1] == off[u])
                dangling += current[u];
        const float share = d * dangling / static_cast<float>(n);
        // Incoming sums accumulated in (source, edge) order per destination,
        // the order of the original's transposed CSR, so the float sums match.
        std::vector<std::uint64_t> in_off(n + 1, 0);
        for (auto v : col)
            ++in_off[static_cast<std::size_t>(v) + 1];
        for (std::size_t v = 1; v <= n; ++v)
            in_off[v] += in_off[v - 1];
        std::vector<std::uint64_t> cur(in_off.begin(), in_off.end() - 1);
        std::vector<float> contrib(col.size());
        for (std::size_t u = 0; u < n; ++u)
        {
            const auto deg = off[u + 1] - off[u];
            for (auto e = off[u]; e < off[u + 1]; ++e)
                contrib[cur[col[e]]++] = current[u] / static_cast<float>(deg);
        }
        std::vector<PageRankContribution> out(n);
        for (std::size_t v = 0; v < n; ++v)
        {
            float sum = 0.0F;
            for (auto p = in_off[v]; p < in_off[v + 1]; ++p)
                sum += contrib[p];
            const float next = base + share + d * sum;
            out[v] = PageRankContribution{v, std::fabs(next - current[v])};
        }
        return out;
    }
} // namespace hytgraph::algorithms
