// Drop-in GPU algorithms over the C ABI: sssp_gpu / pagerank_gpu with the
// signatures of the original include/algorithms/{sssp,pagerank}.hpp, and the
// host contribution functions. (The batched additions are in
// batched_api.cpp.)
//
// This translation unit only uses CSRGraph's public accessors and the
// original's declarations, so it builds against this repository's include/
// or the original project's include/ unchanged (it replaces cuda/sssp.cu and
// cuda/pagerank.cu there).
//
// Error mapping (original cuda/sssp.cu:18-25,81-98, cuda/pagerank.cu:76-98):
// invalid graphs and options -> std::invalid_argument with the original's
// messages; no device or a HIP failure -> std::runtime_error.
#include "algorithms/pagerank.hpp"
#include "algorithms/sssp.hpp"
#include "gpu_call.hpp"
#include "hytgraph_amd.h"

#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace hytgraph::algorithms
{
    using namespace detail;

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
        // tolerance > 0: the original's candidate + tol < best rule, applied
        // in incoming-edge order by the ordered pull rounds
        o.tolerance = options.tolerance;
        check(hyt_sssp(sg.g, &src, 1, bits.data(), &o, &st), "sssp_gpu");
        SSSPResult r;
        r.distances.resize(n);
        for (std::size_t v = 0; v < n; ++v)
            r.distances[v] = bits_to_float(bits[v]);
        r.iterations = static_cast<std::size_t>(st.rounds);
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

    std::vector<PageRankContribution> pagerank_contributions_gpu(const graph::CSRGraph &graph,
                                                                 const std::vector<float> &current,
                                                                 const PageRankOptions &options)
    {
        validate_pagerank(graph, options);
        const auto n = static_cast<std::size_t>(graph.num_vertices());
        if (current.size() != n)
            throw std::invalid_argument("PageRank current_ranks size must match graph vertex count");
        require_device("pagerank_contributions_gpu");
        ScopedGraph sg;
        upload(graph, false, sg, "pagerank_contributions_gpu");
        std::vector<double> c(n);
        check(hyt_pagerank_contributions(sg.g, current.data(), options.damping_factor, c.data(), 0),
              "pagerank_contributions_gpu");
        std::vector<PageRankContribution> out(n);
        for (std::size_t v = 0; v < n; ++v)
            out[v] = PageRankContribution{v, c[v]};
        return out;
    }

    std::vector<SSSPContribution> sssp_contributions_gpu(const std::vector<float> &current,
                                                         const std::vector<float> &candidate, float tolerance)
    {
        if (!(tolerance >= 0.0F) || !std::isfinite(tolerance))
            throw std::invalid_argument("SSSP contribution tolerance must be finite and non-negative");
        if (current.size() != candidate.size())
            throw std::invalid_argument("SSSP contribution distance vectors must have equal size");
        std::vector<SSSPContribution> out(current.size());
        if (current.empty())
            return out;
        require_device("sssp_contributions_gpu");
        std::vector<double> c(current.size());
        check(hyt_sssp_contributions(current.data(), candidate.data(), current.size(), tolerance, c.data(), 0),
              "sssp_contributions_gpu");
        for (std::size_t v = 0; v < c.size(); ++v)
            out[v] = SSSPContribution{v, c[v]};
        return out;
    }
} // namespace hytgraph::algorithms
