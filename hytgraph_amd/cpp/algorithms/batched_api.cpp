// Batched additions of the C++ API (not in the original): K SSSP sources or
// K personalized-PageRank seeds per call over the [V][K] kernels.
#include "algorithms/pagerank.hpp"
#include "algorithms/sssp.hpp"
#include "gpu_call.hpp"
#include "hytgraph_amd.h"

#include <stdexcept>
#include <vector>

namespace hytgraph::algorithms
{
    using namespace detail;

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
} // namespace hytgraph::algorithms
