// SSSP on the MI355X engine, with the interface of the original project's
// include/algorithms/sssp.hpp:11-51 (SSSPOptions, SSSPResult, sssp_gpu,
// sssp_contributions), plus the batched multi-source form.
#pragma once

#include "graph/csr_graph.hpp"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace hytgraph::algorithms
{
    struct SSSPOptions
    {
        graph::CSRGraph::vertex_id source = 0;
        float tolerance = 0.0F;
        std::size_t max_iterations = 0; // 0 selects |V|
    };

    struct SSSPContribution
    {
        std::size_t vertex = 0U;
        double contribution = 0.0;
    };

    // Per-vertex improvement current - candidate when it exceeds tolerance;
    // 1.0 when an unreachable vertex becomes reachable (original
    // src/algorithms/sssp.cpp:236-290).
    [[nodiscard]] std::vector<SSSPContribution>
    sssp_contributions(const std::vector<float> &current_distances,
                       const std::vector<float> &candidate_distances,
                       float tolerance = 0.0F);

    // sssp_contributions computed on the GPU (addition): the same values.
    [[nodiscard]] std::vector<SSSPContribution>
    sssp_contributions_gpu(const std::vector<float> &current_distances,
                           const std::vector<float> &candidate_distances,
                           float tolerance = 0.0F);

    struct SSSPResult
    {
        std::vector<float> distances;
        std::size_t iterations = 0;
        bool converged = false;
    };

    // Synchronous frontier relaxation on the GPU. Same distances (bit-exact),
    // iteration count and convergence flag as the original sssp_cpu/sssp_gpu.
    // std::invalid_argument for bad graphs/options, std::runtime_error when
    // no device is available or a HIP call fails.
    SSSPResult sssp_gpu(const graph::CSRGraph &graph, const SSSPOptions &options = {});

    // The original's CPU Bellman-Ford (src/algorithms/sssp.cpp:109-234),
    // declared for source compatibility only: this library replaces the GPU
    // path, and the definition comes from the original's CPU sources.
    SSSPResult sssp_cpu(const graph::CSRGraph &graph, const SSSPOptions &options = {});

    // Batched form (addition): K <= 64 sources in one [V][K] traversal.
    // distances[k] holds source k's distances (+inf unreachable).
    struct SSSPBatchResult
    {
        std::vector<std::vector<float>> distances;
        std::size_t rounds = 0;
        bool converged = false;
    };
    SSSPBatchResult sssp_gpu_batch(const graph::CSRGraph &graph,
                                   const std::vector<graph::CSRGraph::vertex_id> &sources,
                                   std::size_t max_iterations = 0);
} // namespace hytgraph::algorithms
