// PageRank on the MI355X engine, with the interface of the original
// project's include/algorithms/pagerank.hpp:11-52 (PageRankOptions,
// PageRankResult, pagerank_gpu, pagerank_contributions), plus personalized
// PageRank.
#pragma once

#include "graph/csr_graph.hpp"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace hytgraph::algorithms
{
    struct PageRankOptions
    {
        float damping_factor = 0.85F;
        float tolerance = 1.0e-6F;
        std::size_t max_iterations = 100;
    };

    struct PageRankContribution
    {
        std::size_t vertex = 0U;
        double contribution = 0.0;
    };

    struct PageRankResult
    {
        std::vector<float> ranks;
        std::size_t iterations = 0;
        float residual = 0.0F;
        bool converged = false;
    };

    // |one Jacobi step - current| per vertex (original
    // src/algorithms/pagerank.cpp:165-243, same float evaluation order).
    [[nodiscard]] std::vector<PageRankContribution>
    pagerank_contributions(const graph::CSRGraph &graph,
                           const std::vector<float> &current_ranks,
                           const PageRankOptions &options = {});

    // Delta-push PageRank on the GPU; converges to the same fixed point as the
    // original Jacobi iteration. The per-vertex threshold is derived from
    // options.tolerance so the normalised L1 error stays below it:
    // eps = tolerance * (1-d)^2 / 2. iterations = push rounds; residual =
    // normalised L1 mass left unpushed; converged = no residual above eps
    // within max_iterations rounds (0 = unlimited).
    PageRankResult pagerank_gpu(const graph::CSRGraph &graph, const PageRankOptions &options = {});

    // pagerank_contributions computed on the GPU (addition): the same values,
    // bit for bit (float sums in the original's order; options validated as
    // the original validates them).
    [[nodiscard]] std::vector<PageRankContribution>
    pagerank_contributions_gpu(const graph::CSRGraph &graph,
                               const std::vector<float> &current_ranks,
                               const PageRankOptions &options = {});

    // The original's CPU Jacobi iteration (src/algorithms/pagerank.cpp:89-163),
    // declared for source compatibility only: this library replaces the GPU
    // path, and the definition comes from the original's CPU sources.
    PageRankResult pagerank_cpu(const graph::CSRGraph &graph, const PageRankOptions &options = {});

    // Personalized PageRank (addition) for up to 64 seeds in one [V][K]
    // traversal; ranks[k] sums to 1. epsilon is the per-vertex residual
    // threshold.
    struct PPRResult
    {
        std::vector<std::vector<float>> ranks;
        std::size_t rounds = 0;
        bool converged = false;
    };
    PPRResult ppr_gpu(const graph::CSRGraph &graph,
                      const std::vector<graph::CSRGraph::vertex_id> &seeds,
                      float damping_factor = 0.85F, float epsilon = 1.0e-6F);
} // namespace hytgraph::algorithms
