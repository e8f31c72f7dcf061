// Shared parts of the C++ GPU entry points (drop_in.cpp, batched_api.cpp):
// status -> exception mapping, device check, per-call device graph, and the
// original's argument validation. Internal to the C++ library.
#pragma once

#include "algorithms/pagerank.hpp"
#include "algorithms/sssp.hpp"
#include "hytgraph_amd.h"

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace hytgraph::algorithms::detail
{
    inline void check(int status, const char *what)
    {
        if (status == HYT_OK)
            return;
        std::string msg = std::string(what) + ": " + hyt_last_error();
        if (status == HYT_ERR_INVALID_ARGUMENT)
            throw std::invalid_argument(msg);
        throw std::runtime_error(msg);
    }

    inline void require_device(const char *what)
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
    inline void upload(const graph::CSRGraph &graph, bool with_weights, ScopedGraph &sg, const char *what)
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

    inline void validate_sssp(const graph::CSRGraph &graph, const SSSPOptions &options)
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

    inline void validate_pagerank(const graph::CSRGraph &graph, const PageRankOptions &options)
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

    inline float bits_to_float(std::uint32_t b)
    {
        float f;
        std::memcpy(&f, &b, 4);
        return f;
    }
} // namespace hytgraph::algorithms::detail
