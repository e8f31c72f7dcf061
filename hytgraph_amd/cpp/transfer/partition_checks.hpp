// Coverage checks of a logical partition list shared by the transfer
// planners (original src/transfer/filter_engine.cpp:10-74 and
// zero_copy_engine.cpp:10-98 use the same rules and messages; zero-copy also
// requires CSR row boundaries).
#pragma once

#include "graph/activity_tracker.hpp"
#include "graph/csr_graph.hpp"
#include "graph/partition.hpp"

#include <stdexcept>
#include <vector>

namespace hytgraph::transfer::detail
{
    inline void check_coverage(const graph::CSRGraph &g, const std::vector<graph::LogicalPartition> &parts,
                               bool row_boundaries)
    {
        const auto V = g.num_vertices();
        const auto E = g.num_edges();
        graph::CSRGraph::offset_type next_v = 0, next_e = 0;
        for (const auto &p : parts)
        {
            if (p.vertex_begin() != next_v)
                throw std::invalid_argument("logical partitions must cover vertices contiguously");
            if (p.edge_begin() != next_e)
                throw std::invalid_argument("logical partitions must cover edges contiguously");
            if (p.vertex_begin() > p.vertex_end())
                throw std::invalid_argument("logical partition has invalid vertex range");
            if (p.edge_begin() > p.edge_end())
                throw std::invalid_argument("logical partition has invalid edge range");
            if (static_cast<graph::CSRGraph::offset_type>(p.vertex_end()) > V)
                throw std::out_of_range("logical partition vertex range exceeds graph");
            if (p.edge_end() > E)
                throw std::out_of_range("logical partition edge range exceeds graph");
            if (row_boundaries)
            {
                const auto &off = g.row_offsets();
                if (off[p.vertex_begin()] != p.edge_begin())
                    throw std::invalid_argument("logical partition edge_begin is not a CSR row boundary");
                if (off[p.vertex_end()] != p.edge_end())
                    throw std::invalid_argument("logical partition edge_end is not a CSR row boundary");
            }
            next_v = p.vertex_end();
            next_e = p.edge_end();
        }
        if (next_v != V)
            throw std::invalid_argument("logical partitions do not cover all graph vertices");
        if (next_e != E)
            throw std::invalid_argument("logical partitions do not cover all graph edges");
    }

    inline std::vector<graph::ActivityTracker::VertexRange> ranges_of(const std::vector<graph::LogicalPartition> &parts)
    {
        std::vector<graph::ActivityTracker::VertexRange> r;
        r.reserve(parts.size());
        for (const auto &p : parts)
            r.push_back({p.vertex_begin(), p.vertex_end()});
        return r;
    }
} // namespace hytgraph::transfer::detail
