// ExpTMFilter::plan with the rules of the original
// src/transfer/filter_engine.cpp:146-220: active <=> active_edges != 0; a
// transferred partition moves edge_count() edges and edge_data_bytes().
#include "transfer/filter_engine.hpp"

#include "partition_checks.hpp"

#include <stdexcept>

namespace hytgraph::transfer
{
    std::size_t FilterTransferPlan::active_partition_count() const noexcept
    {
        std::size_t n = 0;
        for (const auto &p : partitions)
            n += p.active ? 1U : 0U;
        return n;
    }

    std::size_t FilterTransferPlan::transferred_partition_count() const noexcept
    {
        std::size_t n = 0;
        for (const auto &p : partitions)
            n += p.transferred ? 1U : 0U;
        return n;
    }

    graph::CSRGraph::offset_type FilterTransferPlan::transferred_edge_count() const noexcept
    {
        graph::CSRGraph::offset_type n = 0;
        for (const auto &p : partitions)
            n += p.transferred_edges;
        return n;
    }

    std::size_t FilterTransferPlan::transferred_byte_count() const noexcept
    {
        std::size_t n = 0;
        for (const auto &p : partitions)
            n += p.transferred_bytes;
        return n;
    }

    ExpTMFilter::ExpTMFilter(bool enabled) noexcept : enabled_(enabled) {}

    bool ExpTMFilter::enabled() const noexcept { return enabled_; }

    FilterTransferPlan ExpTMFilter::plan(const graph::CSRGraph &graph,
                                         const std::vector<graph::LogicalPartition> &partitions,
                                         const graph::ActivityTracker &activity) const
    {
        if (activity.vertex_count() != graph.num_vertices())
            throw std::invalid_argument("activity tracker vertex count does not match graph");
        detail::check_coverage(graph, partitions, false);
        const auto stats = activity.partition_statistics(graph, detail::ranges_of(partitions));
        FilterTransferPlan out;
        out.partitions.resize(partitions.size());
        for (std::size_t i = 0; i < partitions.size(); ++i)
        {
            FilterPartitionDecision &d = out.partitions[i];
            d.partition_index = i;
            d.active_edges = stats[i].active_edges;
            d.active = d.active_edges != 0U;
            d.total_edges = partitions[i].edge_count();
            d.transferred = !enabled_ || d.active;
            if (d.transferred)
            {
                d.transferred_edges = partitions[i].edge_count();
                d.transferred_bytes = static_cast<std::size_t>(partitions[i].edge_data_bytes());
            }
        }
        return out;
    }
} // namespace hytgraph::transfer
