// ExpTM-Filter planning (interface of the original
// include/transfer/filter_engine.hpp:13-69): a logical partition with at
// least one active edge moves whole; one without is skipped (all move when
// the engine is disabled). The device execution of this plan is the hybrid
// engine's filter path (hyt_hybrid_*), which copies selected partitions with
// hipMemcpyAsync on a copy stream.
#pragma once

#include "graph/activity_tracker.hpp"
#include "graph/csr_graph.hpp"
#include "graph/partition.hpp"

#include <cstddef>
#include <vector>

namespace hytgraph::transfer
{
    struct FilterPartitionDecision
    {
        std::size_t partition_index = 0;
        bool active = false;
        bool transferred = false;
        graph::CSRGraph::offset_type active_edges = 0;
        graph::CSRGraph::offset_type total_edges = 0;
        graph::CSRGraph::offset_type transferred_edges = 0;
        std::size_t transferred_bytes = 0;
    };

    struct FilterTransferPlan
    {
        std::vector<FilterPartitionDecision> partitions;

        [[nodiscard]] std::size_t active_partition_count() const noexcept;
        [[nodiscard]] std::size_t transferred_partition_count() const noexcept;
        [[nodiscard]] graph::CSRGraph::offset_type transferred_edge_count() const noexcept;
        [[nodiscard]] std::size_t transferred_byte_count() const noexcept;
    };

    class ExpTMFilter
    {
    public:
        explicit ExpTMFilter(bool enabled = true) noexcept;

        [[nodiscard]] bool enabled() const noexcept;

        [[nodiscard]] FilterTransferPlan plan(const graph::CSRGraph &graph,
                                              const std::vector<graph::LogicalPartition> &partitions,
                                              const graph::ActivityTracker &activity) const;

    private:
        bool enabled_;
    };
} // namespace hytgraph::transfer
