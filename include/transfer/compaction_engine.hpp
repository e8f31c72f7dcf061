// ExpTM-Compaction (interface of the original
// include/transfer/compaction_engine.hpp:12-140): per logical partition with
// at least one active vertex, its active vertices in ascending order, their
// neighbour lists packed in that order and a u64 prefix index of size
// |active| + 1. The device execution is the hybrid engine's compaction path,
// which packs rows (and weights) on host threads into pinned staging and
// copies them with one transfer per batch.
#pragma once

#include "graph/activity_tracker.hpp"
#include "graph/csr_graph.hpp"
#include "graph/partition.hpp"

#include <cstddef>
#include <vector>

namespace hytgraph::transfer
{
    struct CompactedPartition
    {
        std::size_t partition_index = 0;
        graph::CSRGraph::vertex_id vertex_begin = 0;
        graph::CSRGraph::vertex_id vertex_end = 0;
        std::vector<graph::CSRGraph::vertex_id> active_vertices;
        std::vector<graph::CSRGraph::vertex_id> neighbors;
        std::vector<graph::CSRGraph::offset_type> neighbor_index;

        [[nodiscard]] graph::CSRGraph::offset_type active_edge_count() const noexcept;
        [[nodiscard]] std::size_t active_vertex_count() const noexcept;
        [[nodiscard]] std::size_t neighbor_bytes() const noexcept;
        [[nodiscard]] std::size_t index_bytes() const noexcept;
        [[nodiscard]] std::size_t total_bytes() const noexcept;
    };

    struct CompactionResult
    {
        std::vector<CompactedPartition> partitions;
        double compaction_seconds = 0.0; // host wall time of compact()

        [[nodiscard]] std::size_t compacted_partition_count() const noexcept;
        [[nodiscard]] graph::CSRGraph::offset_type active_edge_count() const noexcept;
        [[nodiscard]] graph::CSRGraph::offset_type active_vertex_count() const noexcept;
        [[nodiscard]] std::size_t neighbor_bytes() const noexcept;
        [[nodiscard]] std::size_t index_bytes() const noexcept;
        [[nodiscard]] std::size_t total_bytes() const noexcept;
    };

    class ExpTMCompaction
    {
    public:
        explicit ExpTMCompaction(bool enabled = true) noexcept;

        [[nodiscard]] bool enabled() const noexcept;

        // Disabled: an empty result.
        [[nodiscard]] CompactionResult compact(const graph::CSRGraph &graph,
                                               const std::vector<graph::LogicalPartition> &partitions,
                                               const graph::ActivityTracker &activity) const;

        // The packed neighbour stream of a compacted partition, after checking
        // that its index describes it.
        [[nodiscard]] static std::vector<graph::CSRGraph::vertex_id>
        expanded_neighbors(const CompactedPartition &partition);

    private:
        bool enabled_;
    };
} // namespace hytgraph::transfer
