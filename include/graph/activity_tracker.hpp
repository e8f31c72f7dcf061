// Per-vertex activity flags and per-range activity statistics (interface of
// the original include/graph/activity_tracker.hpp:13-93). An active vertex
// contributes all of its out-edges to the active edge count.
#pragma once

#include "graph/csr_graph.hpp"

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace hytgraph::graph
{
    class ActivityTracker
    {
    public:
        using vertex_id = CSRGraph::vertex_id;
        using count_type = CSRGraph::offset_type;

        struct VertexRange
        {
            vertex_id begin = 0;
            vertex_id end = 0;
        };

        struct PartitionActivity
        {
            count_type active_vertices = 0;
            count_type active_edges = 0;
            count_type total_edges = 0;

            [[nodiscard]] bool has_active_vertices() const noexcept { return active_vertices != 0U; }
            [[nodiscard]] bool has_active_edges() const noexcept { return active_edges != 0U; }
        };

        explicit ActivityTracker(count_type vertex_count);

        [[nodiscard]] count_type vertex_count() const noexcept;
        void clear() noexcept;
        void set_active(vertex_id vertex, bool active = true);
        [[nodiscard]] bool is_active(vertex_id vertex) const;
        void set_active_vertices(const std::vector<vertex_id> &vertices);
        [[nodiscard]] std::vector<vertex_id> active_vertices() const;
        [[nodiscard]] count_type active_vertex_count() const noexcept;
        [[nodiscard]] count_type active_edge_count(const CSRGraph &graph) const;
        [[nodiscard]] std::vector<PartitionActivity>
        partition_statistics(const CSRGraph &graph, const std::vector<VertexRange> &partitions) const;

    private:
        count_type vertex_count_ = 0;
        std::vector<unsigned char> active_;
        count_type active_vertex_count_ = 0;
    };
} // namespace hytgraph::graph
