// Logical partitioning and activity tracking with the semantics of the
// original src/graph/partition.cpp:8-196 and src/graph/activity_tracker.cpp:9-155.
#include "graph/activity_tracker.hpp"
#include "graph/partition.hpp"

#include <algorithm>
#include <stdexcept>

namespace hytgraph::graph
{
    LogicalPartition::LogicalPartition(vertex_id vb, vertex_id ve, offset_type eb, offset_type ee, byte_type target)
        : vertex_begin_(vb), vertex_end_(ve), edge_begin_(eb), edge_end_(ee), target_bytes_(target)
    {
        if (vb > ve)
            throw std::invalid_argument("LogicalPartition: vertex_begin must not exceed vertex_end");
        if (eb > ee)
            throw std::invalid_argument("LogicalPartition: edge_begin must not exceed edge_end");
        if (target == 0U)
            throw std::invalid_argument("LogicalPartition: target_bytes must be greater than zero");
    }
    LogicalPartition::vertex_id LogicalPartition::vertex_begin() const noexcept { return vertex_begin_; }
    LogicalPartition::vertex_id LogicalPartition::vertex_end() const noexcept { return vertex_end_; }
    LogicalPartition::offset_type LogicalPartition::edge_begin() const noexcept { return edge_begin_; }
    LogicalPartition::offset_type LogicalPartition::edge_end() const noexcept { return edge_end_; }
    LogicalPartition::offset_type LogicalPartition::vertex_count() const noexcept
    {
        return static_cast<offset_type>(vertex_end_ - vertex_begin_);
    }
    LogicalPartition::offset_type LogicalPartition::edge_count() const noexcept { return edge_end_ - edge_begin_; }
    LogicalPartition::byte_type LogicalPartition::target_bytes() const noexcept { return target_bytes_; }
    LogicalPartition::byte_type LogicalPartition::edge_data_bytes() const noexcept
    {
        return static_cast<byte_type>(edge_count()) * sizeof(CSRGraph::vertex_id);
    }

    LogicalPartitioner::LogicalPartitioner(byte_type partition_bytes) : partition_bytes_(partition_bytes)
    {
        if (partition_bytes_ == 0U)
            throw std::invalid_argument("LogicalPartitioner: partition_bytes must be greater than zero");
    }

    LogicalPartitioner::byte_type LogicalPartitioner::partition_bytes() const noexcept { return partition_bytes_; }

    std::vector<LogicalPartition> LogicalPartitioner::partition(const CSRGraph &graph) const
    {
        const auto &off = graph.row_offsets();
        const auto n = graph.num_vertices();
        if (n != 0U)
        {
            if (off.size() != static_cast<std::size_t>(n) + 1U)
                throw std::invalid_argument("LogicalPartitioner: invalid CSR row-offset count");
            if (off.back() != graph.num_edges())
                throw std::invalid_argument("LogicalPartitioner: CSR terminal offset does not match the number of edges");
        }
        std::vector<LogicalPartition> parts;
        if (n == 0U)
            return parts;
        // Close the running range before vertex v when including v's row would
        // take its column bytes past the target; a lone oversized row stays whole.
        vertex_id begin = 0;
        offset_type edge_begin = 0;
        const auto V = static_cast<vertex_id>(n);
        for (vertex_id v = 0; v < V; ++v)
        {
            const offset_type bytes = (off[static_cast<std::size_t>(v) + 1] - edge_begin) * sizeof(CSRGraph::vertex_id);
            if (v > begin && bytes > partition_bytes_)
            {
                parts.emplace_back(begin, v, edge_begin, off[v], partition_bytes_);
                begin = v;
                edge_begin = off[v];
            }
        }
        parts.emplace_back(begin, V, edge_begin, off[n], partition_bytes_);
        return parts;
    }

    // ------------------------------------------------------------ activity
    ActivityTracker::ActivityTracker(count_type vertex_count)
        : vertex_count_(vertex_count), active_(static_cast<std::size_t>(vertex_count), 0U)
    {
    }
    ActivityTracker::count_type ActivityTracker::vertex_count() const noexcept { return vertex_count_; }
    void ActivityTracker::clear() noexcept
    {
        std::fill(active_.begin(), active_.end(), 0U);
        active_vertex_count_ = 0U;
    }
    void ActivityTracker::set_active(vertex_id vertex, bool active)
    {
        if (static_cast<count_type>(vertex) >= vertex_count_)
            throw std::out_of_range("activity vertex is out of range");
        unsigned char &flag = active_[vertex];
        if ((flag != 0U) == active)
            return;
        flag = active ? 1U : 0U;
        if (active)
            ++active_vertex_count_;
        else
            --active_vertex_count_;
    }
    bool ActivityTracker::is_active(vertex_id vertex) const
    {
        if (static_cast<count_type>(vertex) >= vertex_count_)
            throw std::out_of_range("activity vertex is out of range");
        return active_[vertex] != 0U;
    }
    void ActivityTracker::set_active_vertices(const std::vector<vertex_id> &vertices)
    {
        clear();
        for (const vertex_id v : vertices)
            set_active(v, true);
    }
    std::vector<ActivityTracker::vertex_id> ActivityTracker::active_vertices() const
    {
        std::vector<vertex_id> out;
        out.reserve(static_cast<std::size_t>(active_vertex_count_));
        for (count_type v = 0; v < vertex_count_; ++v)
            if (active_[static_cast<std::size_t>(v)] != 0U)
                out.push_back(static_cast<vertex_id>(v));
        return out;
    }
    ActivityTracker::count_type ActivityTracker::active_vertex_count() const noexcept { return active_vertex_count_; }

    ActivityTracker::count_type ActivityTracker::active_edge_count(const CSRGraph &graph) const
    {
        if (graph.num_vertices() != vertex_count_)
            throw std::invalid_argument("activity tracker vertex count does not match graph");
        const auto &off = graph.row_offsets();
        count_type edges = 0;
        for (count_type v = 0; v < vertex_count_; ++v)
            if (active_[static_cast<std::size_t>(v)] != 0U)
                edges += off[v + 1] - off[v];
        return edges;
    }

    std::vector<ActivityTracker::PartitionActivity>
    ActivityTracker::partition_statistics(const CSRGraph &graph, const std::vector<VertexRange> &partitions) const
    {
        if (graph.num_vertices() != vertex_count_)
            throw std::invalid_argument("activity tracker vertex count does not match graph");
        const auto &off = graph.row_offsets();
        std::vector<PartitionActivity> stats;
        stats.reserve(partitions.size());
        for (const VertexRange &r : partitions)
        {
            if (r.begin > r.end)
                throw std::invalid_argument("activity partition range has begin greater than end");
            if (static_cast<count_type>(r.end) > vertex_count_)
                throw std::out_of_range("activity partition range exceeds vertex count");
            PartitionActivity a;
            for (count_type v = r.begin; v < r.end; ++v)
            {
                const count_type deg = off[v + 1] - off[v];
                a.total_edges += deg;
                if (active_[static_cast<std::size_t>(v)] != 0U)
                {
                    ++a.active_vertices;
                    a.active_edges += deg;
                }
            }
            stats.push_back(a);
        }
        return stats;
    }
} // namespace hytgraph::graph
