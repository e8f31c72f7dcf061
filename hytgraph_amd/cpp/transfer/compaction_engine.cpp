// ExpTMCompaction with the output format and checks of the original
// src/transfer/compaction_engine.cpp:253-493. One pass over the ascending
// active-vertex list serves every partition (the partitions are contiguous
// ascending vertex ranges), so the cost is O(V + active edges) rather than
// a rescan of the active list per partition.
#include "transfer/compaction_engine.hpp"

#include <algorithm>
#include <chrono>
#include <limits>
#include <stdexcept>
#include <utility>

namespace hytgraph::transfer
{
    namespace
    {
        using VertexId = graph::CSRGraph::vertex_id;
        using Offset = graph::CSRGraph::offset_type;

        void check_partition(const graph::CSRGraph &g, const graph::LogicalPartition &p)
        {
            const VertexId V = static_cast<VertexId>(g.num_vertices());
            if (p.vertex_begin() > p.vertex_end())
                throw std::invalid_argument("ExpTMCompaction: partition vertex range is invalid");
            if (p.vertex_end() > V)
                throw std::invalid_argument("ExpTMCompaction: partition vertex range exceeds graph");
            if (p.edge_begin() > p.edge_end())
                throw std::invalid_argument("ExpTMCompaction: partition edge range is invalid");
            if (p.edge_end() > g.num_edges())
                throw std::invalid_argument("ExpTMCompaction: partition edge range exceeds graph");
            if (p.vertex_begin() < V && g.neighbor_range(p.vertex_begin()).first != p.edge_begin())
                throw std::invalid_argument("ExpTMCompaction: partition edge_begin does not match its first CSR row");
            if (p.vertex_end() > p.vertex_begin() && g.neighbor_range(p.vertex_end() - 1U).second != p.edge_end())
                throw std::invalid_argument("ExpTMCompaction: partition edge_end does not match its final CSR row");
        }

        template <class T>
        T saturating_add(T a, T b)
        {
            return b > std::numeric_limits<T>::max() - a ? std::numeric_limits<T>::max() : a + b;
        }
    } // namespace

    Offset CompactedPartition::active_edge_count() const noexcept
    {
        return static_cast<Offset>(neighbors.size());
    }
    std::size_t CompactedPartition::active_vertex_count() const noexcept { return active_vertices.size(); }
    std::size_t CompactedPartition::neighbor_bytes() const noexcept { return neighbors.size() * sizeof(VertexId); }
    std::size_t CompactedPartition::index_bytes() const noexcept { return neighbor_index.size() * sizeof(Offset); }
    std::size_t CompactedPartition::total_bytes() const noexcept
    {
        return saturating_add(neighbor_bytes(), index_bytes());
    }

    std::size_t CompactionResult::compacted_partition_count() const noexcept { return partitions.size(); }
    Offset CompactionResult::active_edge_count() const noexcept
    {
        Offset t = 0;
        for (const auto &p : partitions)
            t = saturating_add(t, p.active_edge_count());
        return t;
    }
    Offset CompactionResult::active_vertex_count() const noexcept
    {
        Offset t = 0;
        for (const auto &p : partitions)
            t = saturating_add(t, static_cast<Offset>(p.active_vertex_count()));
        return t;
    }
    std::size_t CompactionResult::neighbor_bytes() const noexcept
    {
        std::size_t t = 0;
        for (const auto &p : partitions)
            t = saturating_add(t, p.neighbor_bytes());
        return t;
    }
    std::size_t CompactionResult::index_bytes() const noexcept
    {
        std::size_t t = 0;
        for (const auto &p : partitions)
            t = saturating_add(t, p.index_bytes());
        return t;
    }
    std::size_t CompactionResult::total_bytes() const noexcept
    {
        return saturating_add(neighbor_bytes(), index_bytes());
    }

    ExpTMCompaction::ExpTMCompaction(bool enabled) noexcept : enabled_(enabled) {}

    bool ExpTMCompaction::enabled() const noexcept { return enabled_; }

    CompactionResult ExpTMCompaction::compact(const graph::CSRGraph &graph,
                                              const std::vector<graph::LogicalPartition> &partitions,
                                              const graph::ActivityTracker &activity) const
    {
        CompactionResult out;
        if (!enabled_)
            return out;
        if (activity.vertex_count() != graph.num_vertices())
            throw std::invalid_argument("ExpTMCompaction: activity tracker vertex count does not match graph");
        const auto t0 = std::chrono::steady_clock::now();
        const std::vector<VertexId> active = activity.active_vertices(); // ascending
        const auto &off = graph.row_offsets();
        const auto &col = graph.column_indices();
        for (std::size_t pi = 0; pi < partitions.size(); ++pi)
        {
            const graph::LogicalPartition &p = partitions[pi];
            check_partition(graph, p);
            const auto lo = std::lower_bound(active.begin(), active.end(), p.vertex_begin());
            const auto hi = std::lower_bound(lo, active.end(), p.vertex_end());
            if (lo == hi)
                continue;
            CompactedPartition c;
            c.partition_index = pi;
            c.vertex_begin = p.vertex_begin();
            c.vertex_end = p.vertex_end();
            c.active_vertices.assign(lo, hi);
            Offset total = 0;
            for (VertexId v : c.active_vertices)
            {
                const Offset d = off[v + 1] - off[v];
                if (d > std::numeric_limits<Offset>::max() - total)
                    throw std::overflow_error("ExpTMCompaction: active-edge count overflow");
                total += d;
            }
            c.neighbors.resize(static_cast<std::size_t>(total));
            c.neighbor_index.resize(c.active_vertices.size() + 1U);
            c.neighbor_index[0] = 0;
            Offset at = 0;
            for (std::size_t i = 0; i < c.active_vertices.size(); ++i)
            {
                const VertexId v = c.active_vertices[i];
                const Offset b = off[v], e = off[v + 1];
                std::copy(col.begin() + static_cast<std::ptrdiff_t>(b), col.begin() + static_cast<std::ptrdiff_t>(e),
                          c.neighbors.begin() + static_cast<std::ptrdiff_t>(at));
                at += e - b;
                c.neighbor_index[i + 1] = at;
            }
            out.partitions.push_back(std::move(c));
        }
        out.compaction_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return out;
    }

    std::vector<VertexId> ExpTMCompaction::expanded_neighbors(const CompactedPartition &p)
    {
        if (p.neighbor_index.size() != p.active_vertices.size() + 1U)
            throw std::invalid_argument(
                "ExpTMCompaction::expanded_neighbors: compressed index size does not match active vertex count");
        if (p.neighbor_index.front() != 0U)
            throw std::invalid_argument("ExpTMCompaction::expanded_neighbors: index must start at zero");
        if (p.neighbor_index.back() != static_cast<Offset>(p.neighbors.size()))
            throw std::invalid_argument(
                "ExpTMCompaction::expanded_neighbors: terminal index does not match neighbor count");
        for (std::size_t i = 1; i < p.neighbor_index.size(); ++i)
            if (p.neighbor_index[i] < p.neighbor_index[i - 1])
                throw std::invalid_argument("ExpTMCompaction::expanded_neighbors: index is not monotonic");
        return p.neighbors;
    }
} // namespace hytgraph::transfer
