// CSRGraph: host CSR with the semantics of the original
// src/graph/csr_graph.cpp:11-283 (validation rules, exception types,
// reorder_vertices permutation convention).
#include "graph/csr_graph.hpp"

#include <stdexcept>
#include <utility>

namespace hytgraph::graph
{
    CSRGraph::CSRGraph(offset_type num_vertices, std::vector<offset_type> row_offsets,
                       std::vector<vertex_id> column_indices, std::vector<weight_type> edge_weights)
        : num_vertices_(num_vertices), row_offsets_(std::move(row_offsets)),
          column_indices_(std::move(column_indices)), edge_weights_(std::move(edge_weights))
    {
        validate();
    }

    CSRGraph::offset_type CSRGraph::num_vertices() const noexcept { return num_vertices_; }

    CSRGraph::offset_type CSRGraph::num_edges() const noexcept
    {
        return static_cast<offset_type>(column_indices_.size());
    }

    bool CSRGraph::has_weights() const noexcept { return !edge_weights_.empty(); }

    CSRGraph::offset_type CSRGraph::out_degree(vertex_id vertex) const
    {
        const auto r = neighbor_range(vertex);
        return r.second - r.first;
    }

    std::pair<CSRGraph::offset_type, CSRGraph::offset_type> CSRGraph::neighbor_range(vertex_id vertex) const
    {
        if (static_cast<offset_type>(vertex) >= num_vertices_)
            throw std::out_of_range("CSRGraph::neighbor_range: vertex ID is out of range");
        return {row_offsets_[vertex], row_offsets_[static_cast<std::size_t>(vertex) + 1U]};
    }

    CSRGraph::vertex_id CSRGraph::neighbor_at(offset_type edge_index) const
    {
        if (edge_index >= num_edges())
            throw std::out_of_range("CSRGraph::neighbor_at: edge index is out of range");
        return column_indices_[static_cast<std::size_t>(edge_index)];
    }

    CSRGraph::weight_type CSRGraph::weight_at(offset_type edge_index) const
    {
        if (!has_weights())
            throw std::logic_error("CSRGraph::weight_at: graph does not contain edge weights");
        if (edge_index >= num_edges())
            throw std::out_of_range("CSRGraph::weight_at: edge index is out of range");
        return edge_weights_[static_cast<std::size_t>(edge_index)];
    }

    const std::vector<CSRGraph::offset_type> &CSRGraph::row_offsets() const noexcept { return row_offsets_; }
    const std::vector<CSRGraph::vertex_id> &CSRGraph::column_indices() const noexcept { return column_indices_; }
    const std::vector<CSRGraph::weight_type> &CSRGraph::edge_weights() const noexcept { return edge_weights_; }

    void CSRGraph::validate() const
    {
        if (row_offsets_.size() != static_cast<std::size_t>(num_vertices_) + 1U)
            throw std::invalid_argument("CSRGraph::validate: row_offsets must contain num_vertices + 1 entries");
        if (row_offsets_.front() != 0)
            throw std::invalid_argument("CSRGraph::validate: first row offset must be zero");
        for (std::size_t i = 1; i < row_offsets_.size(); ++i)
            if (row_offsets_[i] < row_offsets_[i - 1])
                throw std::invalid_argument("CSRGraph::validate: row offsets must be non-decreasing");
        if (row_offsets_.back() != num_edges())
            throw std::invalid_argument("CSRGraph::validate: terminal row offset must equal number of edges");
        for (const vertex_id v : column_indices_)
            if (static_cast<offset_type>(v) >= num_vertices_)
                throw std::invalid_argument(
                    "CSRGraph::validate: column index contains an invalid destination vertex ID");
        if (!edge_weights_.empty() && edge_weights_.size() != column_indices_.size())
            throw std::invalid_argument("CSRGraph::validate: edge_weights must contain exactly one weight per edge");
    }

    void CSRGraph::reorder_vertices(const std::vector<vertex_id> &vertex_order)
    {
        const auto n = static_cast<std::size_t>(num_vertices_);
        if (vertex_order.size() != n)
            throw std::invalid_argument(
                "CSRGraph::reorder_vertices: vertex_order must contain exactly one entry per vertex");
        // inverse permutation, validating range and uniqueness first
        std::vector<vertex_id> renamed(n);
        std::vector<unsigned char> seen(n, 0);
        for (std::size_t nv = 0; nv < n; ++nv)
        {
            const auto old = static_cast<std::size_t>(vertex_order[nv]);
            if (old >= n)
                throw std::invalid_argument(
                    "CSRGraph::reorder_vertices: vertex_order contains an out-of-range vertex ID");
            if (seen[old])
                throw std::invalid_argument("CSRGraph::reorder_vertices: vertex_order contains duplicate vertex IDs");
            seen[old] = 1;
            renamed[old] = static_cast<vertex_id>(nv);
        }
        std::vector<offset_type> off(n + 1, 0);
        for (std::size_t nv = 0; nv < n; ++nv)
        {
            const auto old = static_cast<std::size_t>(vertex_order[nv]);
            off[nv + 1] = off[nv] + (row_offsets_[old + 1] - row_offsets_[old]);
        }
        std::vector<vertex_id> col(column_indices_.size());
        std::vector<weight_type> w(edge_weights_.size());
        for (std::size_t nv = 0; nv < n; ++nv)
        {
            const auto old = static_cast<std::size_t>(vertex_order[nv]);
            auto dst = off[nv];
            for (auto e = row_offsets_[old]; e < row_offsets_[old + 1]; ++e, ++dst)
            {
                col[dst] = renamed[column_indices_[e]];
                if (!w.empty())
                    w[dst] = edge_weights_[e];
            }
        }
        row_offsets_ = std::move(off);
        column_indices_ = std::move(col);
        edge_weights_ = std::move(w);
        validate();
    }
} // namespace hytgraph::graph
