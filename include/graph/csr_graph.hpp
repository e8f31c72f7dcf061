// CSRGraph - host CSR graph with the interface of the original project's
// include/graph/csr_graph.hpp:11-69 (same names, types, accessor semantics
// and exceptions), so code written against it compiles unchanged.
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace hytgraph::graph
{
    class CSRGraph
    {
    public:
        using vertex_id = std::uint32_t;
        using offset_type = std::uint64_t;
        using weight_type = float;

        CSRGraph() = default;

        // Validates on construction (std::invalid_argument when malformed).
        CSRGraph(offset_type num_vertices,
                 std::vector<offset_type> row_offsets,
                 std::vector<vertex_id> column_indices,
                 std::vector<weight_type> edge_weights = {});

        [[nodiscard]] offset_type num_vertices() const noexcept;
        [[nodiscard]] offset_type num_edges() const noexcept;
        [[nodiscard]] bool has_weights() const noexcept;

        // Out-degree of a vertex (std::out_of_range for a bad id).
        [[nodiscard]] offset_type out_degree(vertex_id vertex) const;

        // Half-open edge range [begin, end) of a vertex's neighbours.
        [[nodiscard]] std::pair<offset_type, offset_type> neighbor_range(vertex_id vertex) const;

        [[nodiscard]] vertex_id neighbor_at(offset_type edge_index) const;
        [[nodiscard]] weight_type weight_at(offset_type edge_index) const;

        [[nodiscard]] const std::vector<offset_type> &row_offsets() const noexcept;
        [[nodiscard]] const std::vector<vertex_id> &column_indices() const noexcept;
        [[nodiscard]] const std::vector<weight_type> &edge_weights() const noexcept;

        // Throws std::invalid_argument when an invariant does not hold.
        void validate() const;

        // vertex_order[new_vertex] = old_vertex; adjacency lists move with
        // their vertex, destinations are renamed, weights stay with edges.
        void reorder_vertices(const std::vector<vertex_id> &vertex_order);

    private:
        offset_type num_vertices_ = 0;
        std::vector<offset_type> row_offsets_;
        std::vector<vertex_id> column_indices_;
        std::vector<weight_type> edge_weights_;
    };

    // CSRGraph::reorder_vertices computed on the GPU (addition): the same
    // graph afterwards, and the same std::invalid_argument for an order that
    // is not a permutation (the graph is then unchanged).
    void reorder_vertices_gpu(CSRGraph &graph, const std::vector<CSRGraph::vertex_id> &vertex_order);
} // namespace hytgraph::graph
