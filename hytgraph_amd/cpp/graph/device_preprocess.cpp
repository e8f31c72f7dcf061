// GPU forms of the host preprocessing steps (additions to the original's
// interface): HubSorter::sort_gpu (original src/scheduling/hub_sort.cpp:28-183)
// and reorder_vertices_gpu (src/graph/csr_graph.cpp:171-283), over the C ABI
// (hyt_hub_sort, hyt_graph_reorder, hyt_graph_download). Each call uploads the
// graph, runs the device kernels and, for the reorder, downloads the result.
#include "../algorithms/gpu_call.hpp"
#include "graph/csr_graph.hpp"
#include "hytgraph_amd.h"
#include "scheduling/hub_sort.hpp"

#include <cstring>
#include <stdexcept>
#include <vector>

namespace hytgraph::scheduling
{
    HubSortResult HubSorter::sort_gpu(const graph::CSRGraph &graph) const
    {
        namespace d = algorithms::detail;
        HubSortResult r;
        const auto n = static_cast<std::size_t>(graph.num_vertices());
        if (n == 0)
            return r;
        d::require_device("HubSorter::sort_gpu");
        d::ScopedGraph sg;
        d::upload(graph, false, sg, "HubSorter::sort_gpu");
        r.vertex_order.resize(n);
        r.scores.resize(n);
        std::uint64_t hubs = 0;
        d::check(hyt_hub_sort(sg.g, options_.hub_fraction, r.vertex_order.data(), r.scores.data(), &hubs, 0),
                 "HubSorter::sort_gpu");
        r.hub_count = static_cast<std::size_t>(hubs);
        return r;
    }
} // namespace hytgraph::scheduling

namespace hytgraph::graph
{
    void reorder_vertices_gpu(CSRGraph &graph, const std::vector<CSRGraph::vertex_id> &vertex_order)
    {
        namespace d = algorithms::detail;
        const auto n = static_cast<std::size_t>(graph.num_vertices());
        if (vertex_order.size() != n)
            throw std::invalid_argument(
                "CSRGraph::reorder_vertices: vertex_order must contain exactly one entry per vertex");
        if (n == 0)
            return;
        d::require_device("reorder_vertices_gpu");
        // weights travel as bit patterns (u32 mode: carried, not validated -
        // the original's reorder accepts any weights)
        const auto E = static_cast<std::size_t>(graph.num_edges());
        std::vector<std::uint32_t> wbits(graph.has_weights() ? E : 0);
        if (!wbits.empty())
            std::memcpy(wbits.data(), graph.edge_weights().data(), E * sizeof(float));
        d::ScopedGraph sg;
        d::check(hyt_graph_upload(graph.num_vertices(), graph.num_edges(), graph.row_offsets().data(),
                                  graph.column_indices().data(), wbits.empty() ? nullptr : wbits.data(),
                                  HYT_GRAPH_WEIGHTS_U32, &sg.g),
                 "reorder_vertices_gpu");
        d::check(hyt_graph_reorder(sg.g, vertex_order.data(), 0), "reorder_vertices_gpu");
        std::vector<CSRGraph::offset_type> off(n + 1);
        std::vector<CSRGraph::vertex_id> col(E);
        std::vector<std::uint32_t> bits(graph.has_weights() ? E : 0);
        d::check(hyt_graph_download(sg.g, off.data(), col.data(), graph.has_weights() ? bits.data() : nullptr),
                 "reorder_vertices_gpu");
        std::vector<CSRGraph::weight_type> w(bits.size());
        if (!bits.empty())
            std::memcpy(w.data(), bits.data(), bits.size() * sizeof(float));
        graph = CSRGraph(graph.num_vertices(), std::move(off), std::move(col), std::move(w));
    }
} // namespace hytgraph::graph
