// HubSorter - hub-first vertex ordering with the interface of the original
// project's include/scheduling/hub_sort.hpp: score H(v) = Do(v) Di(v) /
// (Do_max Di_max), the top hub_fraction of vertices (rounded up unless the
// product is integral) first by descending score then ascending id, every
// other vertex after them in its original order. The ordering is applied
// with CSRGraph::reorder_vertices.
#pragma once

#include "graph/csr_graph.hpp"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace hytgraph::scheduling
{
    struct HubSortResult
    {
        // new position -> original vertex; the first hub_count are the hubs
        std::vector<graph::CSRGraph::vertex_id> vertex_order;
        // score of every original vertex (0 for every vertex of an edgeless graph)
        std::vector<double> scores;
        std::size_t hub_count = 0U;

        [[nodiscard]] bool empty() const noexcept { return vertex_order.empty(); }
    };

    struct HubSortOptions
    {
        double hub_fraction = 0.08;
    };

    class HubSorter
    {
    public:
        // std::invalid_argument unless hub_fraction is finite and in [0, 1]
        explicit HubSorter(HubSortOptions options = {});

        [[nodiscard]] const HubSortOptions &options() const noexcept;

        [[nodiscard]] HubSortResult sort(const graph::CSRGraph &graph) const;

        // The same result computed on the GPU (addition): scores bit for bit
        // (the long double division is reproduced exactly), same order.
        [[nodiscard]] HubSortResult sort_gpu(const graph::CSRGraph &graph) const;

    private:
        HubSortOptions options_;
    };
} // namespace hytgraph::scheduling
