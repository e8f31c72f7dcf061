// Hub-first ordering (original: src/scheduling/hub_sort.cpp:28-183). Scores
// in long double, one pass for in-degrees, and only the hubs are ranked
// (nth_element, then a sort of the hub prefix), which gives the same order
// as ranking every vertex because (score desc, id asc) is a total order.
#include "scheduling/hub_sort.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace hytgraph::scheduling
{
    HubSorter::HubSorter(HubSortOptions options) : options_(options)
    {
        if (!std::isfinite(options_.hub_fraction) || options_.hub_fraction < 0.0 || options_.hub_fraction > 1.0)
            throw std::invalid_argument("HubSortOptions.hub_fraction must be finite and in [0, 1]");
    }

    const HubSortOptions &HubSorter::options() const noexcept { return options_; }

    HubSortResult HubSorter::sort(const graph::CSRGraph &graph) const
    {
        using vid = graph::CSRGraph::vertex_id;
        const std::size_t n = static_cast<std::size_t>(graph.num_vertices());
        HubSortResult r;
        r.scores.assign(n, 0.0);
        if (n == 0)
            return r;
        const auto &off = graph.row_offsets();
        const auto &col = graph.column_indices();
        std::vector<std::uint64_t> in_deg(n, 0);
        for (const vid v : col)
            ++in_deg[v];
        std::uint64_t max_out = 0, max_in = 0;
        for (std::size_t v = 0; v < n; ++v)
        {
            max_out = std::max<std::uint64_t>(max_out, off[v + 1] - off[v]);
            max_in = std::max(max_in, in_deg[v]);
        }
        const long double denom = static_cast<long double>(max_out) * static_cast<long double>(max_in);
        if (denom > 0.0L)
            for (std::size_t v = 0; v < n; ++v)
                r.scores[v] = static_cast<double>(static_cast<long double>(off[v + 1] - off[v]) *
                                                  static_cast<long double>(in_deg[v]) / denom);

        // hubs = n * fraction, rounded up unless it is an integer (to 1e-12)
        const long double want = static_cast<long double>(n) * static_cast<long double>(options_.hub_fraction);
        const long double nearest = std::round(want);
        const long double count = std::fabs(want - nearest) < 1e-12L ? nearest : std::ceil(want);
        r.hub_count = std::min(n, static_cast<std::size_t>(count));

        std::vector<vid> rank(n);
        for (std::size_t v = 0; v < n; ++v)
            rank[v] = static_cast<vid>(v);
        const auto before = [&](vid a, vid b) {
            return r.scores[a] != r.scores[b] ? r.scores[a] > r.scores[b] : a < b;
        };
        const auto hub_end = rank.begin() + static_cast<std::ptrdiff_t>(r.hub_count);
        if (r.hub_count < n)
            std::nth_element(rank.begin(), hub_end, rank.end(), before);
        std::sort(rank.begin(), hub_end, before);

        std::vector<bool> hub(n, false);
        r.vertex_order.reserve(n);
        for (auto it = rank.begin(); it != hub_end; ++it)
        {
            hub[*it] = true;
            r.vertex_order.push_back(*it);
        }
        for (std::size_t v = 0; v < n; ++v)
            if (!hub[v])
                r.vertex_order.push_back(static_cast<vid>(v));
        return r;
    }
} // namespace hytgraph::scheduling
