// Contribution ordering of work items: the core of the original
// ContributionScheduler::schedule (src/scheduling/contribution_scheduler.cpp:
// 23-121) - a stable sort by descending contribution, equal contributions by
// ascending item index when the tie-break is deterministic - shared by the
// C++ ContributionScheduler and the hybrid engine, which orders a round's
// filter runs by device-computed partition contributions with it.
#pragma once

#include <algorithm>
#include <cstddef>
#include <numeric>
#include <vector>

namespace hyt
{
    struct ContributionItem
    {
        std::size_t item = 0;
        double contribution = 0.0;
    };

    // Positions of `items` in scheduled order; `reordered` = positions whose
    // item moved.
    inline std::vector<std::size_t> contribution_order(const std::vector<ContributionItem> &items,
                                                       bool deterministic_tie_break, std::size_t *reordered)
    {
        std::vector<std::size_t> order(items.size());
        std::iota(order.begin(), order.end(), std::size_t{0});
        std::stable_sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) {
            const double ca = items[a].contribution, cb = items[b].contribution;
            if (ca != cb)
                return ca > cb;
            return deterministic_tie_break && items[a].item < items[b].item;
        });
        if (reordered)
        {
            std::size_t n = 0;
            for (std::size_t pos = 0; pos < order.size(); ++pos)
                n += order[pos] != pos ? 1U : 0U;
            *reordered = n;
        }
        return order;
    }
} // namespace hyt
