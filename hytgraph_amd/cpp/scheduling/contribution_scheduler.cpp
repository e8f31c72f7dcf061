// ContributionScheduler with the ordering and metrics of the original
// src/scheduling/contribution_scheduler.cpp:23-246.
#include "scheduling/contribution_scheduler.hpp"

#include "contribution_order.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace hytgraph::scheduling
{
    namespace
    {
        std::size_t duplicate_items(const std::vector<ContributionPriority> &p)
        {
            std::vector<std::size_t> ids(p.size());
            for (std::size_t i = 0; i < p.size(); ++i)
                ids[i] = p[i].item_index;
            std::sort(ids.begin(), ids.end());
            std::size_t n = 0;
            for (std::size_t i = 1; i < ids.size(); ++i)
                n += ids[i] == ids[i - 1] ? 1U : 0U;
            return n;
        }
    } // namespace

    std::vector<ContributionPriority> make_contribution_priorities(const std::vector<std::size_t> &item_indices,
                                                                   const std::vector<double> &contributions)
    {
        if (item_indices.size() != contributions.size())
            throw std::invalid_argument("contribution priority inputs must have equal size");
        std::vector<ContributionPriority> out(item_indices.size());
        for (std::size_t i = 0; i < out.size(); ++i)
            out[i] = ContributionPriority{item_indices[i], contributions[i]};
        return out;
    }

    ContributionScheduler::ContributionScheduler(ContributionSchedulerOptions options) : options_(options) {}

    const ContributionSchedulerOptions &ContributionScheduler::options() const noexcept { return options_; }

    ContributionSchedulingPlan ContributionScheduler::schedule(const std::vector<ContributionPriority> &priorities) const
    {
        ContributionSchedulingPlan plan;
        plan.metrics.input_item_count = priorities.size();
        plan.metrics.scheduled_item_count = priorities.size();
        plan.metrics.redundant_work_count = duplicate_items(priorities);
        for (const auto &p : priorities)
            if (!std::isfinite(p.contribution))
                throw std::invalid_argument("contribution priority must be finite");
        for (const auto &p : priorities)
            if (std::abs(p.contribution) <= options_.zero_contribution_epsilon)
                ++plan.metrics.zero_contribution_count;
        std::vector<hyt::ContributionItem> items(priorities.size());
        for (std::size_t i = 0; i < items.size(); ++i)
            items[i] = hyt::ContributionItem{priorities[i].item_index, priorities[i].contribution};
        const std::vector<std::size_t> order =
            hyt::contribution_order(items, options_.deterministic_tie_break, &plan.metrics.reordered_item_count);
        plan.ordered_item_indices.reserve(order.size());
        for (const std::size_t pos : order)
            plan.ordered_item_indices.push_back(priorities[pos].item_index);
        plan.metrics.scheduling_overhead = estimate_scheduling_overhead(priorities);
        return plan;
    }

    ContributionSchedulingPlan
    ContributionScheduler::synchronous_reference(const std::vector<ContributionPriority> &priorities)
    {
        ContributionSchedulingPlan plan;
        plan.metrics.input_item_count = priorities.size();
        plan.metrics.scheduled_item_count = priorities.size();
        plan.ordered_item_indices.reserve(priorities.size());
        for (const auto &p : priorities)
        {
            plan.ordered_item_indices.push_back(p.item_index);
            if (p.contribution == 0.0)
                ++plan.metrics.zero_contribution_count;
        }
        return plan;
    }

    ContributionSchedulingMetrics
    ContributionScheduler::measure_work_metrics(const std::vector<ContributionPriority> &priorities)
    {
        ContributionSchedulingMetrics m;
        m.input_item_count = priorities.size();
        m.scheduled_item_count = priorities.size();
        for (const auto &p : priorities)
            if (p.contribution == 0.0)
                ++m.zero_contribution_count;
        m.redundant_work_count = duplicate_items(priorities);
        return m;
    }

    ContributionSchedulingMetrics
    ContributionScheduler::measure_work_metrics(const std::vector<ContributionPriority> &priorities,
                                                const std::vector<ContributionWorkObservation> &observations)
    {
        ContributionSchedulingMetrics m = measure_work_metrics(priorities);
        for (const auto &o : observations)
            if (o.stale)
                ++m.stale_work_count;
        return m;
    }

    std::vector<ContributionPriority>
    ContributionScheduler::make_contribution_priorities(const std::vector<std::size_t> &item_indices,
                                                        const std::vector<double> &contributions)
    {
        return scheduling::make_contribution_priorities(item_indices, contributions);
    }

    double ContributionScheduler::estimate_scheduling_overhead(const std::vector<ContributionPriority> &priorities)
    {
        return priorities.empty() ? 0.0 : std::log2(static_cast<double>(priorities.size()));
    }
} // namespace hytgraph::scheduling
