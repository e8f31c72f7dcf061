// Contribution-driven ordering of work items (interface of the original
// include/scheduling/contribution_scheduler.hpp:10-207): items sorted by
// descending contribution (stable; equal contributions by ascending item
// index when deterministic), with counts of reordered, zero-contribution
// and duplicate items, and an overhead estimate of log2(N).
#pragma once

#include <cstddef>
#include <vector>

namespace hytgraph::scheduling
{
    enum class ContributionSchedulingPolicy
    {
        PageRankDelta,
        SSSPContribution
    };

    struct ContributionPriority
    {
        std::size_t item_index = 0U;
        double contribution = 0.0;
    };

    struct ContributionSchedulingMetrics
    {
        std::size_t input_item_count = 0U;
        std::size_t scheduled_item_count = 0U;
        std::size_t reordered_item_count = 0U;
        std::size_t zero_contribution_count = 0U;
        double scheduling_overhead = 0.0;
        std::size_t stale_work_count = 0U;
        std::size_t redundant_work_count = 0U;
    };

    struct ContributionSchedulingPlan
    {
        std::vector<std::size_t> ordered_item_indices;
        ContributionSchedulingMetrics metrics;

        [[nodiscard]] std::size_t size() const noexcept { return ordered_item_indices.size(); }
        [[nodiscard]] bool empty() const noexcept { return ordered_item_indices.empty(); }
        [[nodiscard]] bool reordered() const noexcept { return metrics.reordered_item_count != 0U; }
    };

    struct ContributionSchedulerOptions
    {
        ContributionSchedulingPolicy policy = ContributionSchedulingPolicy::PageRankDelta;
        bool deterministic_tie_break = true;
        double zero_contribution_epsilon = 0.0;
    };

    struct ContributionWorkObservation
    {
        std::size_t item_index = 0U;
        bool stale = false;
    };

    [[nodiscard]] std::vector<ContributionPriority>
    make_contribution_priorities(const std::vector<std::size_t> &item_indices, const std::vector<double> &contributions);

    class ContributionScheduler
    {
    public:
        explicit ContributionScheduler(ContributionSchedulerOptions options = {});

        [[nodiscard]] const ContributionSchedulerOptions &options() const noexcept;

        [[nodiscard]] ContributionSchedulingPlan schedule(const std::vector<ContributionPriority> &priorities) const;

        [[nodiscard]] static ContributionSchedulingPlan
        synchronous_reference(const std::vector<ContributionPriority> &priorities);

        [[nodiscard]] static ContributionSchedulingMetrics
        measure_work_metrics(const std::vector<ContributionPriority> &priorities);

        [[nodiscard]] static ContributionSchedulingMetrics
        measure_work_metrics(const std::vector<ContributionPriority> &priorities,
                             const std::vector<ContributionWorkObservation> &observations);

        [[nodiscard]] std::vector<ContributionPriority>
        make_contribution_priorities(const std::vector<std::size_t> &item_indices,
                                     const std::vector<double> &contributions);

        [[nodiscard]] static double estimate_scheduling_overhead(const std::vector<ContributionPriority> &priorities);

    private:
        ContributionSchedulerOptions options_;
    };
} // namespace hytgraph::scheduling
