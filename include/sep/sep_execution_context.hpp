// Handles a SEP driver runs against (interface of the original
// include/sep/sep_execution_context.hpp:12-192): opaque graph / activity /
// partition / task handles and an optional frontier. Data-driven variants
// need a frontier.
#pragma once

#include "sep_execution_config.hpp"
#include "sep_frontier.hpp"

namespace hytgraph::sep
{
    class SEPExecutionContext
    {
    public:
        using graph_handle_type = void;
        using activity_handle_type = void;
        using partition_handle_type = void;
        using task_handle_type = void;

        SEPExecutionContext() noexcept = default;
        SEPExecutionContext(graph_handle_type *graph, activity_handle_type *activity = nullptr,
                            partition_handle_type *partitions = nullptr, task_handle_type *tasks = nullptr,
                            SEPFrontier *frontier = nullptr) noexcept
            : graph_(graph), activity_(activity), partitions_(partitions), tasks_(tasks), frontier_(frontier)
        {
        }

        graph_handle_type *graph() noexcept { return graph_; }
        const graph_handle_type *graph() const noexcept { return graph_; }
        activity_handle_type *activity() noexcept { return activity_; }
        const activity_handle_type *activity() const noexcept { return activity_; }
        partition_handle_type *partitions() noexcept { return partitions_; }
        const partition_handle_type *partitions() const noexcept { return partitions_; }
        task_handle_type *tasks() noexcept { return tasks_; }
        const task_handle_type *tasks() const noexcept { return tasks_; }
        SEPFrontier *frontier() noexcept { return frontier_; }
        const SEPFrontier *frontier() const noexcept { return frontier_; }

        bool has_graph() const noexcept { return graph_ != nullptr; }
        bool has_activity() const noexcept { return activity_ != nullptr; }
        bool has_partitions() const noexcept { return partitions_ != nullptr; }
        bool has_tasks() const noexcept { return tasks_ != nullptr; }
        bool has_frontier() const noexcept { return frontier_ != nullptr; }

        bool satisfies_variant_requirements(const SEPExecutionVariant &variant) const noexcept
        {
            return !variant.is_data_driven() || frontier_ != nullptr;
        }

    private:
        graph_handle_type *graph_{nullptr};
        activity_handle_type *activity_{nullptr};
        partition_handle_type *partitions_{nullptr};
        task_handle_type *tasks_{nullptr};
        SEPFrontier *frontier_{nullptr};
    };
} // namespace hytgraph::sep
