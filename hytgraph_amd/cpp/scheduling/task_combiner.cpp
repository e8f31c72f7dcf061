// TaskCombiner with the grouping rules of the original
// src/scheduling/task_combiner.cpp:67-248.
#include "scheduling/task_combiner.hpp"

#include <algorithm>
#include <stdexcept>
#include <utility>

namespace hytgraph::scheduling
{
    namespace
    {
        using transfer::TransferEngine;

        ExecutableTask make_task(TransferEngine engine, std::vector<std::size_t> members)
        {
            ExecutableTask t;
            t.engine = engine;
            t.partition_indices = std::move(members);
            if (!t.partition_indices.empty())
            {
                t.first_partition_index = t.partition_indices.front();
                t.end_partition_index = t.partition_indices.back() + 1U;
            }
            return t;
        }
    } // namespace

    TaskCombiner::TaskCombiner(TaskCombinerOptions options) : options_(std::move(options))
    {
        if (options_.filter_combine_k == 0U)
            throw std::invalid_argument("TaskCombiner: filter_combine_k must be greater than zero");
    }

    const TaskCombinerOptions &TaskCombiner::options() const noexcept { return options_; }

    TaskCombinationPlan TaskCombiner::combine(const std::vector<TransferEngine> &engines) const
    {
        TaskCombinationPlan plan;
        plan.metrics.logical_partition_count = engines.size();
        std::vector<std::size_t> compaction, zero_copy;
        std::vector<ExecutableTask> tasks;
        for (std::size_t i = 0; i < engines.size();)
        {
            switch (engines[i])
            {
            case TransferEngine::ExpTMFilter:
            {
                std::vector<std::size_t> run;
                while (i < engines.size() && engines[i] == TransferEngine::ExpTMFilter &&
                       run.size() < options_.filter_combine_k)
                    run.push_back(i++);
                tasks.push_back(make_task(TransferEngine::ExpTMFilter, std::move(run)));
                break;
            }
            case TransferEngine::ExpTMCompaction:
                compaction.push_back(i++);
                break;
            case TransferEngine::ImpTMZeroCopy:
                zero_copy.push_back(i++);
                break;
            default:
                throw std::invalid_argument("TaskCombiner: unsupported transfer engine");
            }
        }
        if (!compaction.empty())
            tasks.push_back(make_task(TransferEngine::ExpTMCompaction, std::move(compaction)));
        if (!zero_copy.empty())
            tasks.push_back(make_task(TransferEngine::ImpTMZeroCopy, std::move(zero_copy)));
        std::sort(tasks.begin(), tasks.end(), [](const ExecutableTask &a, const ExecutableTask &b) {
            return a.first_partition_index < b.first_partition_index;
        });
        for (std::size_t t = 0; t < tasks.size(); ++t)
        {
            tasks[t].task_index = t;
            const std::size_t n = tasks[t].partition_count();
            if (tasks[t].engine == TransferEngine::ExpTMFilter)
                plan.metrics.filter_partitions_combined += n > 1 ? n - 1 : 0;
            else if (tasks[t].engine == TransferEngine::ExpTMCompaction)
                plan.metrics.compaction_partitions_combined += n;
            else
                plan.metrics.zero_copy_partitions_combined += n;
        }
        plan.tasks = std::move(tasks);
        plan.metrics.executable_task_count = plan.tasks.size();
        return plan;
    }
} // namespace hytgraph::scheduling
