// Validity rules of SEP configurations (interface of the original
// include/sep/sep_execution_requirements.hpp:12-116): a configuration is
// runnable when the context has a graph and, for data-driven variants, a
// frontier.
#pragma once

#include "sep_execution_config.hpp"
#include "sep_execution_context.hpp"

namespace hytgraph::sep
{
    inline bool is_valid_execution_configuration(const SEPExecutionConfig &config,
                                                 const SEPExecutionContext &context) noexcept
    {
        return context.has_graph() && context.satisfies_variant_requirements(config.variant());
    }

    constexpr bool is_valid_variant(const SEPExecutionVariant &v) noexcept
    {
        const int m = static_cast<int>(v.execution_mode());
        const int p = static_cast<int>(v.message_passing());
        const int s = static_cast<int>(v.scheduling_mode());
        return m >= 0 && m <= 1 && p >= 0 && p <= 1 && s >= 0 && s <= 1;
    }

    constexpr bool requires_frontier(const SEPExecutionVariant &v) noexcept { return v.is_data_driven(); }

    constexpr bool is_topology_driven(const SEPExecutionVariant &v) noexcept { return v.is_topology_driven(); }
} // namespace hytgraph::sep
