// Selects an execution plan by variant, configuration or name (interface of
// the original include/sep/sep_execution_selection.hpp:12-83).
#pragma once

#include "sep_execution_plan.hpp"
#include "sep_execution_variant_registry.hpp"

#include <string_view>

namespace hytgraph::sep
{
    class SEPExecutionSelection
    {
    public:
        static constexpr SEPExecutionPlan select(const SEPExecutionVariant &v) noexcept
        {
            return SEPExecutionPlan(SEPExecutionConfig(v));
        }
        static constexpr SEPExecutionPlan select(const SEPExecutionConfig &c) noexcept { return SEPExecutionPlan(c); }
        static SEPExecutionPlan select(std::string_view name)
        {
            return SEPExecutionPlan(SEPExecutionConfig::from_string(name));
        }
        static constexpr SEPExecutionPlan select_default() noexcept { return SEPExecutionPlan(SEPExecutionConfig{}); }
        static bool can_select(std::string_view name) noexcept { return SEPExecutionVariantRegistry::contains(name); }
    };
} // namespace hytgraph::sep
