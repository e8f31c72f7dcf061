// A selected SEP configuration and its compatibility with a context
// (interface of the original include/sep/sep_execution_plan.hpp:12-124).
#pragma once

#include "sep_execution_config.hpp"
#include "sep_execution_context.hpp"
#include "sep_execution_requirements.hpp"

#include <string>

namespace hytgraph::sep
{
    class SEPExecutionPlan
    {
    public:
        explicit constexpr SEPExecutionPlan(SEPExecutionConfig config) noexcept : config_(config) {}

        constexpr const SEPExecutionConfig &config() const noexcept { return config_; }
        constexpr const SEPExecutionVariant &variant() const noexcept { return config_.variant(); }
        constexpr bool is_sync() const noexcept { return variant().is_synchronous(); }
        constexpr bool is_async() const noexcept { return variant().is_asynchronous(); }
        constexpr bool is_push() const noexcept { return variant().is_push(); }
        constexpr bool is_pull() const noexcept { return variant().is_pull(); }
        constexpr bool is_data_driven() const noexcept { return variant().is_data_driven(); }
        constexpr bool is_topology_driven() const noexcept { return variant().is_topology_driven(); }
        constexpr bool requires_frontier() const noexcept { return sep::requires_frontier(variant()); }
        bool is_compatible_with(const SEPExecutionContext &context) const noexcept
        {
            return is_valid_execution_configuration(config_, context);
        }
        std::string to_string() const { return variant().to_string(); }

    private:
        SEPExecutionConfig config_;
    };
} // namespace hytgraph::sep
