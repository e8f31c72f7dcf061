// SEP execution configuration (interface of the original
// include/sep/sep_execution_config.hpp:12-160): one variant, SYNC_PUSH_DD by
// default, with per-axis setters.
#pragma once

#include "sep_execution_variant.hpp"

#include <string>
#include <string_view>

namespace hytgraph::sep
{
    class SEPExecutionConfig
    {
    public:
        constexpr SEPExecutionConfig() noexcept
            : variant_(ExecutionMode::SYNC, MessagePassing::PUSH, SchedulingMode::DATA_DRIVEN)
        {
        }
        explicit constexpr SEPExecutionConfig(SEPExecutionVariant variant) noexcept : variant_(variant) {}

        constexpr const SEPExecutionVariant &variant() const noexcept { return variant_; }
        constexpr ExecutionMode execution_mode() const noexcept { return variant_.execution_mode(); }
        constexpr MessagePassing message_passing() const noexcept { return variant_.message_passing(); }
        constexpr SchedulingMode scheduling_mode() const noexcept { return variant_.scheduling_mode(); }
        constexpr bool is_synchronous() const noexcept { return variant_.is_synchronous(); }
        constexpr bool is_asynchronous() const noexcept { return variant_.is_asynchronous(); }
        constexpr bool is_push() const noexcept { return variant_.is_push(); }
        constexpr bool is_pull() const noexcept { return variant_.is_pull(); }
        constexpr bool is_data_driven() const noexcept { return variant_.is_data_driven(); }
        constexpr bool is_topology_driven() const noexcept { return variant_.is_topology_driven(); }
        std::string to_string() const { return variant_.to_string(); }

        constexpr void set_variant(SEPExecutionVariant v) noexcept { variant_ = v; }
        constexpr void set_execution_mode(ExecutionMode m) noexcept
        {
            variant_ = SEPExecutionVariant(m, variant_.message_passing(), variant_.scheduling_mode());
        }
        constexpr void set_message_passing(MessagePassing p) noexcept
        {
            variant_ = SEPExecutionVariant(variant_.execution_mode(), p, variant_.scheduling_mode());
        }
        constexpr void set_scheduling_mode(SchedulingMode s) noexcept
        {
            variant_ = SEPExecutionVariant(variant_.execution_mode(), variant_.message_passing(), s);
        }

        static SEPExecutionConfig from_string(std::string_view name)
        {
            return SEPExecutionConfig(SEPExecutionVariant::from_string(name));
        }

    private:
        SEPExecutionVariant variant_;
    };
} // namespace hytgraph::sep
