// SEP execution variants (interface of the original
// include/sep/sep_execution_variant.hpp:73-315): execution mode (SYNC /
// ASYNC) x message passing (PUSH / PULL) x scheduling (DATA_DRIVEN /
// TOPOLOGY_DRIVEN), named "SYNC_PUSH_DD" ... "ASYNC_PULL_TD", ordered
// lexicographically by (mode, passing, scheduling).
#pragma once

#include <array>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <string_view>

namespace hytgraph::sep
{
    enum class ExecutionMode
    {
        SYNC = 0,
        ASYNC = 1,
    };

    enum class MessagePassing
    {
        PUSH = 0,
        PULL = 1,
    };

    enum class SchedulingMode
    {
        DATA_DRIVEN = 0,
        TOPOLOGY_DRIVEN = 1,
    };

    class SEPExecutionVariant
    {
    public:
        constexpr SEPExecutionVariant(ExecutionMode m, MessagePassing p, SchedulingMode s) noexcept
            : mode_(m), passing_(p), scheduling_(s)
        {
        }

        constexpr ExecutionMode execution_mode() const noexcept { return mode_; }
        constexpr MessagePassing message_passing() const noexcept { return passing_; }
        constexpr SchedulingMode scheduling_mode() const noexcept { return scheduling_; }

        constexpr bool is_synchronous() const noexcept { return mode_ == ExecutionMode::SYNC; }
        constexpr bool is_asynchronous() const noexcept { return mode_ == ExecutionMode::ASYNC; }
        constexpr bool is_push() const noexcept { return passing_ == MessagePassing::PUSH; }
        constexpr bool is_pull() const noexcept { return passing_ == MessagePassing::PULL; }
        constexpr bool is_data_driven() const noexcept { return scheduling_ == SchedulingMode::DATA_DRIVEN; }
        constexpr bool is_topology_driven() const noexcept { return scheduling_ == SchedulingMode::TOPOLOGY_DRIVEN; }

        // Position in the (mode, passing, scheduling) order, 0..7.
        constexpr int ordinal() const noexcept
        {
            return static_cast<int>(mode_) * 4 + static_cast<int>(passing_) * 2 + static_cast<int>(scheduling_);
        }

        constexpr bool operator==(const SEPExecutionVariant &o) const noexcept { return ordinal() == o.ordinal(); }
        constexpr bool operator!=(const SEPExecutionVariant &o) const noexcept { return ordinal() != o.ordinal(); }
        constexpr bool operator<(const SEPExecutionVariant &o) const noexcept { return ordinal() < o.ordinal(); }

        std::string to_string() const
        {
            std::string s = is_synchronous() ? "SYNC_" : "ASYNC_";
            s += is_push() ? "PUSH_" : "PULL_";
            s += is_data_driven() ? "DD" : "TD";
            return s;
        }

        // Every variant in ordinal order.
        static constexpr std::array<SEPExecutionVariant, 8> all() noexcept
        {
            return {{of(0), of(1), of(2), of(3), of(4), of(5), of(6), of(7)}};
        }

        static SEPExecutionVariant from_string(std::string_view name)
        {
            for (const SEPExecutionVariant &v : all())
                if (v.to_string() == name)
                    return v;
            throw std::invalid_argument("Invalid SEP execution variant: " + std::string(name));
        }

    private:
        static constexpr SEPExecutionVariant of(int ordinal) noexcept
        {
            return SEPExecutionVariant(static_cast<ExecutionMode>(ordinal / 4),
                                       static_cast<MessagePassing>((ordinal / 2) % 2),
                                       static_cast<SchedulingMode>(ordinal % 2));
        }

        ExecutionMode mode_;
        MessagePassing passing_;
        SchedulingMode scheduling_;
    };
} // namespace hytgraph::sep
