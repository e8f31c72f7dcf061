// Outcome of one SEP execution step (interface of the original
// include/sep/sep_execution_result.hpp:10-142).
#pragma once

#include <cstddef>

namespace hytgraph::sep
{
    enum class SEPExecutionStatus
    {
        SUCCESS = 0,
        NOT_INITIALIZED,
        INVALID_CONFIGURATION,
        INVALID_CONTEXT,
        NO_WORK,
        CONVERGED,
    };

    struct SEPExecutionResult
    {
        SEPExecutionStatus status{SEPExecutionStatus::SUCCESS};
        std::size_t vertices_processed{0};
        std::size_t updates_generated{0};
        bool changed{false};
        bool converged{false};

        constexpr bool succeeded() const noexcept
        {
            return status == SEPExecutionStatus::SUCCESS || status == SEPExecutionStatus::CONVERGED;
        }
        constexpr bool has_work() const noexcept { return vertices_processed != 0 || updates_generated != 0; }
        constexpr bool reached_convergence() const noexcept
        {
            return converged || status == SEPExecutionStatus::CONVERGED;
        }

        static constexpr SEPExecutionResult of(SEPExecutionStatus s, bool conv = false) noexcept
        {
            return {s, 0, 0, false, conv};
        }
        static constexpr SEPExecutionResult success(std::size_t vertices = 0, std::size_t updates = 0,
                                                    bool state_changed = false) noexcept
        {
            return {SEPExecutionStatus::SUCCESS, vertices, updates, state_changed, false};
        }
        static constexpr SEPExecutionResult no_work() noexcept { return of(SEPExecutionStatus::NO_WORK); }
        static constexpr SEPExecutionResult converged_result() noexcept
        {
            return of(SEPExecutionStatus::CONVERGED, true);
        }
        static constexpr SEPExecutionResult not_initialized() noexcept
        {
            return of(SEPExecutionStatus::NOT_INITIALIZED);
        }
        static constexpr SEPExecutionResult invalid_configuration() noexcept
        {
            return of(SEPExecutionStatus::INVALID_CONFIGURATION);
        }
        static constexpr SEPExecutionResult invalid_context() noexcept
        {
            return of(SEPExecutionStatus::INVALID_CONTEXT);
        }
    };
} // namespace hytgraph::sep
