// The SEP driver interface (interface of the original
// include/sep/sep_execution_driver.hpp:15-165).
#pragma once

#include "sep_application.hpp"
#include "sep_execution_config.hpp"
#include "sep_execution_context.hpp"
#include "sep_execution_result.hpp"

#include <cstddef>

namespace hytgraph::sep
{
    struct SEPExecutionStepResult
    {
        std::size_t vertices_processed{0};
        std::size_t updates_generated{0};
        bool changed{false};
        bool converged{false};
    };

    template <typename TValue, typename TBuffer, typename TWeight = double>
    class SEPExecutionDriver
    {
    public:
        using application_type = SEPApplication<TValue, TBuffer, TWeight>;
        using execution_config_type = SEPExecutionConfig;
        using execution_variant_type = SEPExecutionVariant;
        using step_result_type = SEPExecutionResult;

        virtual ~SEPExecutionDriver() = default;
        virtual void initialize(application_type &application, const execution_config_type &config) = 0;
        virtual void set_context(SEPExecutionContext &context) = 0;
        virtual step_result_type execute_step() = 0;
        virtual const execution_variant_type &execution_variant() const noexcept = 0;
        virtual bool initialized() const noexcept = 0;
        virtual bool converged() const noexcept = 0;
        virtual void finalize() noexcept = 0;
    };
} // namespace hytgraph::sep
