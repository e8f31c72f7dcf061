// Shared state of SEP drivers (interface of the original
// include/sep/sep_execution_driver_base.hpp:12-252): the application, the
// configuration, the context and the initialised / converged flags.
#pragma once

#include "sep_execution_driver.hpp"
#include "sep_execution_requirements.hpp"

#include <stdexcept>

namespace hytgraph::sep
{
    template <typename TValue, typename TBuffer, typename TWeight = double>
    class SEPExecutionDriverBase : public SEPExecutionDriver<TValue, TBuffer, TWeight>
    {
    public:
        using base_type = SEPExecutionDriver<TValue, TBuffer, TWeight>;
        using application_type = typename base_type::application_type;
        using execution_config_type = typename base_type::execution_config_type;
        using execution_variant_type = typename base_type::execution_variant_type;
        using step_result_type = typename base_type::step_result_type;

        SEPExecutionDriverBase() = default;
        ~SEPExecutionDriverBase() override = default;

        void initialize(application_type &application, const execution_config_type &config) override
        {
            if (!is_valid_variant(config.variant()))
                throw std::invalid_argument("Invalid SEP execution variant");
            app_ = &application;
            config_ = config;
            initialized_ = true;
            converged_ = false;
            app_->set_execution_variant(config.variant());
        }

        step_result_type execute_step() override = 0;

        const execution_variant_type &execution_variant() const noexcept override { return config_.variant(); }
        bool initialized() const noexcept override { return initialized_; }
        bool converged() const noexcept override { return converged_; }
        void finalize() noexcept override
        {
            app_ = nullptr;
            ctx_ = nullptr;
            initialized_ = false;
            converged_ = false;
        }

    protected:
        void set_context(SEPExecutionContext &context) noexcept override { ctx_ = &context; }

        application_type &application()
        {
            if (app_ == nullptr)
                throw std::logic_error("SEP execution application is not initialized");
            return *app_;
        }
        const application_type &application() const
        {
            if (app_ == nullptr)
                throw std::logic_error("SEP execution application is not initialized");
            return *app_;
        }
        const execution_config_type &config() const noexcept { return config_; }
        SEPExecutionContext &context()
        {
            if (ctx_ == nullptr)
                throw std::logic_error("SEP execution context is not initialized");
            return *ctx_;
        }
        const SEPExecutionContext &context() const
        {
            if (ctx_ == nullptr)
                throw std::logic_error("SEP execution context is not initialized");
            return *ctx_;
        }
        bool context_is_valid() const noexcept
        {
            return initialized_ && app_ != nullptr && ctx_ != nullptr &&
                   is_valid_execution_configuration(config_, *ctx_);
        }
        void mark_converged() noexcept { converged_ = true; }
        void clear_converged() noexcept { converged_ = false; }

    private:
        application_type *app_{nullptr};
        execution_config_type config_{};
        SEPExecutionContext *ctx_{nullptr};
        bool initialized_{false};
        bool converged_{false};
    };
} // namespace hytgraph::sep
