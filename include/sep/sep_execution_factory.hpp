// SEP driver factory (interface of the original
// include/sep/sep_execution_factory.hpp:26-157). The original hands out
// deferred placeholders for every variant; here the factory's driver binds
// to the GPU at initialize() when the application is one the kernels run
// (SEPSSSPApplication with <float, float, float>, SEPPageRankApplication with
// <float, float, double>, sep_gpu_drivers.hpp) and then executes one round
// per execute_step(). For any other application it stays deferred, as in the
// original: it checks the configuration on initialize() and reports
// INVALID_CONFIGURATION from execute_step().
#pragma once

#include "sep_execution_driver_base.hpp"
#include "sep_execution_requirements.hpp"
#include "sep_gpu_drivers.hpp"

#include <memory>
#include <stdexcept>
#include <string_view>
#include <type_traits>

namespace hytgraph::sep
{
    template <typename TValue, typename TBuffer, typename TWeight = double>
    class SEPDeferredExecutionDriver final : public SEPExecutionDriverBase<TValue, TBuffer, TWeight>
    {
    public:
        using base_type = SEPExecutionDriverBase<TValue, TBuffer, TWeight>;
        using step_result_type = typename base_type::step_result_type;

        explicit SEPDeferredExecutionDriver(SEPExecutionVariant variant) : variant_(variant) {}

        void initialize(typename base_type::application_type &application,
                        const typename base_type::execution_config_type &config) override
        {
            if (config.variant() != variant_)
                throw std::invalid_argument("SEP execution factory variant does not match configuration");
            base_type::initialize(application, config);
        }
        void set_context(SEPExecutionContext &context) noexcept override { base_type::set_context(context); }
        step_result_type execute_step() override { return step_result_type::invalid_configuration(); }
        const typename base_type::execution_variant_type &execution_variant() const noexcept override
        {
            return variant_;
        }

    private:
        SEPExecutionVariant variant_;
    };

    // The factory's driver: one fixed variant; at initialize() it binds to the
    // GPU driver of a kernel-backed application, otherwise it behaves as the
    // deferred driver.
    template <typename TValue, typename TBuffer, typename TWeight = double>
    class SEPFactoryExecutionDriver final : public SEPExecutionDriverBase<TValue, TBuffer, TWeight>
    {
    public:
        using base_type = SEPExecutionDriverBase<TValue, TBuffer, TWeight>;
        using step_result_type = typename base_type::step_result_type;
        using driver_type = SEPExecutionDriver<TValue, TBuffer, TWeight>;

        explicit SEPFactoryExecutionDriver(SEPExecutionVariant variant) : variant_(variant) {}

        void initialize(typename base_type::application_type &application,
                        const typename base_type::execution_config_type &config) override
        {
            if (config.variant() != variant_)
                throw std::invalid_argument("SEP execution factory variant does not match configuration");
            base_type::initialize(application, config);
            bound_.reset();
            if constexpr (std::is_same_v<TValue, float> && std::is_same_v<TBuffer, float>)
            {
                if constexpr (std::is_same_v<TWeight, float>)
                {
                    if (dynamic_cast<SEPSSSPApplication *>(&application) != nullptr)
                        bound_ = std::make_unique<SEPSSSPGpuDriver>();
                }
                else if constexpr (std::is_same_v<TWeight, double>)
                {
                    if (dynamic_cast<SEPPageRankApplication *>(&application) != nullptr)
                        bound_ = std::make_unique<SEPPageRankGpuDriver>();
                }
            }
            if (bound_)
            {
                bound_->initialize(application, config);
                if (ctx_ != nullptr)
                    bound_->set_context(*ctx_);
            }
        }
        void set_context(SEPExecutionContext &context) noexcept override
        {
            base_type::set_context(context);
            ctx_ = &context;
            if (bound_)
                bound_->set_context(context);
        }
        step_result_type execute_step() override
        {
            if (!bound_)
                return step_result_type::invalid_configuration();
            const step_result_type r = bound_->execute_step();
            if (bound_->converged())
                this->mark_converged();
            return r;
        }
        void finalize() noexcept override
        {
            if (bound_)
                bound_->finalize();
            bound_.reset();
            ctx_ = nullptr;
            base_type::finalize();
        }
        const typename base_type::execution_variant_type &execution_variant() const noexcept override
        {
            return variant_;
        }
        // The GPU driver this driver runs (null while deferred).
        const driver_type *bound() const noexcept { return bound_.get(); }

    private:
        SEPExecutionVariant variant_;
        SEPExecutionContext *ctx_{nullptr};
        std::unique_ptr<driver_type> bound_;
    };

    // The state of the GPU driver a driver runs (values() of
    // sep_gpu_drivers.hpp), or null for a deferred or host driver.
    template <typename TValue, typename TBuffer, typename TWeight>
    const std::vector<float> *gpu_values(const SEPExecutionDriver<TValue, TBuffer, TWeight> &driver)
    {
        const SEPExecutionDriver<TValue, TBuffer, TWeight> *d = &driver;
        if (const auto *f = dynamic_cast<const SEPFactoryExecutionDriver<TValue, TBuffer, TWeight> *>(d))
            d = f->bound();
        if constexpr (std::is_same_v<TValue, float> && std::is_same_v<TBuffer, float>)
            if (const auto *g = dynamic_cast<const SEPGpuDriverBase<TWeight> *>(d))
                return &g->values();
        return nullptr;
    }

    template <typename TValue, typename TBuffer, typename TWeight = double>
    class SEPExecutionFactory
    {
    public:
        using driver_type = SEPExecutionDriver<TValue, TBuffer, TWeight>;
        using driver_ptr = std::unique_ptr<driver_type>;

        static driver_ptr create(const SEPExecutionVariant &variant)
        {
            if (!is_valid_variant(variant))
                throw std::invalid_argument("Invalid SEP execution variant");
            return std::make_unique<SEPFactoryExecutionDriver<TValue, TBuffer, TWeight>>(variant);
        }
        static driver_ptr create(const SEPExecutionConfig &config) { return create(config.variant()); }
        static driver_ptr create(std::string_view name) { return create(SEPExecutionVariant::from_string(name)); }
    };
} // namespace hytgraph::sep
