// SEP driver factory (interface of the original
// include/sep/sep_execution_factory.hpp:26-157). For an arbitrary
// application the factory returns a deferred driver, as the original does:
// it checks the configuration on initialize() and reports
// INVALID_CONFIGURATION from execute_step(). The SSSP and PageRank
// applications of sep_gpu_drivers.hpp get drivers that run the variant on the
// GPU.
#pragma once

#include "sep_execution_driver_base.hpp"
#include "sep_execution_requirements.hpp"

#include <memory>
#include <stdexcept>
#include <string_view>

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
            return std::make_unique<SEPDeferredExecutionDriver<TValue, TBuffer, TWeight>>(variant);
        }
        static driver_ptr create(const SEPExecutionConfig &config) { return create(config.variant()); }
        static driver_ptr create(std::string_view name) { return create(SEPExecutionVariant::from_string(name)); }
    };
} // namespace hytgraph::sep
