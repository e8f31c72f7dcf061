// Presents any object with the SEP hook set as a SEPApplication (interface
// of the original include/sep/sep_application_adapter.hpp:12-214); every
// hook forwards to the wrapped object.
#pragma once

#include "sep_application.hpp"

namespace hytgraph::sep
{
    template <typename Application>
    class SEPApplicationAdapter final
        : public SEPApplication<typename Application::value_type, typename Application::buffer_type,
                                typename Application::weight_type>
    {
    public:
        using application_type = Application;
        using value_type = typename Application::value_type;
        using buffer_type = typename Application::buffer_type;
        using weight_type = typename Application::weight_type;
        using base_type = SEPApplication<value_type, buffer_type, weight_type>;
        using node_id_type = typename base_type::node_id_type;
        using CombineResult = typename base_type::CombineResult;
        using AccumulateResult = typename base_type::AccumulateResult;

        explicit SEPApplicationAdapter(Application &application) noexcept : app_(&application) {}
        ~SEPApplicationAdapter() override = default;
        SEPApplicationAdapter(const SEPApplicationAdapter &) = delete;
        SEPApplicationAdapter &operator=(const SEPApplicationAdapter &) = delete;
        // Movable so make_sep_application_adapter can return one by value.
        SEPApplicationAdapter(SEPApplicationAdapter &&) noexcept = default;

        void set_execution_variant(const SEPExecutionVariant &v) override { app_->set_execution_variant(v); }
        const SEPExecutionVariant &execution_variant() const noexcept override { return app_->execution_variant(); }
        value_type init_value(node_id_type n) const override { return app_->init_value(n); }
        buffer_type init_buffer(node_id_type n) const override { return app_->init_buffer(n); }
        buffer_type identity_element() const override { return app_->identity_element(); }
        CombineResult combine_value_buffer(node_id_type n, const value_type &v, const buffer_type &b) const override
        {
            return app_->combine_value_buffer(n, v, b);
        }
        AccumulateResult accumulate_buffer(node_id_type s, node_id_type d, buffer_type &dst,
                                           const buffer_type &msg) const override
        {
            return app_->accumulate_buffer(s, d, dst, msg);
        }
        AccumulateResult accumulate_buffer(node_id_type s, node_id_type d, weight_type w, buffer_type &dst,
                                           const buffer_type &msg) const override
        {
            return app_->accumulate_buffer(s, d, w, dst, msg);
        }
        bool is_active(node_id_type n, const buffer_type &b) const override { return app_->is_active(n, b); }
        void post_computation() override { app_->post_computation(); }
        bool is_high_priority(const buffer_type &cur, const buffer_type &b) const override
        {
            return app_->is_high_priority(cur, b);
        }
        bool requires_edge_weights() const noexcept override { return app_->requires_edge_weights(); }
        bool supports_activity_predicate() const noexcept override { return app_->supports_activity_predicate(); }

        Application &underlying() noexcept { return *app_; }
        const Application &underlying() const noexcept { return *app_; }

    private:
        Application *app_;
    };

    template <typename Application>
    SEPApplicationAdapter<Application> make_sep_application_adapter(Application &application) noexcept
    {
        return SEPApplicationAdapter<Application>(application);
    }
} // namespace hytgraph::sep
