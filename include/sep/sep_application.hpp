// The SEP application contract (interface of the original
// include/sep/sep_application.hpp:45-244): per-vertex value and buffer,
// combine / accumulate / activity hooks, optional weighted accumulate,
// priority predicate and capability flags.
#pragma once

#include "sep_execution_variant.hpp"

#include <cstddef>

namespace hytgraph::sep
{
    template <typename TValue, typename TBuffer, typename TWeight = double>
    class SEPApplication
    {
    public:
        using value_type = TValue;
        using buffer_type = TBuffer;
        using weight_type = TWeight;
        using node_id_type = std::size_t;

        struct CombineResult
        {
            TBuffer new_buffer{};
            bool activate{false};
        };

        struct AccumulateResult
        {
            bool changed{false};
            bool activate{false};
        };

        virtual ~SEPApplication() = default;

        virtual void set_execution_variant(const SEPExecutionVariant &variant) = 0;
        virtual const SEPExecutionVariant &execution_variant() const noexcept = 0;
        virtual TValue init_value(node_id_type node) const = 0;
        virtual TBuffer init_buffer(node_id_type node) const = 0;
        virtual TBuffer identity_element() const = 0;
        virtual CombineResult combine_value_buffer(node_id_type node, const TValue &value,
                                                   const TBuffer &buffer) const = 0;
        virtual AccumulateResult accumulate_buffer(node_id_type src, node_id_type dst, TBuffer &destination,
                                                   const TBuffer &message) const = 0;
        // Weighted form; applications that read edge weights override it.
        virtual AccumulateResult accumulate_buffer(node_id_type, node_id_type, TWeight, TBuffer &,
                                                   const TBuffer &) const
        {
            return AccumulateResult{};
        }
        virtual bool is_active(node_id_type node, const TBuffer &buffer) const = 0;
        virtual void post_computation() {}
        virtual bool is_high_priority(const TBuffer &, const TBuffer &) const { return true; }
        virtual bool requires_edge_weights() const noexcept { return false; }
        virtual bool supports_activity_predicate() const noexcept { return true; }
    };
} // namespace hytgraph::sep
