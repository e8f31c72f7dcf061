// Registry of the eight SEP variants (interface of the original
// include/sep/sep_execution_variant_registry.hpp:12-150).
#pragma once

#include "sep_execution_variant.hpp"

#include <array>
#include <cstddef>
#include <string_view>

namespace hytgraph::sep
{
    class SEPExecutionVariantRegistry
    {
    public:
        static constexpr std::size_t kVariantCount = 8;
        using variant_array = std::array<SEPExecutionVariant, kVariantCount>;

        static constexpr const variant_array &all() noexcept { return variants_; }
        static constexpr std::size_t size() noexcept { return kVariantCount; }
        static constexpr const SEPExecutionVariant &at(std::size_t index) noexcept { return variants_[index]; }

        static const SEPExecutionVariant *find(std::string_view name) noexcept
        {
            for (const SEPExecutionVariant &v : variants_)
                if (v.to_string() == name)
                    return &v;
            return nullptr;
        }
        static bool contains(std::string_view name) noexcept { return find(name) != nullptr; }

    private:
        inline static constexpr variant_array variants_ = SEPExecutionVariant::all();
    };
} // namespace hytgraph::sep
