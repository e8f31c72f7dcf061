// Compile-time traits of SEP applications (interface of the original
// include/sep/sep_application_traits.hpp:11-103).
#pragma once

#include <cstddef>
#include <type_traits>

namespace hytgraph::sep
{
    template <typename Application>
    struct SEPApplicationTraits
    {
        static constexpr bool requires_edge_weights = false;
        static constexpr bool supports_activity_predicate = true;
    };

    template <typename Application>
    using sep_application_value_t = typename Application::value_type;
    template <typename Application>
    using sep_application_buffer_t = typename Application::buffer_type;
    template <typename Application>
    using sep_application_weight_t = typename Application::weight_type;

    // True when Application names value_type, buffer_type and weight_type.
    template <typename Application, typename = void>
    struct is_sep_application_compatible : std::false_type
    {
    };
    template <typename Application>
    struct is_sep_application_compatible<Application,
                                         std::void_t<typename Application::value_type,
                                                     typename Application::buffer_type,
                                                     typename Application::weight_type>> : std::true_type
    {
    };
    template <typename Application>
    inline constexpr bool is_sep_application_compatible_v = is_sep_application_compatible<Application>::value;

    template <typename Application>
    struct SEPSSSPTraits
    {
        static constexpr bool requires_edge_weights = true;
        static constexpr bool supports_activity_predicate = true;
    };

    template <typename Application>
    struct SEPPageRankTraits
    {
        static constexpr bool requires_edge_weights = false;
        static constexpr bool supports_activity_predicate = true;
    };
} // namespace hytgraph::sep
