// Host-side planning helpers of the transfer engine (shared by the C ABI and
// the device engine).
#pragma once

#include "common.hpp"
#include "hytm_core.hpp"

#include <cstdint>
#include <vector>

namespace hyt
{
    struct Task
    {
        std::int32_t engine;
        std::vector<std::uint64_t> members;
    };

    hytm::Params params_from(const hyt_cost_params &c);
    hyt_cost_params default_cost_params();
    std::vector<hyt_partition> partition_graph(const std::uint64_t *off, std::uint64_t V, std::uint64_t target);
    void combine_tasks(const std::int32_t *engines, std::uint64_t n, std::uint64_t k, std::vector<Task> &out);
} // namespace hyt
