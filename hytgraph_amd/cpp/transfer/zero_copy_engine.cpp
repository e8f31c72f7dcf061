// ImpTMZeroCopy::prepare with the request model and checks of the original
// src/transfer/zero_copy_engine.cpp:119-393: per active vertex
// ceil(deg * 4 / m) requests plus one alignment request when
// row_offsets[v] * 4 is not a multiple of the alignment (degree > 0 only);
// per-partition TLP counts stay 0 and the total is ceil(requests / MR).
#include "transfer/zero_copy_engine.hpp"

#include "partition_checks.hpp"

#include <chrono>
#include <limits>
#include <stdexcept>
#include <utility>

namespace hytgraph::transfer
{
    namespace
    {
        std::uint64_t ceil_div(std::uint64_t a, std::uint64_t b)
        {
            if (b == 0U)
                throw std::invalid_argument("zero-copy division denominator must be greater than zero");
            return a / b + (a % b != 0U ? 1U : 0U);
        }
    } // namespace

    ImpTMZeroCopy::ImpTMZeroCopy(ZeroCopyOptions options) noexcept : options_(std::move(options)) {}

    const ZeroCopyOptions &ImpTMZeroCopy::options() const noexcept { return options_; }

    ZeroCopyResult ImpTMZeroCopy::prepare(const graph::CSRGraph &graph,
                                          const std::vector<graph::LogicalPartition> &partitions,
                                          const graph::ActivityTracker &activity) const
    {
        const auto t0 = std::chrono::steady_clock::now();
        if (options_.request_payload_bytes == 0U)
            throw std::invalid_argument("zero-copy request payload size must be greater than zero");
        if (options_.max_requests_per_tlp == 0U)
            throw std::invalid_argument("zero-copy maximum requests per TLP must be greater than zero");
        if (options_.alignment_bytes == 0U)
            throw std::invalid_argument("zero-copy alignment size must be greater than zero");
        if (activity.vertex_count() != graph.num_vertices())
            throw std::invalid_argument("activity tracker vertex count does not match graph");
        detail::check_coverage(graph, partitions, true);

        const auto stats = activity.partition_statistics(graph, detail::ranges_of(partitions));
        const auto &off = graph.row_offsets();
        const std::uint64_t m = options_.request_payload_bytes, align = options_.alignment_bytes;
        constexpr std::uint64_t kIdBytes = sizeof(graph::CSRGraph::vertex_id);
        ZeroCopyResult r;
        r.mode = ZeroCopyMode::Modeled; // the model only; mapped execution is the device engine's
        r.host_memory_mapped = false;
        r.partitions.resize(partitions.size());
        for (std::size_t i = 0; i < partitions.size(); ++i)
        {
            const graph::LogicalPartition &p = partitions[i];
            ZeroCopyPartition &zp = r.partitions[i];
            zp.partition_index = i;
            zp.vertex_begin = p.vertex_begin();
            zp.vertex_end = p.vertex_end();
            if (stats[i].active_vertices == 0U)
                continue;
            zp.active_vertices.reserve(static_cast<std::size_t>(stats[i].active_vertices));
            zp.vertex_metrics.reserve(static_cast<std::size_t>(stats[i].active_vertices));
            for (graph::CSRGraph::vertex_id v = p.vertex_begin(); v < p.vertex_end(); ++v)
            {
                if (!activity.is_active(v))
                    continue;
                ZeroCopyVertexMetrics vm;
                vm.vertex = v;
                vm.degree = off[v + 1] - off[v];
                if (vm.degree > std::numeric_limits<std::uint64_t>::max() / kIdBytes ||
                    off[v] > std::numeric_limits<std::uint64_t>::max() / kIdBytes)
                    throw std::overflow_error("zero-copy logical neighbor byte offset overflow");
                vm.memory_requests = ceil_div(vm.degree * kIdBytes, m);
                vm.alignment_overhead = (vm.degree != 0U && (off[v] * kIdBytes) % align != 0U) ? 1U : 0U;
                zp.active_vertices.push_back(v);
                zp.vertex_metrics.push_back(vm);
                zp.active_edge_count += vm.degree;
                zp.memory_request_count += vm.memory_requests;
                zp.alignment_overhead_count += vm.alignment_overhead;
            }
            zp.modeled_tlp_count = 0;
            if (!zp.active_vertices.empty())
                ++r.active_partition_count;
            r.active_vertex_count += zp.active_vertices.size();
            r.active_edge_count += zp.active_edge_count;
            r.memory_request_count += zp.memory_request_count;
            r.alignment_overhead_count += zp.alignment_overhead_count;
        }
        r.modeled_tlp_count = ceil_div(r.memory_request_count, options_.max_requests_per_tlp);
        r.preparation_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return r;
    }
} // namespace hytgraph::transfer
