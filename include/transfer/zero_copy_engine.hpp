// ImpTM zero-copy request model (interface of the original
// include/transfer/zero_copy_engine.hpp:14-158): per active vertex,
// requests = ceil(deg * 4 / m) and one alignment request when its logical
// neighbour byte offset (row_offsets[v] * 4) is not a multiple of the
// alignment; totals per partition and overall, modeled TLPs =
// ceil(requests / MR) over the whole graph (alignment excluded). prepare()
// computes the model; the device execution of zero-copy reads is the hybrid
// engine's zero-copy path over hipHostMalloc-mapped host arrays.
#pragma once

#include "graph/activity_tracker.hpp"
#include "graph/csr_graph.hpp"
#include "graph/partition.hpp"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace hytgraph::transfer
{
    enum class ZeroCopyMode
    {
        Modeled,
        MappedHost
    };

    struct ZeroCopyVertexMetrics
    {
        graph::CSRGraph::vertex_id vertex = 0;
        graph::CSRGraph::offset_type degree = 0;
        std::uint64_t memory_requests = 0;
        std::uint64_t alignment_overhead = 0;
    };

    struct ZeroCopyPartition
    {
        std::size_t partition_index = 0;
        graph::CSRGraph::vertex_id vertex_begin = 0;
        graph::CSRGraph::vertex_id vertex_end = 0;
        std::vector<graph::CSRGraph::vertex_id> active_vertices;
        std::vector<ZeroCopyVertexMetrics> vertex_metrics;
        graph::CSRGraph::offset_type active_edge_count = 0;
        std::uint64_t memory_request_count = 0;
        std::uint64_t alignment_overhead_count = 0;
        std::uint64_t modeled_tlp_count = 0;
    };

    struct ZeroCopyResult
    {
        ZeroCopyMode mode = ZeroCopyMode::Modeled;
        std::vector<ZeroCopyPartition> partitions;
        std::size_t active_partition_count = 0;
        graph::CSRGraph::offset_type active_vertex_count = 0;
        graph::CSRGraph::offset_type active_edge_count = 0;
        std::uint64_t memory_request_count = 0;
        std::uint64_t alignment_overhead_count = 0;
        std::uint64_t modeled_tlp_count = 0;
        bool host_memory_mapped = false;
        double preparation_seconds = 0.0;
    };

    struct ZeroCopyOptions
    {
        std::size_t request_payload_bytes = 128U; // m
        std::size_t max_requests_per_tlp = 256U;  // MR
        bool enable_host_mapping = true;
        std::size_t alignment_bytes = 128U;
    };

    class ImpTMZeroCopy
    {
    public:
        explicit ImpTMZeroCopy(ZeroCopyOptions options = {}) noexcept;

        [[nodiscard]] const ZeroCopyOptions &options() const noexcept;

        [[nodiscard]] ZeroCopyResult prepare(const graph::CSRGraph &graph,
                                             const std::vector<graph::LogicalPartition> &partitions,
                                             const graph::ActivityTracker &activity) const;

    private:
        ZeroCopyOptions options_;
    };
} // namespace hytgraph::transfer
