// HyTM cost model (interface of the original
// include/transfer/hytm_cost_model.hpp:14-211): per-partition filter /
// compaction / zero-copy costs and the strict engine selector. The formulas
// live in hytgraph_amd/csrc/host/hytm_core.hpp and are shared with the device
// transfer engine, which evaluates them on counts gathered on the GPU.
#pragma once

#include "graph/activity_tracker.hpp"
#include "graph/csr_graph.hpp"
#include "graph/partition.hpp"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace hytgraph::transfer
{
    enum class TransferEngine
    {
        ExpTMFilter,
        ExpTMCompaction,
        ImpTMZeroCopy
    };

    struct HyTMCostModelOptions
    {
        double alpha = 0.80;
        double beta = 0.40;
        double gamma = 0.625;
        std::size_t request_payload_bytes = 128U;
        std::size_t max_requests_per_tlp = 256U;
        double rtt = 1.0;
        double cpu_compaction_throughput_bytes_per_second = 1.0;
        std::size_t destination_entry_bytes = sizeof(graph::CSRGraph::vertex_id);
        std::size_t vertex_index_bytes = sizeof(graph::CSRGraph::offset_type);
        std::size_t alignment_bytes = 128U;
    };

    struct HyTMPartitionMetrics
    {
        std::size_t partition_index = 0U;
        graph::CSRGraph::vertex_id vertex_begin = 0U;
        graph::CSRGraph::vertex_id vertex_end = 0U;
        graph::CSRGraph::offset_type active_vertices = 0U;
        graph::CSRGraph::offset_type active_edges = 0U;
        graph::CSRGraph::offset_type total_edges = 0U;
        std::uint64_t filter_transfer_bytes = 0U;
        std::uint64_t compaction_bytes = 0U;
        std::uint64_t zero_copy_memory_requests = 0U;
        std::uint64_t zero_copy_alignment_overhead = 0U;
        std::uint64_t zero_copy_total_requests = 0U;
    };

    struct HyTMPartitionCosts
    {
        HyTMPartitionMetrics metrics;
        double filter_cost = 0.0;
        double compaction_cost = 0.0;
        double zero_copy_cost = 0.0;
        double zero_copy_rtt = 0.0;
        TransferEngine selected_engine = TransferEngine::ExpTMFilter;
    };

    class HyTMCostModel
    {
    public:
        explicit HyTMCostModel(HyTMCostModelOptions options = {});

        [[nodiscard]] const HyTMCostModelOptions &options() const noexcept;

        [[nodiscard]] HyTMPartitionCosts evaluate_partition(const graph::CSRGraph &graph,
                                                            const graph::LogicalPartition &partition,
                                                            const graph::ActivityTracker &activity,
                                                            std::size_t partition_index = 0U) const;

        [[nodiscard]] std::vector<HyTMPartitionCosts>
        evaluate_partitions(const graph::CSRGraph &graph, const std::vector<graph::LogicalPartition> &partitions,
                            const graph::ActivityTracker &activity) const;

        [[nodiscard]] static TransferEngine select_engine(double filter_cost, double compaction_cost,
                                                          double zero_copy_cost, double alpha, double beta);

    private:
        HyTMCostModelOptions options_;
    };
} // namespace hytgraph::transfer
