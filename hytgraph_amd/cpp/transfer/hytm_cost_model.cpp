// HyTMCostModel over the shared formulas (hytm_core.hpp); validation and
// per-partition metric collection follow the original
// src/transfer/hytm_cost_model.cpp:94-452.
#include "transfer/hytm_cost_model.hpp"

#include "hytm_core.hpp"

#include <stdexcept>
#include <utility>

namespace hytgraph::transfer
{
    namespace
    {
        hyt::hytm::Params params_of(const HyTMCostModelOptions &o)
        {
            hyt::hytm::Params p;
            p.alpha = o.alpha;
            p.beta = o.beta;
            p.gamma = o.gamma;
            p.request_bytes = o.request_payload_bytes;
            p.requests_per_tlp = o.max_requests_per_tlp;
            p.rtt = o.rtt;
            p.cpu_bytes_per_second = o.cpu_compaction_throughput_bytes_per_second;
            p.entry_bytes = o.destination_entry_bytes;
            p.index_bytes = o.vertex_index_bytes;
            p.alignment_bytes = o.alignment_bytes;
            return p;
        }

        TransferEngine engine_of(int e)
        {
            return e == hyt::hytm::kCompaction ? TransferEngine::ExpTMCompaction
                   : e == hyt::hytm::kZeroCopy ? TransferEngine::ImpTMZeroCopy
                                               : TransferEngine::ExpTMFilter;
        }

        void validate_partition(const graph::CSRGraph &g, const graph::LogicalPartition &p)
        {
            if (p.vertex_begin() > p.vertex_end())
                throw std::invalid_argument("HyTMCostModel: partition vertex range is invalid");
            if (static_cast<graph::CSRGraph::offset_type>(p.vertex_end()) > g.num_vertices())
                throw std::invalid_argument("HyTMCostModel: partition vertex range exceeds graph");
            if (p.edge_begin() > p.edge_end())
                throw std::invalid_argument("HyTMCostModel: partition edge range is invalid");
            if (p.edge_end() > g.num_edges())
                throw std::invalid_argument("HyTMCostModel: partition edge range exceeds graph");
            const auto &off = g.row_offsets();
            if (off[p.vertex_begin()] != p.edge_begin())
                throw std::invalid_argument("HyTMCostModel: partition edge_begin does not match its CSR row boundary");
            if (off[p.vertex_end()] != p.edge_end())
                throw std::invalid_argument("HyTMCostModel: partition edge_end does not match its CSR row boundary");
        }
    } // namespace

    HyTMCostModel::HyTMCostModel(HyTMCostModelOptions options) : options_(std::move(options))
    {
        hyt::hytm::validate(params_of(options_));
    }

    const HyTMCostModelOptions &HyTMCostModel::options() const noexcept { return options_; }

    HyTMPartitionCosts HyTMCostModel::evaluate_partition(const graph::CSRGraph &graph,
                                                         const graph::LogicalPartition &partition,
                                                         const graph::ActivityTracker &activity,
                                                         std::size_t partition_index) const
    {
        const hyt::hytm::Params p = params_of(options_);
        hyt::hytm::validate(p);
        if (activity.vertex_count() != graph.num_vertices())
            throw std::invalid_argument("HyTMCostModel: activity tracker vertex count does not match graph");
        validate_partition(graph, partition);

        hyt::hytm::Metrics m;
        m.total_edges = partition.edge_count();
        const auto &off = graph.row_offsets();
        for (auto v = partition.vertex_begin(); v < partition.vertex_end(); ++v)
        {
            if (!activity.is_active(v))
                continue;
            const std::uint64_t deg = off[v + 1] - off[v];
            ++m.active_vertices;
            m.active_edges = hyt::hytm::checked_add(m.active_edges, deg, "HyTMCostModel: active edge count overflow");
            m.zc_requests = hyt::hytm::checked_add(m.zc_requests, hyt::hytm::vertex_requests(deg, p),
                                                   "HyTMCostModel: zero-copy request count overflow");
            m.zc_alignment = hyt::hytm::checked_add(m.zc_alignment, hyt::hytm::vertex_alignment(off[v], deg, p),
                                                    "HyTMCostModel: zero-copy alignment count overflow");
        }
        hyt::hytm::finish_metrics(m, p);
        const hyt::hytm::Costs c = hyt::hytm::evaluate(m, p);

        HyTMPartitionCosts r;
        r.metrics.partition_index = partition_index;
        r.metrics.vertex_begin = partition.vertex_begin();
        r.metrics.vertex_end = partition.vertex_end();
        r.metrics.active_vertices = m.active_vertices;
        r.metrics.active_edges = m.active_edges;
        r.metrics.total_edges = m.total_edges;
        r.metrics.filter_transfer_bytes = m.filter_bytes;
        r.metrics.compaction_bytes = m.compaction_bytes;
        r.metrics.zero_copy_memory_requests = m.zc_requests;
        r.metrics.zero_copy_alignment_overhead = m.zc_alignment;
        r.metrics.zero_copy_total_requests = m.zc_total_requests;
        r.filter_cost = c.filter;
        r.compaction_cost = c.compaction;
        r.zero_copy_cost = c.zero_copy;
        r.zero_copy_rtt = c.zc_rtt;
        r.selected_engine = engine_of(c.engine);
        return r;
    }

    std::vector<HyTMPartitionCosts> HyTMCostModel::evaluate_partitions(const graph::CSRGraph &graph,
                                                                       const std::vector<graph::LogicalPartition> &parts,
                                                                       const graph::ActivityTracker &activity) const
    {
        std::vector<HyTMPartitionCosts> out;
        out.reserve(parts.size());
        for (std::size_t i = 0; i < parts.size(); ++i)
            out.push_back(evaluate_partition(graph, parts[i], activity, i));
        return out;
    }

    TransferEngine HyTMCostModel::select_engine(double f, double c, double z, double alpha, double beta)
    {
        return engine_of(hyt::hytm::select(f, c, z, alpha, beta));
    }
} // namespace hytgraph::transfer
