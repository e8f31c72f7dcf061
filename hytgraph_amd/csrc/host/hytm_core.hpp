// HyTM transfer cost model: the three cost equations and the strict engine
// selector of the original src/transfer/hytm_cost_model.cpp:233-715 (from the
// HyTGraph paper), as plain functions over per-partition counts. Shared by the
// C++ API (HyTMCostModel) and the device transfer engine, which fills the
// counts on the GPU each round.
//
//   T_filter     = ceil(E_p * d1 / (m * MR)) * RTT
//   T_compaction = ceil(B_c / (m * MR)) * RTT + B_c / T_cpu,
//                  B_c = active_edges * d1 + active_vertices * d2
//   T_zero_copy  = ceil((sum_v ceil(deg(v) * d1 / m) + sum_v am(v)) / MR) * RTT_zc,
//                  RTT_zc = gamma * RTT + (1 - gamma) * (active_edges / E_p) * RTT
//                  (gamma * RTT when E_p = 0); am(v) = 1 when v has edges and
//                  its first neighbour byte offset is not a multiple of align
//   select: compaction if T_c < alpha * T_f and T_c < beta * T_z,
//           else filter if T_f < T_z, else zero-copy (strict comparisons).
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>

namespace hyt::hytm
{
    enum Engine : int
    {
        kFilter = 0,
        kCompaction = 1,
        kZeroCopy = 2
    };

    struct Params
    {
        double alpha = 0.80;
        double beta = 0.40;
        double gamma = 0.625;
        std::uint64_t request_bytes = 128;  // m
        std::uint64_t requests_per_tlp = 256; // MR
        double rtt = 1.0;
        double cpu_bytes_per_second = 1.0;
        std::uint64_t entry_bytes = 4;       // d1
        std::uint64_t index_bytes = 8;       // d2
        std::uint64_t alignment_bytes = 128;
    };

    struct Metrics
    {
        std::uint64_t active_vertices = 0;
        std::uint64_t active_edges = 0;
        std::uint64_t total_edges = 0;
        std::uint64_t filter_bytes = 0;
        std::uint64_t compaction_bytes = 0;
        std::uint64_t zc_requests = 0;
        std::uint64_t zc_alignment = 0;
        std::uint64_t zc_total_requests = 0;
    };

    struct Costs
    {
        double filter = 0.0;
        double compaction = 0.0;
        double zero_copy = 0.0;
        double zc_rtt = 0.0;
        int engine = kFilter;
    };

    inline std::uint64_t checked_mul(std::uint64_t a, std::uint64_t b, const char *what)
    {
        if (a != 0 && b > std::numeric_limits<std::uint64_t>::max() / a)
            throw std::overflow_error(what);
        return a * b;
    }

    inline std::uint64_t checked_add(std::uint64_t a, std::uint64_t b, const char *what)
    {
        if (b > std::numeric_limits<std::uint64_t>::max() - a)
            throw std::overflow_error(what);
        return a + b;
    }

    inline std::uint64_t ceil_div(std::uint64_t a, std::uint64_t b)
    {
        if (b == 0)
            throw std::invalid_argument("HyTMCostModel: division by zero");
        return a / b + (a % b != 0 ? 1 : 0);
    }

    inline void validate(const Params &p)
    {
        if (!std::isfinite(p.alpha) || !std::isfinite(p.beta) || !std::isfinite(p.gamma))
            throw std::invalid_argument("HyTMCostModel: alpha, beta, and gamma must be finite");
        if (p.alpha <= 0.0)
            throw std::invalid_argument("HyTMCostModel: alpha must be greater than zero");
        if (p.beta <= 0.0)
            throw std::invalid_argument("HyTMCostModel: beta must be greater than zero");
        if (p.gamma < 0.0 || p.gamma > 1.0)
            throw std::invalid_argument("HyTMCostModel: gamma must be in the range [0, 1]");
        if (!std::isfinite(p.rtt) || p.rtt < 0.0)
            throw std::invalid_argument("HyTMCostModel: RTT must be finite and non-negative");
        if (!std::isfinite(p.cpu_bytes_per_second) || p.cpu_bytes_per_second <= 0.0)
            throw std::invalid_argument("HyTMCostModel: CPU compaction throughput must be finite and greater than zero");
        if (p.request_bytes == 0)
            throw std::invalid_argument("HyTMCostModel: request payload size must be greater than zero");
        if (p.requests_per_tlp == 0)
            throw std::invalid_argument("HyTMCostModel: maximum requests per TLP must be greater than zero");
        if (p.entry_bytes == 0)
            throw std::invalid_argument("HyTMCostModel: destination entry size must be greater than zero");
        if (p.index_bytes == 0)
            throw std::invalid_argument("HyTMCostModel: vertex index size must be greater than zero");
        if (p.alignment_bytes == 0)
            throw std::invalid_argument("HyTMCostModel: alignment size must be greater than zero");
    }

    // Requests to read one active vertex's neighbour list: ceil(deg*d1/m).
    inline std::uint64_t vertex_requests(std::uint64_t degree, const Params &p)
    {
        if (degree == 0)
            return 0;
        return ceil_div(checked_mul(degree, p.entry_bytes, "HyTMCostModel: zero-copy vertex byte count overflow"),
                        p.request_bytes);
    }

    // am(v): one extra request when the list does not start on an alignment
    // boundary (logical byte offset first_edge * d1).
    inline std::uint64_t vertex_alignment(std::uint64_t first_edge, std::uint64_t degree, const Params &p)
    {
        if (degree == 0)
            return 0;
        const std::uint64_t off =
            checked_mul(first_edge, p.entry_bytes, "HyTMCostModel: logical neighbor byte offset overflow");
        return (off % p.alignment_bytes != 0) ? 1 : 0;
    }

    // Completes the derived byte/request totals from the raw counts.
    inline void finish_metrics(Metrics &m, const Params &p)
    {
        m.filter_bytes = checked_mul(m.total_edges, p.entry_bytes, "HyTMCostModel: filter transfer byte count overflow");
        const std::uint64_t eb =
            checked_mul(m.active_edges, p.entry_bytes, "HyTMCostModel: compaction edge byte count overflow");
        const std::uint64_t ib =
            checked_mul(m.active_vertices, p.index_bytes, "HyTMCostModel: compaction index byte count overflow");
        m.compaction_bytes = checked_add(eb, ib, "HyTMCostModel: compaction transfer byte count overflow");
        m.zc_total_requests =
            checked_add(m.zc_requests, m.zc_alignment, "HyTMCostModel: zero-copy total request count overflow");
    }

    inline int select(double f, double c, double z, double alpha, double beta)
    {
        if (!std::isfinite(f) || !std::isfinite(c) || !std::isfinite(z))
            throw std::invalid_argument("HyTMCostModel::select_engine: costs must be finite");
        if (f < 0.0 || c < 0.0 || z < 0.0)
            throw std::invalid_argument("HyTMCostModel::select_engine: costs must be non-negative");
        if (!std::isfinite(alpha) || alpha <= 0.0)
            throw std::invalid_argument("HyTMCostModel::select_engine: alpha must be finite and greater than zero");
        if (!std::isfinite(beta) || beta <= 0.0)
            throw std::invalid_argument("HyTMCostModel::select_engine: beta must be finite and greater than zero");
        if (c < alpha * f && c < beta * z)
            return kCompaction;
        if (f < z)
            return kFilter;
        return kZeroCopy;
    }

    inline Costs evaluate(const Metrics &m, const Params &p)
    {
        Costs c;
        const std::uint64_t tlp_bytes = checked_mul(p.request_bytes, p.requests_per_tlp, "HyTMCostModel: TLP byte capacity overflow");
        c.filter = static_cast<double>(ceil_div(m.filter_bytes, tlp_bytes)) * p.rtt;
        c.compaction = static_cast<double>(ceil_div(m.compaction_bytes, tlp_bytes)) * p.rtt +
                       static_cast<double>(m.compaction_bytes) / p.cpu_bytes_per_second;
        if (m.total_edges == 0)
            c.zc_rtt = p.gamma * p.rtt;
        else
            c.zc_rtt = p.gamma * p.rtt +
                       (1.0 - p.gamma) * (static_cast<double>(m.active_edges) / static_cast<double>(m.total_edges)) * p.rtt;
        c.zero_copy = static_cast<double>(ceil_div(m.zc_total_requests, p.requests_per_tlp)) * c.zc_rtt;
        c.engine = select(c.filter, c.compaction, c.zero_copy, p.alpha, p.beta);
        return c;
    }
} // namespace hyt::hytm
