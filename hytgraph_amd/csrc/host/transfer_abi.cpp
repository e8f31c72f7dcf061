// C-ABI planning functions of the hybrid transfer engine: partitioning, host
// metric collection, cost evaluation and task combination (see hytm_core.hpp
// for the equations).
#include "common.hpp"
#include "hytm_core.hpp"
#include "transfer_plan.hpp"

#include <algorithm>
#include <vector>

namespace hyt
{
    hytm::Params params_from(const hyt_cost_params &c)
    {
        hytm::Params p;
        p.alpha = c.alpha;
        p.beta = c.beta;
        p.gamma = c.gamma;
        p.request_bytes = c.request_bytes;
        p.requests_per_tlp = c.requests_per_tlp;
        p.rtt = c.rtt;
        p.cpu_bytes_per_second = c.cpu_bytes_per_second;
        p.entry_bytes = c.entry_bytes;
        p.index_bytes = c.index_bytes;
        p.alignment_bytes = c.alignment_bytes;
        return p;
    }

    hyt_cost_params default_cost_params()
    {
        const hytm::Params p;
        hyt_cost_params c{};
        c.alpha = p.alpha;
        c.beta = p.beta;
        c.gamma = p.gamma;
        c.request_bytes = p.request_bytes;
        c.requests_per_tlp = p.requests_per_tlp;
        c.rtt = p.rtt;
        c.cpu_bytes_per_second = p.cpu_bytes_per_second;
        c.entry_bytes = p.entry_bytes;
        c.index_bytes = p.index_bytes;
        c.alignment_bytes = p.alignment_bytes;
        return c;
    }

    std::vector<hyt_partition> partition_graph(const std::uint64_t *off, std::uint64_t V, std::uint64_t target)
    {
        require(target > 0, "LogicalPartitioner: partition_bytes must be greater than zero");
        std::vector<hyt_partition> parts;
        if (V == 0)
            return parts;
        std::uint64_t begin = 0, edge_begin = 0;
        for (std::uint64_t v = 0; v < V; ++v)
        {
            const std::uint64_t bytes = (off[v + 1] - edge_begin) * 4ULL;
            if (v > begin && bytes > target)
            {
                parts.push_back(hyt_partition{static_cast<std::uint32_t>(begin), static_cast<std::uint32_t>(v),
                                              edge_begin, off[v]});
                begin = v;
                edge_begin = off[v];
            }
        }
        parts.push_back(hyt_partition{static_cast<std::uint32_t>(begin), static_cast<std::uint32_t>(V), edge_begin,
                                      off[V]});
        return parts;
    }

    void combine_tasks(const std::int32_t *engines, std::uint64_t n, std::uint64_t k, std::vector<Task> &out)
    {
        require(k > 0, "TaskCombiner: filter_combine_k must be greater than zero");
        out.clear();
        Task comp{HYT_ENGINE_COMPACTION, {}}, zc{HYT_ENGINE_ZERO_COPY, {}};
        for (std::uint64_t i = 0; i < n;)
        {
            const std::int32_t e = engines[i];
            if (e == HYT_ENGINE_FILTER)
            {
                Task t{HYT_ENGINE_FILTER, {}};
                while (i < n && engines[i] == HYT_ENGINE_FILTER && t.members.size() < k)
                    t.members.push_back(i++);
                out.push_back(std::move(t));
            }
            else if (e == HYT_ENGINE_COMPACTION)
                comp.members.push_back(i++);
            else if (e == HYT_ENGINE_ZERO_COPY)
                zc.members.push_back(i++);
            else
                fail(HYT_ERR_INVALID_ARGUMENT, "TaskCombiner: unsupported transfer engine");
        }
        if (!comp.members.empty())
            out.push_back(std::move(comp));
        if (!zc.members.empty())
            out.push_back(std::move(zc));
        std::stable_sort(out.begin(), out.end(),
                         [](const Task &a, const Task &b) { return a.members.front() < b.members.front(); });
    }
} // namespace hyt

extern "C" {

int hyt_cost_params_default(hyt_cost_params *out)
{
    return hyt::guarded("hyt_cost_params_default", [&] {
        hyt::require(out != nullptr, "null output");
        *out = hyt::default_cost_params();
    });
}

int hyt_partition_graph(const uint64_t *row_offsets, uint64_t num_vertices, uint64_t target_bytes, hyt_partition *out,
                        uint64_t capacity, uint64_t *count)
{
    return hyt::guarded("hyt_partition_graph", [&] {
        hyt::require(row_offsets != nullptr && count != nullptr, "null pointer");
        const auto parts = hyt::partition_graph(row_offsets, num_vertices, target_bytes);
        *count = parts.size();
        hyt::require(out != nullptr || parts.empty(), "null output");
        hyt::require(parts.size() <= capacity, "partition output capacity too small");
        std::copy(parts.begin(), parts.end(), out);
    });
}

int hyt_partition_metrics_host(const uint64_t *row_offsets, uint64_t num_vertices, const uint8_t *active,
                               const hyt_partition *parts, uint64_t n, const hyt_cost_params *params,
                               hyt_partition_metrics *out)
{
    return hyt::guarded("hyt_partition_metrics_host", [&] {
        hyt::require(row_offsets && active && params && (n == 0 || (parts && out)), "null pointer");
        const hyt::hytm::Params p = hyt::params_from(*params);
        hyt::hytm::validate(p);
        for (uint64_t i = 0; i < n; ++i)
        {
            const hyt_partition &pt = parts[i];
            hyt::require(pt.vertex_begin <= pt.vertex_end && pt.vertex_end <= num_vertices,
                         "HyTMCostModel: partition vertex range is invalid");
            hyt::hytm::Metrics m;
            m.total_edges = pt.edge_end - pt.edge_begin;
            for (uint64_t v = pt.vertex_begin; v < pt.vertex_end; ++v)
            {
                if (!active[v])
                    continue;
                const uint64_t deg = row_offsets[v + 1] - row_offsets[v];
                ++m.active_vertices;
                m.active_edges += deg;
                m.zc_requests += hyt::hytm::vertex_requests(deg, p);
                m.zc_alignment += hyt::hytm::vertex_alignment(row_offsets[v], deg, p);
            }
            hyt::hytm::finish_metrics(m, p);
            out[i] = hyt_partition_metrics{m.active_vertices, m.active_edges, m.total_edges, m.filter_bytes,
                                           m.compaction_bytes, m.zc_requests, m.zc_alignment, m.zc_total_requests};
        }
    });
}

int hyt_cost_evaluate(const hyt_cost_params *params, const hyt_partition_metrics *metrics, uint64_t n,
                      hyt_partition_costs *out)
{
    return hyt::guarded("hyt_cost_evaluate", [&] {
        hyt::require(params && (n == 0 || (metrics && out)), "null pointer");
        const hyt::hytm::Params p = hyt::params_from(*params);
        hyt::hytm::validate(p);
        for (uint64_t i = 0; i < n; ++i)
        {
            hyt::hytm::Metrics m{metrics[i].active_vertices, metrics[i].active_edges, metrics[i].total_edges,
                                 metrics[i].filter_bytes,    metrics[i].compaction_bytes, metrics[i].zc_requests,
                                 metrics[i].zc_alignment,    metrics[i].zc_total_requests};
            const hyt::hytm::Costs c = hyt::hytm::evaluate(m, p);
            out[i] = hyt_partition_costs{c.filter, c.compaction, c.zero_copy, c.zc_rtt, c.engine, 0};
        }
    });
}

int hyt_select_engine(double filter_cost, double compaction_cost, double zero_copy_cost, double alpha, double beta,
                      int32_t *engine)
{
    return hyt::guarded("hyt_select_engine", [&] {
        hyt::require(engine != nullptr, "null output");
        *engine = hyt::hytm::select(filter_cost, compaction_cost, zero_copy_cost, alpha, beta);
    });
}

int hyt_combine_tasks(const int32_t *engines, uint64_t n, uint64_t k, int32_t *task_engine, uint64_t *task_offset,
                      uint64_t *members, uint64_t *num_tasks)
{
    return hyt::guarded("hyt_combine_tasks", [&] {
        hyt::require(num_tasks != nullptr && (n == 0 || (engines && task_engine && task_offset && members)),
                     "null pointer");
        std::vector<hyt::Task> tasks;
        hyt::combine_tasks(engines, n, k, tasks);
        *num_tasks = tasks.size();
        uint64_t at = 0;
        for (std::size_t t = 0; t < tasks.size(); ++t)
        {
            task_engine[t] = tasks[t].engine;
            task_offset[t] = at;
            for (uint64_t m : tasks[t].members)
                members[at++] = m;
        }
        if (task_offset)
            task_offset[tasks.size()] = at;
    });
}

} // extern "C"
