// ref_driver.cpp - extern "C" entry points over the ORIGINAL project's CPU
// functions, compiled from the original sources where they lie (see
// oracle/Makefile, target `ref`). TEST INFRASTRUCTURE ONLY: used to pin the
// oracle restatement and as the cpu_baseline of bench.py ("reference" kind).
//
// Wraps: sssp_cpu (src/algorithms/sssp.cpp:109), pagerank_cpu
// (src/algorithms/pagerank.cpp:89), via CSRGraph (src/graph/csr_graph.cpp:11);
// LogicalPartitioner::partition (src/graph/partition.cpp:105),
// HyTMCostModel::evaluate_partitions (src/transfer/hytm_cost_model.cpp:617),
// TaskCombiner::combine (src/scheduling/task_combiner.cpp:67);
// HubSorter::sort (src/scheduling/hub_sort.cpp:28), CSRGraph::reorder_vertices
// (src/graph/csr_graph.cpp:171), pagerank_contributions
// (src/algorithms/pagerank.cpp:165), sssp_contributions
// (src/algorithms/sssp.cpp:236).
#include "algorithms/pagerank.hpp"
#include "algorithms/sssp.hpp"
#include "graph/activity_tracker.hpp"
#include "graph/csr_graph.hpp"
#include "graph/partition.hpp"
#include "scheduling/hub_sort.hpp"
#include "scheduling/task_combiner.hpp"
#include "transfer/hytm_cost_model.hpp"

#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <vector>

using hytgraph::graph::CSRGraph;

namespace
{
    CSRGraph make_graph(std::uint64_t V, const std::uint64_t *off, const std::uint32_t *col, const float *w)
    {
        const std::uint64_t E = off[V];
        std::vector<CSRGraph::offset_type> o(off, off + V + 1);
        std::vector<CSRGraph::vertex_id> c(col, col + E);
        std::vector<CSRGraph::weight_type> ww;
        if (w)
            ww.assign(w, w + E);
        return CSRGraph(V, std::move(o), std::move(c), std::move(ww));
    }

    int code_of(const std::exception_ptr &ep)
    {
        try
        {
            std::rethrow_exception(ep);
        }
        catch (const std::invalid_argument &)
        {
            return 1;
        }
        catch (...)
        {
            return 7;
        }
    }
} // namespace

extern "C" {

// Returns 0, 1 (std::invalid_argument) or 7 (any other exception).
int ref_sssp_cpu(std::uint64_t V, const std::uint64_t *off, const std::uint32_t *col, const float *w,
                 std::uint32_t source, float tol, std::uint64_t max_iter, float *out, std::uint64_t *iterations,
                 int *converged)
{
    try
    {
        const CSRGraph g = make_graph(V, off, col, w);
        hytgraph::algorithms::SSSPOptions o;
        o.source = source;
        o.tolerance = tol;
        o.max_iterations = max_iter;
        const auto r = hytgraph::algorithms::sssp_cpu(g, o);
        std::memcpy(out, r.distances.data(), r.distances.size() * sizeof(float));
        *iterations = r.iterations;
        *converged = r.converged ? 1 : 0;
        return 0;
    }
    catch (...)
    {
        return code_of(std::current_exception());
    }
}

int ref_pagerank_cpu(std::uint64_t V, const std::uint64_t *off, const std::uint32_t *col, float damping, float tol,
                     std::uint64_t max_iter, float *out, std::uint64_t *iterations, float *residual, int *converged)
{
    try
    {
        const CSRGraph g = make_graph(V, off, col, nullptr);
        hytgraph::algorithms::PageRankOptions o;
        o.damping_factor = damping;
        o.tolerance = tol;
        o.max_iterations = max_iter;
        const auto r = hytgraph::algorithms::pagerank_cpu(g, o);
        std::memcpy(out, r.ranks.data(), r.ranks.size() * sizeof(float));
        *iterations = r.iterations;
        *residual = r.residual;
        *converged = r.converged ? 1 : 0;
        return 0;
    }
    catch (...)
    {
        return code_of(std::current_exception());
    }
}

// Partitions: out arrays of capacity cap: vb, ve (u32), eb, ee (u64).
int ref_partition(std::uint64_t V, const std::uint64_t *off, const std::uint32_t *col, std::uint64_t bytes,
                  std::uint32_t *vb, std::uint32_t *ve, std::uint64_t *eb, std::uint64_t *ee, std::uint64_t cap,
                  std::uint64_t *count)
{
    try
    {
        const CSRGraph g = make_graph(V, off, col, nullptr);
        const auto parts = hytgraph::graph::LogicalPartitioner(bytes).partition(g);
        *count = parts.size();
        if (parts.size() > cap)
            return 7;
        for (std::size_t i = 0; i < parts.size(); ++i)
        {
            vb[i] = parts[i].vertex_begin();
            ve[i] = parts[i].vertex_end();
            eb[i] = parts[i].edge_begin();
            ee[i] = parts[i].edge_end();
        }
        return 0;
    }
    catch (...)
    {
        return code_of(std::current_exception());
    }
}

// Cost evaluation of the partitions of `bytes` for an activity byte array.
// params: alpha beta gamma rtt cpu (doubles) ; m MR d1 d2 align (u64).
// out per partition: 8 u64 metrics + 4 doubles (f c z rtt_zc) + engine (i32).
int ref_cost_evaluate(std::uint64_t V, const std::uint64_t *off, const std::uint32_t *col, const std::uint8_t *active,
                      std::uint64_t bytes, const double *dp, const std::uint64_t *up, std::uint64_t *metrics,
                      double *costs, std::int32_t *engines, std::uint64_t cap, std::uint64_t *count)
{
    try
    {
        const CSRGraph g = make_graph(V, off, col, nullptr);
        const auto parts = hytgraph::graph::LogicalPartitioner(bytes).partition(g);
        hytgraph::graph::ActivityTracker act(V);
        for (std::uint64_t v = 0; v < V; ++v)
            if (active[v])
                act.set_active(static_cast<std::uint32_t>(v));
        hytgraph::transfer::HyTMCostModelOptions o;
        o.alpha = dp[0];
        o.beta = dp[1];
        o.gamma = dp[2];
        o.rtt = dp[3];
        o.cpu_compaction_throughput_bytes_per_second = dp[4];
        o.request_payload_bytes = up[0];
        o.max_requests_per_tlp = up[1];
        o.destination_entry_bytes = up[2];
        o.vertex_index_bytes = up[3];
        o.alignment_bytes = up[4];
        const auto res = hytgraph::transfer::HyTMCostModel(o).evaluate_partitions(g, parts, act);
        *count = res.size();
        if (res.size() > cap)
            return 7;
        for (std::size_t i = 0; i < res.size(); ++i)
        {
            const auto &m = res[i].metrics;
            std::uint64_t *mo = metrics + 8 * i;
            mo[0] = m.active_vertices;
            mo[1] = m.active_edges;
            mo[2] = m.total_edges;
            mo[3] = m.filter_transfer_bytes;
            mo[4] = m.compaction_bytes;
            mo[5] = m.zero_copy_memory_requests;
            mo[6] = m.zero_copy_alignment_overhead;
            mo[7] = m.zero_copy_total_requests;
            costs[4 * i + 0] = res[i].filter_cost;
            costs[4 * i + 1] = res[i].compaction_cost;
            costs[4 * i + 2] = res[i].zero_copy_cost;
            costs[4 * i + 3] = res[i].zero_copy_rtt;
            engines[i] = static_cast<std::int32_t>(res[i].selected_engine);
        }
        return 0;
    }
    catch (...)
    {
        return code_of(std::current_exception());
    }
}

// Task combination; engines 0 filter, 1 compaction, 2 zero-copy.
int ref_combine(const std::int32_t *engines, std::uint64_t n, std::uint64_t k, std::int32_t *task_engine,
                std::uint64_t *task_offset, std::uint64_t *members, std::uint64_t *num_tasks)
{
    try
    {
        std::vector<hytgraph::transfer::TransferEngine> e(n);
        for (std::uint64_t i = 0; i < n; ++i)
            e[i] = static_cast<hytgraph::transfer::TransferEngine>(engines[i]);
        hytgraph::scheduling::TaskCombinerOptions o;
        o.filter_combine_k = k;
        const auto plan = hytgraph::scheduling::TaskCombiner(o).combine(e);
        *num_tasks = plan.tasks.size();
        std::uint64_t at = 0;
        for (std::size_t t = 0; t < plan.tasks.size(); ++t)
        {
            task_engine[t] = static_cast<std::int32_t>(plan.tasks[t].engine);
            task_offset[t] = at;
            for (auto m : plan.tasks[t].partition_indices)
                members[at++] = m;
        }
        task_offset[plan.tasks.size()] = at;
        return 0;
    }
    catch (...)
    {
        return code_of(std::current_exception());
    }
}

// HubSorter::sort: order u32[V], scores f64[V], hub count.
int ref_hub_sort(std::uint64_t V, const std::uint64_t *off, const std::uint32_t *col, double fraction,
                 std::uint32_t *order, double *scores, std::uint64_t *hub_count)
{
    try
    {
        const CSRGraph g = make_graph(V, off, col, nullptr);
        hytgraph::scheduling::HubSortOptions o;
        o.hub_fraction = fraction;
        const auto r = hytgraph::scheduling::HubSorter(o).sort(g);
        for (std::size_t i = 0; i < r.vertex_order.size(); ++i)
            order[i] = r.vertex_order[i];
        std::memcpy(scores, r.scores.data(), r.scores.size() * sizeof(double));
        *hub_count = r.hub_count;
        return 0;
    }
    catch (...)
    {
        return code_of(std::current_exception());
    }
}

// CSRGraph::reorder_vertices on a copy: new offsets u64[V+1], columns, and
// float weights (w may be NULL).
int ref_reorder(std::uint64_t V, const std::uint64_t *off, const std::uint32_t *col, const float *w,
                const std::uint32_t *order, std::uint64_t *off2, std::uint32_t *col2, float *w2)
{
    try
    {
        CSRGraph g = make_graph(V, off, col, w);
        g.reorder_vertices(std::vector<CSRGraph::vertex_id>(order, order + V));
        std::memcpy(off2, g.row_offsets().data(), (V + 1) * sizeof(std::uint64_t));
        std::memcpy(col2, g.column_indices().data(), g.column_indices().size() * sizeof(std::uint32_t));
        if (w)
            std::memcpy(w2, g.edge_weights().data(), g.edge_weights().size() * sizeof(float));
        return 0;
    }
    catch (...)
    {
        return code_of(std::current_exception());
    }
}

int ref_pagerank_contributions(std::uint64_t V, const std::uint64_t *off, const std::uint32_t *col, float damping,
                               const float *x, double *out)
{
    try
    {
        const CSRGraph g = make_graph(V, off, col, nullptr);
        hytgraph::algorithms::PageRankOptions o;
        o.damping_factor = damping;
        const auto c = hytgraph::algorithms::pagerank_contributions(g, std::vector<float>(x, x + V), o);
        for (std::size_t i = 0; i < c.size(); ++i)
            out[c[i].vertex] = c[i].contribution;
        return 0;
    }
    catch (...)
    {
        return code_of(std::current_exception());
    }
}

int ref_sssp_contributions(std::uint64_t n, const float *cur, const float *cand, float tol, double *out)
{
    try
    {
        const auto c = hytgraph::algorithms::sssp_contributions(std::vector<float>(cur, cur + n),
                                                                std::vector<float>(cand, cand + n), tol);
        for (std::size_t i = 0; i < c.size(); ++i)
            out[c[i].vertex] = c[i].contribution;
        return 0;
    }
    catch (...)
    {
        return code_of(std::current_exception());
    }
}

} // extern "C"
