// GPU execution of the SEP SSSP and PageRank applications through the C ABI.
#include "sep/sep_gpu_drivers.hpp"

#include "hytgraph_amd.h"

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace hytgraph::sep
{
    namespace
    {
        void check(int status, const char *what)
        {
            if (status == HYT_OK)
                return;
            const std::string msg = std::string(what) + ": " + hyt_last_error();
            if (status == HYT_ERR_INVALID_ARGUMENT)
                throw std::invalid_argument(msg);
            throw std::runtime_error(msg);
        }

        struct DeviceGraph
        {
            hyt_graph_t g = nullptr;
            DeviceGraph(const graph::CSRGraph &graph, bool weights, const char *what)
            {
                std::vector<std::uint32_t> bits;
                const std::uint32_t *w = nullptr;
                if (weights && graph.has_weights())
                {
                    bits.resize(graph.edge_weights().size());
                    if (!bits.empty())
                        std::memcpy(bits.data(), graph.edge_weights().data(), bits.size() * 4);
                    w = bits.data();
                }
                check(hyt_graph_upload(graph.num_vertices(), graph.num_edges(), graph.row_offsets().data(),
                                       graph.column_indices().data(), w, HYT_GRAPH_WEIGHTS_F32, &g),
                      what);
            }
            ~DeviceGraph()
            {
                if (g)
                    hyt_graph_free(g);
            }
        };

        // Rounds of a variant: SYNC/ASYNC mode (SSSP), direction, and
        // topology-driven = pull every round.
        hyt_run_options options_for(const SEPExecutionVariant &v)
        {
            hyt_run_options o{};
            o.mode = v.is_synchronous() ? HYT_SSSP_SYNC : HYT_SSSP_ASYNC;
            o.direction = (v.is_pull() || v.is_topology_driven()) ? HYT_DIR_PULL : HYT_DIR_PUSH;
            return o;
        }

        SEPExecutionResult finished(const hyt_run_stats &st)
        {
            SEPExecutionResult r = SEPExecutionResult::converged_result();
            r.vertices_processed = static_cast<std::size_t>(st.frontier_total);
            r.updates_generated = static_cast<std::size_t>(st.updates_total);
            r.changed = st.updates_total != 0;
            r.converged = st.converged != 0;
            if (!r.converged)
                r.status = SEPExecutionStatus::SUCCESS;
            return r;
        }
    } // namespace

    void SEPSSSPGpuDriver::initialize(application_type &application, const execution_config_type &config)
    {
        if (dynamic_cast<SEPSSSPApplication *>(&application) == nullptr)
            throw std::invalid_argument("SEPSSSPGpuDriver needs a SEPSSSPApplication");
        base_type::initialize(application, config);
        values_.clear();
        rounds_ = 0;
    }

    SEPExecutionResult SEPSSSPGpuDriver::execute_step()
    {
        if (!initialized())
            return SEPExecutionResult::not_initialized();
        const graph::CSRGraph *g = graph_or_null();
        if (g == nullptr)
            return SEPExecutionResult::invalid_context();
        if (converged())
            return SEPExecutionResult::converged_result();
        const auto &app = static_cast<const SEPSSSPApplication &>(application());
        if (app.source() >= g->num_vertices())
            throw std::invalid_argument("SSSP source vertex is out of range");
        for (const float w : g->edge_weights())
            if (!std::isfinite(w) || w < 0.0F)
                throw std::invalid_argument("SSSP requires finite non-negative edge weights");
        DeviceGraph dg(*g, true, "SEPSSSPGpuDriver");
        const auto n = static_cast<std::size_t>(g->num_vertices());
        std::vector<std::uint32_t> bits(n);
        const std::uint32_t src = static_cast<std::uint32_t>(app.source());
        hyt_run_options o = options_for(config().variant());
        hyt_run_stats st{};
        check(hyt_sssp(dg.g, &src, 1, bits.data(), &o, &st), "SEPSSSPGpuDriver");
        values_.resize(n);
        std::memcpy(values_.data(), bits.data(), n * 4);
        rounds_ = st.rounds;
        mark_converged();
        return finished(st);
    }

    void SEPPageRankGpuDriver::initialize(application_type &application, const execution_config_type &config)
    {
        const auto *app = dynamic_cast<SEPPageRankApplication *>(&application);
        if (app == nullptr)
            throw std::invalid_argument("SEPPageRankGpuDriver needs a SEPPageRankApplication");
        if (!(app->damping() >= 0.0F && app->damping() < 1.0F))
            throw std::invalid_argument("PageRank damping_factor must be in [0, 1)");
        if (!(app->epsilon() > 0.0F) || !std::isfinite(app->epsilon()))
            throw std::invalid_argument("PageRank epsilon must be finite and positive");
        base_type::initialize(application, config);
        values_.clear();
        rounds_ = 0;
    }

    SEPExecutionResult SEPPageRankGpuDriver::execute_step()
    {
        if (!initialized())
            return SEPExecutionResult::not_initialized();
        const graph::CSRGraph *g = graph_or_null();
        if (g == nullptr)
            return SEPExecutionResult::invalid_context();
        if (converged())
            return SEPExecutionResult::converged_result();
        const auto &app = static_cast<const SEPPageRankApplication &>(application());
        DeviceGraph dg(*g, false, "SEPPageRankGpuDriver");
        values_.assign(static_cast<std::size_t>(g->num_vertices()), 0.0F);
        hyt_run_options o = options_for(config().variant());
        hyt_run_stats st{};
        check(hyt_pagerank(dg.g, app.damping(), app.epsilon(), values_.data(), &o, &st), "SEPPageRankGpuDriver");
        rounds_ = st.rounds;
        mark_converged();
        return finished(st);
    }
} // namespace hytgraph::sep
