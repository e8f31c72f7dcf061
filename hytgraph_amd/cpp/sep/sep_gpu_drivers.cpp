// GPU execution of the SEP SSSP and PageRank applications through the C ABI,
// one round per execute_step() (hyt_stepper_*).
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

        // Rounds of a variant: SYNC/ASYNC mode (SSSP), direction, and
        // topology-driven = pull every round.
        hyt_run_options options_for(const SEPExecutionVariant &v)
        {
            hyt_run_options o{};
            o.mode = v.is_synchronous() ? HYT_SSSP_SYNC : HYT_SSSP_ASYNC;
            o.direction = (v.is_pull() || v.is_topology_driven()) ? HYT_DIR_PULL : HYT_DIR_PUSH;
            return o;
        }

        SEPExecutionResult round_result(const hyt_run_stats &st, bool finished)
        {
            SEPExecutionResult r = finished ? SEPExecutionResult::converged_result()
                                            : SEPExecutionResult::success();
            r.vertices_processed = static_cast<std::size_t>(st.frontier_total);
            r.updates_generated = static_cast<std::size_t>(st.updates_total);
            r.changed = st.updates_total != 0;
            return r;
        }
    } // namespace

    // The device graph of the context's CSRGraph and the stepper on it.
    struct SEPGpuRun
    {
        const graph::CSRGraph *host = nullptr;
        hyt_graph_t g = nullptr;
        hyt_stepper_t s = nullptr;

        SEPGpuRun(const graph::CSRGraph &graph, bool weights, std::uint32_t kind, std::uint32_t vertex,
                  float damping, float epsilon, const hyt_run_options &o, const char *what)
            : host(&graph)
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
            const int rc = hyt_stepper_create(g, kind, vertex, damping, epsilon, &o, &s);
            if (rc != HYT_OK)
            {
                const std::string msg = std::string(what) + ": " + hyt_last_error();
                hyt_graph_free(g);
                g = nullptr;
                if (rc == HYT_ERR_INVALID_ARGUMENT)
                    throw std::invalid_argument(msg);
                throw std::runtime_error(msg);
            }
        }
        ~SEPGpuRun()
        {
            if (s)
                hyt_stepper_free(s);
            if (g)
                hyt_graph_free(g);
        }
        SEPGpuRun(const SEPGpuRun &) = delete;
        SEPGpuRun &operator=(const SEPGpuRun &) = delete;

        // One round; the state after it into values.
        hyt_run_stats step(std::vector<float> &values, bool &finished, const char *what)
        {
            hyt_run_stats st{};
            std::int32_t fin = 0;
            check(hyt_stepper_step(s, 1, &st, &fin), what);
            values.resize(static_cast<std::size_t>(host->num_vertices()));
            check(hyt_stepper_read(s, values.data(), nullptr), what);
            finished = fin != 0;
            return st;
        }
    };

    void SEPSSSPGpuDriver::initialize(application_type &application, const execution_config_type &config)
    {
        if (dynamic_cast<SEPSSSPApplication *>(&application) == nullptr)
            throw std::invalid_argument("SEPSSSPGpuDriver needs a SEPSSSPApplication");
        base_type::initialize(application, config);
        run_.reset();
        values_.clear();
        rounds_ = 0;
        persistent_ = false;
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
        if (!run_ || run_->host != g)
        {
            const auto &app = static_cast<const SEPSSSPApplication &>(application());
            if (app.source() >= g->num_vertices())
                throw std::invalid_argument("SSSP source vertex is out of range");
            for (const float w : g->edge_weights())
                if (!std::isfinite(w) || w < 0.0F)
                    throw std::invalid_argument("SSSP requires finite non-negative edge weights");
            run_ = std::make_shared<SEPGpuRun>(*g, true, HYT_STEP_SSSP, static_cast<std::uint32_t>(app.source()),
                                               0.0F, 0.0F, options_for(config().variant()), "SEPSSSPGpuDriver");
            rounds_ = 0;
        }
        bool finished = false;
        const hyt_run_stats st = run_->step(values_, finished, "SEPSSSPGpuDriver");
        rounds_ += st.rounds;
        if (finished)
            mark_converged();
        return round_result(st, finished);
    }

    SEPExecutionResult SEPSSSPGpuDriver::run_to_convergence()
    {
        persistent_ = false;
        if (!initialized())
            return SEPExecutionResult::not_initialized();
        const graph::CSRGraph *g = graph_or_null();
        if (g == nullptr)
            return SEPExecutionResult::invalid_context();
        if (converged())
            return SEPExecutionResult::converged_result();
        const SEPExecutionVariant &v = config().variant();
        if ((!run_ || run_->host != g) && !v.is_pull() && !v.is_topology_driven())
        {
            // a fresh run: every round in one persistent launch
            const auto &app = static_cast<const SEPSSSPApplication &>(application());
            if (app.source() >= g->num_vertices())
                throw std::invalid_argument("SSSP source vertex is out of range");
            for (const float w : g->edge_weights())
                if (!std::isfinite(w) || w < 0.0F)
                    throw std::invalid_argument("SSSP requires finite non-negative edge weights");
            hyt_run_options o = options_for(v);
            run_ = std::make_shared<SEPGpuRun>(*g, true, HYT_STEP_SSSP, static_cast<std::uint32_t>(app.source()),
                                               0.0F, 0.0F, o, "SEPSSSPGpuDriver");
            o.persistent = 1;
            const std::uint32_t src = static_cast<std::uint32_t>(app.source());
            std::vector<std::uint32_t> bits(static_cast<std::size_t>(g->num_vertices()));
            hyt_run_stats st{};
            check(hyt_sssp(run_->g, &src, 1, bits.data(), &o, &st), "SEPSSSPGpuDriver");
            values_.resize(bits.size());
            if (!bits.empty())
                std::memcpy(values_.data(), bits.data(), bits.size() * 4);
            rounds_ = st.rounds;
            persistent_ = true;
            // the stepper of run_ never ran: a later step finds the run converged
            mark_converged();
            return round_result(st, true);
        }
        hyt_run_stats total{};
        SEPExecutionResult r;
        do
        {
            r = execute_step();
            if (!r.succeeded())
                return r;
            total.frontier_total += r.vertices_processed;
            total.updates_total += r.updates_generated;
        } while (!r.reached_convergence());
        return round_result(total, true);
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
        run_.reset();
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
        if (!run_ || run_->host != g)
        {
            const auto &app = static_cast<const SEPPageRankApplication &>(application());
            run_ = std::make_shared<SEPGpuRun>(*g, false, HYT_STEP_PAGERANK, 0xFFFFFFFFU, app.damping(), app.epsilon(),
                                               options_for(config().variant()), "SEPPageRankGpuDriver");
            rounds_ = 0;
        }
        bool finished = false;
        const hyt_run_stats st = run_->step(values_, finished, "SEPPageRankGpuDriver");
        rounds_ += st.rounds;
        if (finished)
            mark_converged();
        return round_result(st, finished);
    }
} // namespace hytgraph::sep
