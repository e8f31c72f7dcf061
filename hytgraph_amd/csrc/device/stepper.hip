// Round-by-round single-source runs (hyt_stepper_*): the rounds of hyt_sssp
// (one source) or hyt_pagerank / hyt_ppr (one seed), advanced a given number
// of rounds per call on the graph's workspace and read between calls. The SEP
// drivers' execute_step() runs one round through it.
#include "dgraph.hpp"

#include <chrono>
#include <cmath>
#include <memory>

struct hyt_stepper
{
    hyt_graph *graph = nullptr; // borrowed
    std::uint32_t kind = HYT_STEP_SSSP;
    std::uint32_t vertex = 0;
    float damping = 0.85F, epsilon = 1e-6F;
    hyt_run_options o{};
    bool started = false, finished = false;
    hyt_run_stats total{};
};

namespace
{
    void add_stats(hyt_run_stats &t, const hyt_run_stats &r)
    {
        t.rounds += r.rounds;
        t.frontier_total += r.frontier_total;
        t.edges_total += r.edges_total;
        t.updates_total += r.updates_total;
        t.expand_ms += r.expand_ms;
        t.expand_launches += r.expand_launches;
        t.pull_rounds += r.pull_rounds;
        t.pull_active_edges += r.pull_active_edges;
        t.pull_edges += r.pull_edges;
        t.pull_frontier += r.pull_frontier;
        t.heavy_edges += r.heavy_edges;
        t.overflow |= r.overflow;
        t.converged = r.converged;
    }
} // namespace

extern "C" {

int hyt_stepper_create(hyt_graph_t g, uint32_t kind, uint32_t vertex, float damping, float epsilon,
                       const hyt_run_options *opts, hyt_stepper_t *out)
{
    return hyt::guarded("hyt_stepper_create", [&] {
        using namespace hyt;
        require(out != nullptr, "out is null");
        *out = nullptr;
        require(g != nullptr, "null graph");
        require(kind == HYT_STEP_SSSP || kind == HYT_STEP_PAGERANK, "unknown stepper kind");
        if (kind == HYT_STEP_SSSP)
            require(vertex < g->g.V, "SSSP source vertex is out of range");
        else
        {
            require(vertex == 0xFFFFFFFFU || vertex < g->g.V, "PPR seed vertex is out of range");
            require(damping >= 0.0F && damping < 1.0F, "PageRank damping_factor must be in [0, 1)");
            require(epsilon > 0.0F && std::isfinite(epsilon), "PageRank epsilon must be finite and positive");
        }
        auto s = std::make_unique<hyt_stepper>();
        s->graph = g;
        s->kind = kind;
        s->vertex = vertex;
        s->damping = damping;
        s->epsilon = epsilon;
        if (opts)
            s->o = *opts;
        require(s->o.tolerance == 0.0F && s->o.delta == 0, "steppers run tolerance-0 rounds without buckets");
        require(s->o.mode == HYT_SSSP_SYNC || s->o.mode == HYT_SSSP_ASYNC, "unknown SSSP mode");
        s->o.max_rounds = 0;
        *out = s.release();
    });
}

int hyt_stepper_step(hyt_stepper_t s, uint64_t rounds, hyt_run_stats *stats, int32_t *finished)
{
    return hyt::guarded("hyt_stepper_step", [&] {
        using namespace hyt;
        require(s != nullptr, "null stepper");
        require(rounds >= 1, "rounds must be >= 1");
        DeviceGraph &g = s->graph->g;
        hip_check(hipSetDevice(g.device), "hipSetDevice");
        hyt_run_stats st{};
        if (!s->finished)
        {
            const auto t0 = std::chrono::steady_clock::now();
            if (s->kind == HYT_STEP_SSSP)
                sssp_steps(g, s->vertex, s->o, s->started, rounds, st);
            else
                pagerank_steps(g, s->vertex, s->damping, s->epsilon, s->o, s->started, rounds, st);
            st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            s->started = true;
            s->finished = st.converged != 0;
            add_stats(s->total, st);
        }
        else
            st.converged = 1;
        if (stats)
            *stats = st;
        if (finished)
            *finished = s->finished ? 1 : 0;
    });
}

int hyt_stepper_read(hyt_stepper_t s, void *out, hyt_run_stats *total)
{
    return hyt::guarded("hyt_stepper_read", [&] {
        using namespace hyt;
        require(s != nullptr && out != nullptr, "null stepper or output");
        require(s->started, "the stepper has run no round");
        DeviceGraph &g = s->graph->g;
        hip_check(hipSetDevice(g.device), "hipSetDevice");
        if (s->kind == HYT_STEP_SSSP)
        {
            hip_check(hipMemcpyAsync(out, g.ws.dist.ptr, g.V * 4,
                                     s->o.out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, g.stream),
                      "copy distances");
            hip_check(hipStreamSynchronize(g.stream), "stepper read");
        }
        else
        {
            hyt_run_stats st{};
            finalize_ranks(g, static_cast<float *>(out), s->o.out_on_device != 0, st);
            s->total.residual = st.residual;
        }
        if (total)
            *total = s->total;
    });
}

int hyt_stepper_free(hyt_stepper_t s)
{
    return hyt::guarded("hyt_stepper_free", [&] { delete s; });
}

} // extern "C"
