// Autotuner over the library's own entry points: every candidate setting is
// run on the caller's sources and timed on the host around the whole call
// (the calls synchronise before returning); the fastest mean wins.
#include "common.hpp"
#include "splits.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <limits>
#include <vector>

namespace hyt
{
    namespace
    {
        using Clock = std::chrono::steady_clock;

        void ok(int status)
        {
            if (status != HYT_OK)
                fail(status, hyt_last_error());
        }

        // Device memory for results, so timings exclude copies to the host.
        struct DeviceOut
        {
            void *p = nullptr;
            explicit DeviceOut(std::size_t bytes)
            {
                if (hipMalloc(&p, bytes ? bytes : 4) != hipSuccess)
                    fail(HYT_ERR_OUT_OF_MEMORY, "hipMalloc(tuning output)");
            }
            ~DeviceOut() { (void)hipFree(p); }
            std::uint32_t *data() const { return static_cast<std::uint32_t *>(p); }
        };

        // Best (over reps) mean milliseconds per query of single-source async
        // SSSP from every source with the given split.
        double time_split(hyt_graph_t g, const std::uint32_t *src, std::uint32_t n, std::uint32_t light,
                          std::uint32_t chunk, std::uint32_t reps, const DeviceOut &out)
        {
            hyt_run_options o{};
            o.mode = HYT_SSSP_ASYNC;
            o.out_on_device = 1;
            o.light_max_degree = light;
            o.heavy_chunk = chunk;
            double best = std::numeric_limits<double>::infinity();
            for (std::uint32_t r = 0; r < reps; ++r)
            {
                const auto t0 = Clock::now();
                for (std::uint32_t i = 0; i < n; ++i)
                    ok(hyt_sssp(g, &src[i], 1, out.data(), &o, nullptr));
                best = std::min(best, std::chrono::duration<double, std::milli>(Clock::now() - t0).count() / n);
            }
            return best;
        }
    } // namespace
} // namespace hyt

extern "C" {

int hyt_tune(hyt_graph_t g, const uint32_t *sources, uint32_t num_sources, const hyt_tune_options *opts,
             hyt_tune_result *out)
{
    return hyt::guarded("hyt_tune", [&] {
        using namespace hyt;
        require(g != nullptr && sources != nullptr && out != nullptr && opts != nullptr, "null pointer");
        require(num_sources >= 1, "at least one source is needed");
        const std::uint32_t reps = opts->reps ? opts->reps : 2;
        const std::uint32_t max_lanes = opts->max_lanes ? opts->max_lanes : 64;
        require(max_lanes >= 2 && max_lanes <= 64, "max_lanes must be in [2, 64]");
        std::uint64_t V = 0, E = 0;
        ok(hyt_graph_info(g, &V, &E, nullptr));
        for (std::uint32_t i = 0; i < num_sources; ++i)
            require(sources[i] < V, "source vertex is out of range");
        hyt_tune_result r{};
        std::uint32_t light0 = 0, chunk0 = 0;
        default_sssp_split(E, light0, chunk0);
        r.light_max_degree = light0;
        r.heavy_chunk = chunk0;
        DeviceOut buf(V * max_lanes * 4);
        if (opts->tune_split)
        {
            // warm-up (workspace allocation), then the graph's default split
            (void)time_split(g, sources, 1, light0, chunk0, 1, buf);
            r.default_sssp_ms = time_split(g, sources, num_sources, light0, chunk0, reps, buf);
            r.sssp_ms = r.default_sssp_ms;
            r.candidates = 1;
            for (std::uint32_t light : {64U, 128U, 256U, 512U, 1024U, 2048U})
            {
                if (light == light0)
                    continue;
                const double t = time_split(g, sources, num_sources, light, chunk0, reps, buf);
                ++r.candidates;
                if (t < r.sssp_ms)
                {
                    r.sssp_ms = t;
                    r.light_max_degree = light;
                }
            }
            for (std::uint32_t chunk : {256U, 512U, 1024U, 2048U, 4096U, 8192U})
            {
                if (chunk == chunk0)
                    continue;
                const double t = time_split(g, sources, num_sources, r.light_max_degree, chunk, reps, buf);
                ++r.candidates;
                if (t < r.sssp_ms)
                {
                    r.sssp_ms = t;
                    r.heavy_chunk = chunk;
                }
            }
        }
        if (opts->tune_lanes)
        {
            hyt_run_options o{};
            o.out_on_device = 1;
            o.light_max_degree = r.light_max_degree;
            o.heavy_chunk = r.heavy_chunk;
            std::vector<std::uint32_t> src(max_lanes);
            r.lanes = 2;
            r.lanes_qps = 0;
            for (std::uint32_t K = 2; K <= max_lanes; K *= 2)
            {
                for (std::uint32_t k = 0; k < K; ++k)
                    src[k] = sources[k % num_sources];
                ok(hyt_sssp(g, src.data(), K, buf.data(), &o, nullptr)); // warm-up (K-lane workspace)
                double best = std::numeric_limits<double>::infinity();
                for (std::uint32_t rep = 0; rep < reps; ++rep)
                {
                    const auto t0 = Clock::now();
                    ok(hyt_sssp(g, src.data(), K, buf.data(), &o, nullptr));
                    best = std::min(best, std::chrono::duration<double>(Clock::now() - t0).count());
                }
                ++r.candidates;
                const double qps = K / best;
                if (qps > r.lanes_qps)
                {
                    r.lanes_qps = qps;
                    r.lanes = K;
                }
                if (K < max_lanes && K * 2 > max_lanes)
                    K = max_lanes / 2; // last step lands on max_lanes
            }
        }
        *out = r;
    });
}

int hyt_tune_hybrid(hyt_host_graph_t host, uint32_t weights_f32, const hyt_hybrid_config *config,
                    const uint64_t *candidates, uint32_t num_candidates, const uint32_t *sources,
                    uint32_t num_sources, hyt_tune_result *out)
{
    return hyt::guarded("hyt_tune_hybrid", [&] {
        using namespace hyt;
        require(host != nullptr && config != nullptr && candidates != nullptr && sources != nullptr &&
                    out != nullptr,
                "null pointer");
        require(num_candidates >= 1 && num_sources >= 1, "candidates and sources must be non-empty");
        std::uint64_t V = 0;
        ok(hyt_host_graph_view(host, &V, nullptr, nullptr, nullptr, nullptr));
        for (std::uint32_t i = 0; i < num_sources; ++i)
            require(sources[i] < V, "source vertex is out of range");
        hyt_tune_result r{};
        r.hybrid_ms = std::numeric_limits<double>::infinity();
        DeviceOut dist(V * 4);
        hyt_run_options o{};
        o.out_on_device = 1;
        for (std::uint32_t c = 0; c < num_candidates; ++c)
        {
            hyt_hybrid_config cfg = *config;
            cfg.partition_edges = candidates[c];
            hyt_hybrid_t h = nullptr;
            ok(hyt_hybrid_create(host, weights_f32, &cfg, &h));
            double total = 0;
            try
            {
                for (std::uint32_t i = 0; i < num_sources; ++i)
                {
                    const auto t0 = Clock::now();
                    ok(hyt_hybrid_sssp(h, sources[i], dist.data(), &o, nullptr));
                    total += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
                }
            }
            catch (...)
            {
                hyt_hybrid_free(h);
                throw;
            }
            ok(hyt_hybrid_free(h));
            ++r.candidates;
            if (total / num_sources < r.hybrid_ms)
            {
                r.hybrid_ms = total / num_sources;
                r.partition_edges = candidates[c];
            }
        }
        *out = r;
    });
}

} // extern "C"
