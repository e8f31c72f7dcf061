// hyt_experiment - one experiment, one structured result.
//
// Same command line as the original project's experiment executable
// (src/main.cpp:41-96): --output <path> [--experiment-id <id>]
// [--algorithm <name>] [--config key=value]..., and the same record
// (runtime::Result, docs/result_schema.json). Where the original writes a
// placeholder measurement, this one runs the workload on the GPU through the
// C ABI and records what it measured.
//
// Config keys (defaults in brackets):
//   graph        rmat | er | <file>.hytcsr | <text edge list>   [rmat]
//   scale, edge_factor, seed, weight_min, weight_max           [16, 16, 1, 1, 64]
//   vertices, weighted   text edge lists only (vertex count, weight column)
//   sources      queries, degree-weighted from seed + 1        [1]
//   source       one fixed source vertex (overrides sources)
//   mode         sync | async (SSSP)                           [sync]
//   max_iterations                                             [0 = |V|]
//   damping, epsilon                                           [0.85, 1e-7 / 1e-6 for ppr]
//   repeats      timed repetitions of the whole query set      [1]
// Algorithms: sssp, pagerank, ppr, sssp_batched (all sources in one call).
#include "hytgraph_amd.h"
#include "runtime/config.hpp"
#include "runtime/logging.hpp"
#include "runtime/result.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
    using hytgraph::runtime::Config;
    using hytgraph::runtime::LogLevel;
    using hytgraph::runtime::Result;
    using Clock = std::chrono::steady_clock;

    std::string utc_timestamp()
    {
        const std::time_t t = std::time(nullptr);
        std::tm utc{};
        gmtime_r(&t, &utc);
        char buffer[32]{};
        std::strftime(buffer, sizeof(buffer), "%Y-%m-%dT%H:%M:%SZ", &utc);
        return buffer;
    }

    void usage(const char *program)
    {
        std::cout << "Usage: " << program
                  << " --output <path> [--experiment-id <id>] [--algorithm <name>]"
                     " [--config key=value]...\n";
    }

    void check(int status)
    {
        if (status != HYT_OK)
            throw std::runtime_error(hyt_last_error());
    }

    double seconds_since(Clock::time_point t0)
    {
        return std::chrono::duration<double>(Clock::now() - t0).count();
    }

    std::int64_t int_or(const Config &c, const std::string &key, std::int64_t fallback)
    {
        return c.contains(key) ? c.get_int(key) : fallback;
    }

    double double_or(const Config &c, const std::string &key, double fallback)
    {
        return c.contains(key) ? c.get_double(key) : fallback;
    }

    // Host graph from the config: generated, binary CSR file or text edge list.
    hyt_host_graph_t host_graph(const Config &c, bool &weights_f32)
    {
        const std::string graph = c.get_or("graph", "rmat");
        hyt_host_graph_t h = nullptr;
        weights_f32 = false;
        if (graph == "rmat" || graph == "er")
        {
            hyt_gen_params p{};
            p.kind = graph == "rmat" ? HYT_GEN_RMAT : HYT_GEN_ER;
            p.scale = static_cast<std::uint32_t>(int_or(c, "scale", 16));
            p.edge_factor = static_cast<std::uint32_t>(int_or(c, "edge_factor", 16));
            p.a = 0.57, p.b = 0.19, p.c = 0.19;
            p.seed = static_cast<std::uint64_t>(int_or(c, "seed", 1));
            p.weight_min = static_cast<std::uint32_t>(int_or(c, "weight_min", 1));
            p.weight_max = static_cast<std::uint32_t>(int_or(c, "weight_max", 64));
            p.memory = HYT_HOST_PAGEABLE;
            check(hyt_generate(&p, &h));
        }
        else if (graph.size() > 7 && graph.compare(graph.size() - 7, 7, ".hytcsr") == 0)
        {
            std::uint32_t f32 = 0;
            check(hyt_host_graph_load(graph.c_str(), HYT_HOST_PAGEABLE, 1, &h, &f32));
            weights_f32 = f32 != 0;
        }
        else
        {
            if (!c.contains("vertices"))
                throw std::invalid_argument("a text edge list needs --config vertices=<count>");
            const bool weighted = c.contains("weighted") && c.get_bool("weighted");
            check(hyt_load_edge_list(graph.c_str(), static_cast<std::uint64_t>(c.get_int("vertices")),
                                     weighted ? 1U : 0U, HYT_HOST_PAGEABLE, &h));
            weights_f32 = weighted;
        }
        return h;
    }

    void run(const Config &c, const std::string &algorithm, Result &r)
    {
        if (algorithm != "sssp" && algorithm != "sssp_batched" && algorithm != "pagerank" && algorithm != "ppr")
            throw std::invalid_argument("unknown algorithm '" + algorithm + "' (sssp, sssp_batched, pagerank, ppr)");
        bool f32 = false;
        auto t0 = Clock::now();
        hyt_host_graph_t h = host_graph(c, f32);
        r.measurements["graph_seconds"] = seconds_since(t0);
        std::uint64_t V = 0, E = 0;
        const std::uint64_t *off = nullptr;
        const std::uint32_t *col = nullptr, *w = nullptr;
        check(hyt_host_graph_view(h, &V, &E, &off, &col, &w));
        r.measurements["vertices"] = static_cast<double>(V);
        r.measurements["edges"] = static_cast<double>(E);

        std::vector<std::uint32_t> sources;
        if (c.contains("source"))
            sources.push_back(static_cast<std::uint32_t>(c.get_int("source")));
        else
        {
            sources.resize(static_cast<std::size_t>(int_or(c, "sources", 1)));
            if (sources.empty() || sources.size() > 64)
                throw std::invalid_argument("sources must be in [1, 64]");
            check(hyt_sample_sources(off, V, static_cast<std::uint64_t>(int_or(c, "seed", 1)) + 1,
                                     static_cast<std::uint32_t>(sources.size()), sources.data()));
        }

        t0 = Clock::now();
        hyt_graph_t g = nullptr;
        check(hyt_graph_upload(V, E, off, col, w, f32 ? HYT_GRAPH_WEIGHTS_F32 : HYT_GRAPH_WEIGHTS_U32, &g));
        r.measurements["upload_seconds"] = seconds_since(t0);

        hyt_run_options o{};
        o.mode = c.get_or("mode", "sync") == "async" ? HYT_SSSP_ASYNC : HYT_SSSP_SYNC;
        o.max_rounds = static_cast<std::uint64_t>(int_or(c, "max_iterations", 0));
        const int repeats = static_cast<int>(int_or(c, "repeats", 1));
        if (repeats < 1)
            throw std::invalid_argument("repeats must be >= 1");
        const float damping = static_cast<float>(double_or(c, "damping", 0.85));
        const std::uint32_t K = static_cast<std::uint32_t>(sources.size());

        double wall = 0, rounds = 0, relaxed = 0, reached_edges = 0, reached = 0, checksum = 0;
        double converged = 1, residual = 0, mass = 0;
        if (algorithm == "sssp" || algorithm == "sssp_batched")
        {
            const bool batched = algorithm == "sssp_batched";
            std::vector<std::uint32_t> dist(static_cast<std::size_t>(V) * (batched ? K : 1));
            for (int rep = 0; rep < repeats; ++rep)
            {
                rounds = relaxed = reached_edges = reached = checksum = 0;
                converged = 1;
                wall = 0;
                for (std::uint32_t k = 0; k < (batched ? 1U : K); ++k)
                {
                    // timed: the call, including its copy of the distances
                    // to host memory; not the checksum pass below
                    hyt_run_stats st{};
                    const auto q0 = Clock::now();
                    check(hyt_sssp(g, batched ? sources.data() : &sources[k], batched ? K : 1, dist.data(), &o,
                                   &st));
                    wall += seconds_since(q0);
                    rounds += static_cast<double>(st.rounds);
                    relaxed += static_cast<double>(st.edges_total);
                    converged = std::min(converged, static_cast<double>(st.converged));
                    const std::uint32_t cols = batched ? K : 1;
                    const std::uint32_t inf = f32 ? 0x7F800000U : 0xFFFFFFFFU;
                    for (std::uint64_t v = 0; v < V; ++v)
                        for (std::uint32_t j = 0; j < cols; ++j)
                        {
                            const std::uint32_t d = dist[v * cols + j];
                            if (d != inf)
                            {
                                reached += 1;
                                reached_edges += static_cast<double>(off[v + 1] - off[v]);
                                float df;
                                std::memcpy(&df, &d, 4);
                                checksum += f32 ? static_cast<double>(df) : static_cast<double>(d);
                            }
                        }
                }
            }
            r.measurements["gteps"] = reached_edges / wall / 1e9;
            r.measurements["queries_per_second"] = K / wall;
            r.measurements["reached_vertices"] = reached;
            r.measurements["distance_sum"] = checksum;
        }
        else if (algorithm == "pagerank" || algorithm == "ppr")
        {
            const bool ppr = algorithm == "ppr";
            const float eps = static_cast<float>(double_or(c, "epsilon", ppr ? 1e-6 : 1e-7));
            std::vector<float> ranks(static_cast<std::size_t>(V) * (ppr ? K : 1));
            for (int rep = 0; rep < repeats; ++rep)
            {
                hyt_run_stats st{};
                const auto q0 = Clock::now();
                if (ppr)
                    check(hyt_ppr(g, sources.data(), K, damping, eps, ranks.data(), &o, &st));
                else
                    check(hyt_pagerank(g, damping, eps, ranks.data(), &o, &st));
                wall = seconds_since(q0);
                rounds = static_cast<double>(st.rounds);
                relaxed = static_cast<double>(st.edges_total);
                converged = st.converged;
                residual = st.residual;
            }
            for (const float x : ranks)
                mass += x;
            r.measurements["rank_sum"] = mass / (ppr ? K : 1);
            r.measurements["residual"] = residual;
            if (ppr)
                r.measurements["queries_per_second"] = K / wall;
        }
        r.measurements["wall_seconds"] = wall;
        r.measurements["rounds"] = rounds;
        r.measurements["edges_relaxed"] = relaxed;
        r.measurements["converged"] = converged;
        std::uint64_t dev_bytes = 0;
        check(hyt_graph_info(g, nullptr, nullptr, &dev_bytes));
        r.measurements["device_bytes"] = static_cast<double>(dev_bytes);
        check(hyt_graph_free(g));
        check(hyt_host_graph_free(h));
    }
} // namespace

int main(int argc, char **argv)
{
    using namespace hytgraph::runtime;

    std::string output;
    std::string experiment_id = "experiment";
    Config config;
    for (int i = 1; i < argc; ++i)
    {
        const std::string arg = argv[i];
        if (arg == "--output" && i + 1 < argc)
            output = argv[++i];
        else if (arg == "--experiment-id" && i + 1 < argc)
            experiment_id = argv[++i];
        else if (arg == "--algorithm" && i + 1 < argc)
            config.set("algorithm", argv[++i]);
        else if (arg == "--config" && i + 1 < argc)
        {
            const std::string pair = argv[++i];
            const auto pos = pair.find('=');
            if (pos == std::string::npos || pos == 0)
            {
                std::cerr << "Invalid --config value: " << pair << '\n';
                return 2;
            }
            config.set(pair.substr(0, pos), pair.substr(pos + 1));
        }
        else if (arg == "--help")
        {
            usage(argv[0]);
            return 0;
        }
        else
        {
            std::cerr << "Unknown or incomplete argument: " << arg << '\n';
            usage(argv[0]);
            return 2;
        }
    }
    if (output.empty())
    {
        std::cerr << "--output is required\n";
        usage(argv[0]);
        return 2;
    }
    const std::filesystem::path output_path(output);
    if (output_path.has_parent_path())
        std::filesystem::create_directories(output_path.parent_path());

    Result result;
    result.experiment_id = experiment_id;
    result.timestamp_utc = utc_timestamp();
    result.config = config.values();
    result.config["library_version"] = hyt_version();
    const std::string algorithm = config.get_or("algorithm", "sssp");
    int rc = 0;
    try
    {
        log(LogLevel::Info, "running " + algorithm);
        run(config, algorithm, result);
    }
    catch (const std::exception &e)
    {
        result.status = "failure";
        result.config["error"] = e.what();
        log(LogLevel::Error, std::string("experiment failed: ") + e.what());
        rc = 1;
    }
    if (!write_result_json(result, output))
    {
        log(LogLevel::Error, "failed to write result: " + output);
        return 1;
    }
    log(LogLevel::Info, "wrote structured result: " + output);
    return rc;
}
