// C++ API tests in the original project's style (expect(cond, msg) throwing,
// exit code 77 when a GPU test finds no device - tests/cuda_algorithm_tests.cu:91-99).
//   cpp_api_test cpu   host-side API: CSRGraph, contributions
//   cpp_api_test gpu   drop-in sssp_gpu / pagerank_gpu vs the original's known answers
#include "algorithms/pagerank.hpp"
#include "algorithms/sssp.hpp"
#include "graph/csr_graph.hpp"
#include "hytgraph_amd.h"
#include "scheduling/hub_sort.hpp"
#include "sep/sep_execution_factory.hpp"
#include "sep/sep_execution_variant_registry.hpp"
#include "sep/sep_gpu_drivers.hpp"

#include <cmath>
#include <cstring>
#include <deque>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
    using hytgraph::graph::CSRGraph;
    namespace alg = hytgraph::algorithms;

    void expect(bool condition, const std::string &message)
    {
        if (!condition)
            throw std::runtime_error("FAIL: " + message);
    }

    template <class Ex, class Fn>
    void expect_throw(Fn fn, const std::string &message)
    {
        bool threw = false;
        try
        {
            fn();
        }
        catch (const Ex &)
        {
            threw = true;
        }
        expect(threw, message);
    }

    // ------------------------------------------------------------- host
    void test_csr_validation()
    {
        CSRGraph g(4, {0, 2, 4, 5, 5}, {1, 2, 2, 3, 3}, {1, 4, 2, 6, 1});
        expect(g.num_vertices() == 4 && g.num_edges() == 5 && g.has_weights(), "CSR sizes");
        expect(g.out_degree(1) == 2 && g.neighbor_at(4) == 3 && g.weight_at(1) == 4.0F, "CSR accessors");
        expect_throw<std::out_of_range>([&] { (void)g.neighbor_range(4); }, "neighbor_range range check");
        expect_throw<std::out_of_range>([&] { (void)g.neighbor_at(5); }, "neighbor_at range check");
        expect_throw<std::invalid_argument>([] { CSRGraph(2, {0, 1}, {1}); }, "offset count");
        expect_throw<std::invalid_argument>([] { CSRGraph(2, {1, 1, 1}, {1}); }, "first offset zero");
        expect_throw<std::invalid_argument>([] { CSRGraph(2, {0, 2, 1}, {1, 0}); }, "monotone offsets");
        expect_throw<std::invalid_argument>([] { CSRGraph(2, {0, 1, 1}, {2}); }, "column range");
        expect_throw<std::invalid_argument>([] { CSRGraph(2, {0, 1, 1}, {1}, {1, 2}); }, "weight count");
        CSRGraph empty(0, {0}, {});
        expect(empty.num_vertices() == 0, "empty graph is valid");
        CSRGraph unweighted(3, {0, 1, 2, 2}, {1, 0});
        expect_throw<std::logic_error>([&] { (void)unweighted.weight_at(0); }, "weight_at on unweighted");
    }

    void test_reorder()
    {
        // 0->1 (1), 0->2 (2), 1->2 (3); order[new] = old = {2, 0, 1}
        CSRGraph g(3, {0, 2, 3, 3}, {1, 2, 2}, {1, 2, 3});
        g.reorder_vertices({2, 0, 1});
        expect(g.row_offsets() == std::vector<CSRGraph::offset_type>({0, 0, 2, 3}), "reordered offsets");
        expect(g.column_indices() == std::vector<CSRGraph::vertex_id>({2, 0, 0}), "renamed destinations");
        expect(g.edge_weights() == std::vector<float>({1, 2, 3}), "weights follow their edges");
        expect_throw<std::invalid_argument>([&] { g.reorder_vertices({0, 0, 1}); }, "duplicate ids");
        expect_throw<std::invalid_argument>([&] { g.reorder_vertices({0, 1}); }, "short order");
    }

    void test_contributions()
    {
        // unit_tests.cpp:2922-2962: one step from uniform on 0->1, 0->2, 1->0, 2->0
        CSRGraph g(3, {0, 2, 3, 4}, {1, 2, 0, 0});
        auto c = alg::pagerank_contributions(g, {1.0F, 0.0F, 0.0F});
        // next = 0.05 + 0.85 * in-sum: v0 <- 0, v1,v2 <- 0.5
        expect(std::fabs(c[0].contribution - 0.95) < 1e-6, "contribution v0");
        expect(std::fabs(c[1].contribution - 0.475) < 1e-6, "contribution v1");
        expect(std::fabs(c[2].contribution - 0.475) < 1e-6, "contribution v2");
        const float inf = std::numeric_limits<float>::infinity();
        auto s = alg::sssp_contributions({5.0F, inf, 3.0F, inf}, {2.0F, 7.0F, 3.0F, inf});
        expect(s[0].contribution == 3.0 && s[1].contribution == 1.0 && s[2].contribution == 0.0 &&
                   s[3].contribution == 0.0,
               "sssp contributions");
        expect_throw<std::invalid_argument>([] { (void)alg::sssp_contributions({1.0F}, {}); }, "size mismatch");
    }

    // A FIFO frontier for the data-driven SEP variants.
    class QueueFrontier final : public hytgraph::sep::SEPFrontier
    {
    public:
        void clear() noexcept override { q_.clear(); }
        void push(node_id_type n) override { q_.push_back(n); }
        bool pop(node_id_type &n) override
        {
            if (q_.empty())
                return false;
            n = q_.front();
            q_.pop_front();
            return true;
        }
        std::size_t size() const noexcept override { return q_.size(); }
        bool empty() const noexcept override { return q_.empty(); }

    private:
        std::deque<node_id_type> q_;
    };

    void test_sep_host()
    {
        namespace sep = hytgraph::sep;
        sep::SEPSSSPApplication app(0);
        float d = 10.0F;
        auto r = app.accumulate_buffer(0, 1, 2.0F, d, 3.0F);
        expect(r.changed && r.activate && d == 5.0F, "SEP SSSP accumulate lowers the buffer");
        r = app.accumulate_buffer(0, 1, 2.0F, d, 4.0F);
        expect(!r.changed && d == 5.0F, "SEP SSSP accumulate keeps a smaller buffer");
        expect(app.init_value(0) == 0.0F && std::isinf(app.init_value(1)) && app.requires_edge_weights(),
               "SEP SSSP init");
        sep::SEPPageRankApplication pr(0.85F, 1e-3F);
        float res = 0.0F;
        expect(pr.accumulate_buffer(0, 1, res, 2e-3F).activate && !pr.accumulate_buffer(0, 1, res, 1e-3F).activate,
               "SEP PageRank activates on crossing epsilon");
        sep::SEPSSSPGpuDriver drv;
        expect(drv.execute_step().status == sep::SEPExecutionStatus::NOT_INITIALIZED, "driver not initialised");
        drv.initialize(app, sep::SEPExecutionConfig{});
        expect(drv.execute_step().status == sep::SEPExecutionStatus::INVALID_CONTEXT, "driver without context");
        CSRGraph g(2, {0, 1, 1}, {1}, {1.0F});
        sep::SEPExecutionContext no_frontier(const_cast<CSRGraph *>(&g));
        drv.set_context(no_frontier);
        expect(drv.execute_step().status == sep::SEPExecutionStatus::INVALID_CONTEXT,
               "data-driven variant needs a frontier");
        sep::SEPPageRankApplication bad(1.0F, 1e-6F);
        expect_throw<std::invalid_argument>([&] { sep::SEPPageRankGpuDriver().initialize(bad, {}); },
                                            "PageRank driver rejects damping 1");
        auto deferred = sep::SEPExecutionFactory<float, float, float>::create("ASYNC_PUSH_DD");
        expect(deferred->execute_step().status == sep::SEPExecutionStatus::INVALID_CONFIGURATION,
               "generic factory drivers stay deferred");
    }

    // -------------------------------------------------------------- gpu
    void test_sssp_gpu_known_answers()
    {
        // algorithm_tests.cpp:100-129 / cuda_algorithm_tests.cu:58-87
        CSRGraph g(4, {0, 2, 4, 5, 5}, {1, 2, 2, 3, 3}, {1.0F, 4.0F, 2.0F, 6.0F, 1.0F});
        alg::SSSPOptions o;
        o.source = 0;
        const auto r = alg::sssp_gpu(g, o);
        expect(r.converged, "GPU SSSP converged");
        expect(r.distances == std::vector<float>({0.0F, 1.0F, 3.0F, 4.0F}), "GPU SSSP distances");
        // algorithm_tests.cpp:131-157: unweighted chain, unreachable vertex
        CSRGraph chain(4, {0, 1, 2, 2, 2}, {1, 2});
        const auto c = alg::sssp_gpu(chain, o);
        expect(c.distances[0] == 0.0F && c.distances[1] == 1.0F && c.distances[2] == 2.0F &&
                   std::isinf(c.distances[3]),
               "GPU SSSP unweighted / unreachable");
        expect(c.iterations == 3, "GPU SSSP iteration count equals the original's");
        // algorithm_tests.cpp:159-206
        expect_throw<std::invalid_argument>([] { (void)alg::sssp_gpu(CSRGraph(2, {0, 1, 1}, {1}, {-1.0F})); },
                                            "negative weight rejected");
        alg::SSSPOptions bad;
        bad.source = 2;
        expect_throw<std::invalid_argument>([&] { (void)alg::sssp_gpu(CSRGraph(2, {0, 0, 0}, {}), bad); },
                                            "out-of-range source rejected");
        auto batch = alg::sssp_gpu_batch(g, {0, 1, 3});
        expect(batch.distances[0] == r.distances, "batched column 0");
        expect(batch.distances[1][2] == 2.0F && std::isinf(batch.distances[1][0]), "batched column 1");
        expect(batch.distances[2][3] == 0.0F, "batched column 2");
    }

    void test_pagerank_gpu_known_answers()
    {
        // cuda_algorithm_tests.cu:23-56 graph
        CSRGraph g(3, {0, 2, 3, 4}, {1, 2, 2, 0});
        alg::PageRankOptions o;
        o.max_iterations = 200;
        o.tolerance = 1.0e-6F;
        const auto r = alg::pagerank_gpu(g, o);
        expect(r.converged, "GPU PageRank converged");
        // fixed point of this graph (FP64 power iteration to 1e-14)
        const double x0 = 0.387789712, x1 = 0.214810627, x2 = 0.397399661;
        expect(std::fabs(r.ranks[0] - x0) <= 1e-5 && std::fabs(r.ranks[1] - x1) <= 1e-5 &&
                   std::fabs(r.ranks[2] - x2) <= 1e-5,
               "GPU PageRank matches the fixed point within 1e-5");
        // algorithm_tests.cpp:22-47: 3-cycle -> 1/3
        const auto cyc = alg::pagerank_gpu(CSRGraph(3, {0, 1, 2, 3}, {1, 2, 0}), o);
        for (float v : cyc.ranks)
            expect(std::fabs(v - 1.0F / 3.0F) <= 1e-5F, "cycle ranks uniform");
        // algorithm_tests.cpp:49-75: dangling vertex, mass preserved
        const auto dg = alg::pagerank_gpu(CSRGraph(3, {0, 1, 2, 2}, {1, 0}), o);
        expect(std::fabs(dg.ranks[0] + dg.ranks[1] + dg.ranks[2] - 1.0F) <= 1e-5F, "mass preserved");
        // algorithm_tests.cpp:77-98
        alg::PageRankOptions bad;
        bad.damping_factor = 1.0F;
        expect_throw<std::invalid_argument>([&] { (void)alg::pagerank_gpu(CSRGraph(1, {0, 0}, {}), bad); },
                                            "damping 1 rejected");
        auto ppr = alg::ppr_gpu(g, {0, 2}, 0.85F, 1e-9F);
        expect(std::fabs(ppr.ranks[0][0] + ppr.ranks[0][1] + ppr.ranks[0][2] - 1.0F) <= 1e-5F, "ppr mass");
    }
    void test_sep_gpu()
    {
        namespace sep = hytgraph::sep;
        CSRGraph g(4, {0, 2, 4, 5, 5}, {1, 2, 2, 3, 3}, {1.0F, 4.0F, 2.0F, 6.0F, 1.0F});
        CSRGraph cyc(3, {0, 1, 2, 3}, {1, 2, 0});
        QueueFrontier frontier;
        for (const auto &v : sep::SEPExecutionVariantRegistry::all())
        {
            sep::SEPSSSPApplication app(0);
            sep::SEPSSSPGpuDriver drv;
            drv.initialize(app, sep::SEPExecutionConfig(v));
            sep::SEPExecutionContext ctx(const_cast<CSRGraph *>(&g), nullptr, nullptr, nullptr, &frontier);
            drv.set_context(ctx);
            // one round per step until a round finds nothing left
            int steps = 0;
            sep::SEPExecutionResult r;
            do
            {
                r = drv.execute_step();
                ++steps;
                expect(r.succeeded(), "SEP SSSP step succeeds: " + v.to_string());
            } while (!r.reached_convergence() && steps < 100);
            expect(drv.converged() && static_cast<std::uint64_t>(steps) == drv.rounds(),
                   "SEP SSSP converges one round per step: " + v.to_string());
            expect(drv.values() == std::vector<float>({0.0F, 1.0F, 3.0F, 4.0F}), "SEP SSSP distances: " + v.to_string());
            expect(drv.execute_step().status == sep::SEPExecutionStatus::CONVERGED, "SEP SSSP stays converged");
            sep::SEPPageRankApplication pr(0.85F, 1e-9F);
            sep::SEPPageRankGpuDriver pd;
            pd.initialize(pr, sep::SEPExecutionConfig(v));
            sep::SEPExecutionContext pctx(const_cast<CSRGraph *>(&cyc), nullptr, nullptr, nullptr, &frontier);
            pd.set_context(pctx);
            for (int i = 0; i < 10000 && !pd.converged(); ++i)
                expect(pd.execute_step().succeeded(), "SEP PageRank runs: " + v.to_string());
            expect(pd.converged(), "SEP PageRank converges: " + v.to_string());
            for (float x : pd.values())
                expect(std::fabs(x - 1.0F / 3.0F) <= 1e-5F, "SEP PageRank cycle ranks: " + v.to_string());
        }
        // SYNC push rounds: the chain's states after 1, 2, 3 steps are the
        // original's max_iterations = 1, 2, 3 states, and 3 rounds converge
        // (algorithm_tests.cpp:131-157)
        sep::SEPSSSPApplication app(0);
        sep::SEPSSSPGpuDriver drv;
        drv.initialize(app, sep::SEPExecutionConfig::from_string("SYNC_PUSH_DD"));
        CSRGraph chain(4, {0, 1, 2, 2, 2}, {1, 2});
        sep::SEPExecutionContext ctx(const_cast<CSRGraph *>(&chain), nullptr, nullptr, nullptr, &frontier);
        drv.set_context(ctx);
        const float inf = std::numeric_limits<float>::infinity();
        const std::vector<std::vector<float>> states = {{0.0F, 1.0F, inf, inf}, {0.0F, 1.0F, 2.0F, inf},
                                                        {0.0F, 1.0F, 2.0F, inf}};
        for (std::size_t k = 0; k < states.size(); ++k)
        {
            const auto r = drv.execute_step();
            expect(drv.values() == states[k], "SEP SYNC_PUSH_DD state after round " + std::to_string(k + 1));
            expect(r.reached_convergence() == (k == 2), "SEP SYNC_PUSH_DD converges on round 3");
        }
        expect(drv.rounds() == 3, "SEP SYNC_PUSH_DD rounds equal sssp_cpu iterations");
        // run_to_convergence: the DATA_DRIVEN PUSH variants run every round in
        // one persistent launch (SYNC with the original's iteration count),
        // the others through the stepper; the same distances either way
        CSRGraph big = [] {
            // a 3000-vertex graph with a hub and a long chain (many rounds)
            std::vector<CSRGraph::offset_type> off{0};
            std::vector<CSRGraph::vertex_id> col;
            std::vector<float> w;
            for (std::uint32_t u = 0; u < 3000; ++u)
            {
                if (u == 0)
                    for (std::uint32_t v = 1; v < 1500; ++v)
                    {
                        col.push_back(v);
                        w.push_back(static_cast<float>((v * 37) % 50 + 1));
                    }
                if (u + 1 < 3000)
                {
                    col.push_back(u + 1);
                    w.push_back(static_cast<float>((u * 13) % 7 + 1));
                }
                off.push_back(col.size());
            }
            return CSRGraph(3000, off, col, w);
        }();
        hytgraph::algorithms::SSSPOptions so;
        so.source = 0;
        // sssp_gpu: the original's distances and iteration count (pinned by
        // the golden cases)
        const auto ref = hytgraph::algorithms::sssp_gpu(big, so);
        for (const char *name : {"SYNC_PUSH_DD", "ASYNC_PUSH_DD", "SYNC_PULL_TD", "ASYNC_PUSH_TD"})
        {
            sep::SEPSSSPApplication rapp(0);
            sep::SEPSSSPGpuDriver rd;
            rd.initialize(rapp, sep::SEPExecutionConfig::from_string(name));
            sep::SEPExecutionContext rctx(const_cast<CSRGraph *>(&big), nullptr, nullptr, nullptr, &frontier);
            rd.set_context(rctx);
            const auto r = rd.run_to_convergence();
            const bool dd_push = std::string(name).find("PUSH_DD") != std::string::npos;
            expect(r.reached_convergence() && rd.converged(), std::string("run_to_convergence converges: ") + name);
            expect(rd.ran_persistent() == dd_push, std::string("persistent launch for the DD push variants: ") + name);
            expect(rd.values() == ref.distances, std::string("run_to_convergence distances: ") + name);
            if (std::string(name) == "SYNC_PUSH_DD")
                expect(rd.rounds() == ref.iterations, "persistent SYNC rounds equal the original's iterations");
            expect(rd.execute_step().status == sep::SEPExecutionStatus::CONVERGED, "converged run stays converged");
        }
        // the factory binds SSSP / PageRank applications to the GPU drivers
        auto fd = sep::SEPExecutionFactory<float, float, float>::create("SYNC_PUSH_DD");
        sep::SEPSSSPApplication fapp(0);
        fd->initialize(fapp, sep::SEPExecutionConfig::from_string("SYNC_PUSH_DD"));
        sep::SEPExecutionContext fctx(const_cast<CSRGraph *>(&g), nullptr, nullptr, nullptr, &frontier);
        fd->set_context(fctx);
        int fsteps = 0;
        while (!fd->converged() && fsteps < 100)
        {
            expect(fd->execute_step().succeeded(), "factory SSSP driver steps");
            ++fsteps;
        }
        const auto *fv = sep::gpu_values(*fd);
        expect(fd->converged() && fv != nullptr && *fv == std::vector<float>({0.0F, 1.0F, 3.0F, 4.0F}),
               "factory driver runs SSSP on the GPU");
        auto fp = sep::SEPExecutionFactory<float, float, double>::create("ASYNC_PUSH_DD");
        sep::SEPPageRankApplication fpr(0.85F, 1e-9F);
        fp->initialize(fpr, sep::SEPExecutionConfig::from_string("ASYNC_PUSH_DD"));
        sep::SEPExecutionContext fpctx(const_cast<CSRGraph *>(&cyc), nullptr, nullptr, nullptr, &frontier);
        fp->set_context(fpctx);
        for (int i = 0; i < 10000 && !fp->converged(); ++i)
            expect(fp->execute_step().succeeded(), "factory PageRank driver steps");
        const auto *pv = sep::gpu_values(*fp);
        expect(fp->converged() && pv != nullptr && std::fabs((*pv)[0] - 1.0F / 3.0F) <= 1e-5F,
               "factory driver runs PageRank on the GPU");
    }

    // device hub sort / reorder / contributions against the host forms
    void test_device_preprocess_gpu()
    {
        namespace sch = hytgraph::scheduling;
        // the original's unit_tests.cpp:2498-2537, 2710-2750 graph, then a
        // larger pseudo-random graph with parallel edges and self-loops
        std::vector<CSRGraph> graphs;
        graphs.emplace_back(5, std::vector<CSRGraph::offset_type>{0, 3, 4, 5, 6, 7},
                            std::vector<CSRGraph::vertex_id>{1, 2, 3, 0, 0, 0, 0});
        {
            const std::size_t n = 3000;
            std::vector<CSRGraph::offset_type> off(n + 1, 0);
            std::vector<CSRGraph::vertex_id> col;
            std::vector<float> w;
            std::uint64_t x = 88172645463325252ULL;
            for (std::size_t u = 0; u < n; ++u)
            {
                x ^= x << 13, x ^= x >> 7, x ^= x << 17;
                const std::size_t deg = (u % 7 == 0) ? 0 : (x % 40) * ((u % 13 == 0) ? 8 : 1);
                for (std::size_t k = 0; k < deg; ++k)
                {
                    x ^= x << 13, x ^= x >> 7, x ^= x << 17;
                    col.push_back(static_cast<CSRGraph::vertex_id>((x % 5 == 0) ? x % 50 : x % n));
                    w.push_back(static_cast<float>(x % 1000) * 0.125F);
                }
                off[u + 1] = col.size();
            }
            graphs.emplace_back(n, off, col, w);
        }
        for (const CSRGraph &g : graphs)
        {
            for (double frac : {0.0, 0.08, 0.4, 1.0})
            {
                sch::HubSortOptions o;
                o.hub_fraction = frac;
                const auto a = sch::HubSorter(o).sort(g);
                const auto b = sch::HubSorter(o).sort_gpu(g);
                expect(a.vertex_order == b.vertex_order && a.hub_count == b.hub_count,
                       "device hub order equals the host order");
                expect(std::memcmp(a.scores.data(), b.scores.data(), a.scores.size() * sizeof(double)) == 0,
                       "device hub scores equal the host scores bit for bit");
            }
            const auto order = sch::HubSorter().sort(g).vertex_order;
            CSRGraph h = g, d = g;
            h.reorder_vertices(order);
            hytgraph::graph::reorder_vertices_gpu(d, order);
            expect(h.row_offsets() == d.row_offsets() && h.column_indices() == d.column_indices() &&
                       h.edge_weights() == d.edge_weights(),
                   "device reorder equals CSRGraph::reorder_vertices");
            std::vector<CSRGraph::vertex_id> dup(order);
            dup.back() = dup.front();
            CSRGraph e = g;
            expect_throw<std::invalid_argument>([&] { hytgraph::graph::reorder_vertices_gpu(e, dup); },
                                                "device reorder rejects a duplicate id");
            expect(e.column_indices() == g.column_indices(), "a rejected reorder leaves the graph");
            std::vector<float> x(static_cast<std::size_t>(g.num_vertices()));
            for (std::size_t v = 0; v < x.size(); ++v)
                x[v] = 1.0F / static_cast<float>(x.size()) * static_cast<float>(1 + v % 5);
            const auto pc = alg::pagerank_contributions(g, x);
            const auto pg = alg::pagerank_contributions_gpu(g, x);
            bool same = pc.size() == pg.size();
            for (std::size_t v = 0; same && v < pc.size(); ++v)
                same = pc[v].vertex == pg[v].vertex &&
                       std::memcmp(&pc[v].contribution, &pg[v].contribution, sizeof(double)) == 0;
            expect(same, "device PageRank contributions equal the host ones bit for bit");
        }
        const float inf = std::numeric_limits<float>::infinity();
        const std::vector<float> cur{5.0F, 5.0F, inf, inf, 3.0F, 2.0F}, cand{3.0F, 4.75F, 7.0F, inf, inf, 2.0F};
        const auto sc = alg::sssp_contributions(cur, cand, 0.25F);
        const auto sg = alg::sssp_contributions_gpu(cur, cand, 0.25F);
        bool same = sc.size() == sg.size();
        for (std::size_t v = 0; same && v < sc.size(); ++v)
            same = sc[v].contribution == sg[v].contribution;
        expect(same, "device SSSP contributions equal the host ones");
        expect_throw<std::invalid_argument>([&] { (void)alg::sssp_contributions_gpu(cur, cand, -1.0F); },
                                            "device SSSP contributions reject a negative tolerance");
    }
} // namespace

int main(int argc, char **argv)
{
    const std::string what = argc > 1 ? argv[1] : "cpu";
    try
    {
        if (what == "cpu")
        {
            test_csr_validation();
            test_reorder();
            test_contributions();
            test_sep_host();
        }
        else
        {
            int n = 0;
            if (hyt_device_count(&n) != HYT_OK || n == 0)
            {
                std::cout << "GPU API tests skipped: no HIP device available.\n";
                return 77;
            }
            test_sssp_gpu_known_answers();
            test_pagerank_gpu_known_answers();
            test_sep_gpu();
            test_device_preprocess_gpu();
        }
        std::cout << "C++ API tests (" << what << ") passed.\n";
        return 0;
    }
    catch (const std::exception &e)
    {
        std::cerr << e.what() << '\n';
        return 1;
    }
}
