// Host planning API in the original project's test style: the known answers
// of its tests/unit_tests.cpp for ExpTM-Filter (918-1121), ExpTM-Compaction
// (1289-1563), ImpTM zero-copy (1737-2056) and the contribution scheduler
// (2751-2887) and the text edge-list loader (373-592), plus seeded random
// graphs checked against direct per-vertex recomputation (partition walking,
// per-partition totals, coverage errors) and a binary CSR round trip.
#include "graph/activity_tracker.hpp"
#include "graph/csr_graph.hpp"
#include "graph/graph_loader.hpp"
#include "graph/partition.hpp"
#include "runtime/config.hpp"
#include "runtime/logging.hpp"
#include "runtime/result.hpp"
#include "scheduling/contribution_scheduler.hpp"
#include "scheduling/hub_sort.hpp"
#include "transfer/compaction_engine.hpp"
#include "transfer/filter_engine.hpp"
#include "transfer/zero_copy_engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <iterator>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
    using hytgraph::graph::ActivityTracker;
    using hytgraph::graph::CSRGraph;
    using hytgraph::graph::LogicalPartition;
    using hytgraph::graph::LogicalPartitioner;
    namespace tr = hytgraph::transfer;
    namespace sc = hytgraph::scheduling;
    using Ids = std::vector<CSRGraph::vertex_id>;
    using Offs = std::vector<CSRGraph::offset_type>;

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

    void test_filter_known_answers()
    {
        CSRGraph g(4, {0, 2, 3, 4, 5}, {1, 2, 2, 0, 0});
        const auto parts = LogicalPartitioner(2U * sizeof(CSRGraph::vertex_id)).partition(g);
        ActivityTracker a(g.num_vertices());
        a.set_active(0);
        const auto plan = tr::ExpTMFilter(true).plan(g, parts, a);
        expect(plan.partitions.size() == parts.size(), "filter: every partition has a decision");
        expect(plan.active_partition_count() >= 1U, "filter: active partition");
        expect(plan.transferred_partition_count() == plan.active_partition_count(), "filter: moves active ones");
        expect(plan.transferred_edge_count() > 0U && plan.transferred_edge_count() < g.num_edges(),
               "filter: skips inactive partitions");
        std::size_t bytes = 0;
        CSRGraph::offset_type edges = 0;
        for (std::size_t i = 0; i < parts.size(); ++i)
            if (plan.partitions[i].transferred)
            {
                edges += parts[i].edge_count();
                bytes += parts[i].edge_data_bytes();
                expect(plan.partitions[i].transferred_edges == parts[i].edge_count(), "filter: whole partition");
            }
        expect(plan.transferred_edge_count() == edges && plan.transferred_byte_count() == bytes, "filter totals");
        const auto all = tr::ExpTMFilter(false).plan(g, parts, a);
        expect(all.transferred_partition_count() == parts.size() && all.transferred_edge_count() == g.num_edges(),
               "disabled filter moves everything");
        ActivityTracker wrong(3);
        expect_throw<std::invalid_argument>([&] { (void)tr::ExpTMFilter().plan(g, parts, wrong); },
                                            "filter: activity size checked");
        std::vector<LogicalPartition> gap(parts.begin() + 1, parts.end());
        expect_throw<std::invalid_argument>([&] { (void)tr::ExpTMFilter().plan(g, gap, a); },
                                            "filter: coverage checked");
    }

    void test_compaction_known_answers()
    {
        CSRGraph g(4, {0, 2, 3, 5, 6}, {1, 2, 2, 0, 3, 1});
        const auto parts = LogicalPartitioner(16U * sizeof(CSRGraph::vertex_id)).partition(g);
        expect(parts.size() == 1U, "compaction: one partition");
        ActivityTracker a(g.num_vertices());
        a.set_active(0);
        a.set_active(2);
        const auto r = tr::ExpTMCompaction(true).compact(g, parts, a);
        expect(r.compacted_partition_count() == 1U && r.active_vertex_count() == 2U && r.active_edge_count() == 4U,
               "compaction counts");
        const auto &p = r.partitions.front();
        expect(p.active_vertices == Ids({0, 2}), "compaction: active vertices ascending");
        expect(p.neighbors == Ids({1, 2, 0, 3}), "compaction: neighbours in source order");
        expect(p.neighbor_index == Offs({0, 2, 4}), "compaction: prefix index");
        expect(tr::ExpTMCompaction::expanded_neighbors(p) == Ids({1, 2, 0, 3}), "compaction: expansion");
        expect(r.total_bytes() == 4 * 4 + 3 * 8, "compaction: payload bytes");
        // zero-degree active vertex
        CSRGraph z(3, {0, 1, 1, 2}, {1, 0});
        ActivityTracker az(3);
        az.set_active(1);
        const auto rz = tr::ExpTMCompaction(true).compact(z, LogicalPartitioner(64).partition(z), az);
        expect(rz.compacted_partition_count() == 1U && rz.partitions[0].active_vertices == Ids({1}) &&
                   rz.partitions[0].neighbors.empty() && rz.partitions[0].neighbor_index == Offs({0, 0}),
               "compaction: zero-degree vertex keeps an empty range");
        // all active = the original CSR
        ActivityTracker aa(4);
        for (CSRGraph::vertex_id v = 0; v < 4; ++v)
            aa.set_active(v);
        const auto ra = tr::ExpTMCompaction(true).compact(g, parts, aa);
        expect(ra.partitions[0].neighbors == g.column_indices() && ra.partitions[0].neighbor_index == g.row_offsets(),
               "compaction: all active reproduces the CSR");
        expect(tr::ExpTMCompaction(false).compact(g, parts, aa).partitions.empty(), "disabled compaction is empty");
        tr::CompactedPartition bad = p;
        bad.neighbor_index = {0, 3, 2};
        expect_throw<std::invalid_argument>([&] { (void)tr::ExpTMCompaction::expanded_neighbors(bad); },
                                            "expansion: monotone index");
        bad.neighbor_index = {0, 2};
        expect_throw<std::invalid_argument>([&] { (void)tr::ExpTMCompaction::expanded_neighbors(bad); },
                                            "expansion: index size");
    }

    void test_zero_copy_known_answers()
    {
        CSRGraph g(4, {0, 2, 3, 4, 4}, {1, 2, 2, 0});
        const auto parts = LogicalPartitioner(2U * sizeof(CSRGraph::vertex_id)).partition(g);
        ActivityTracker a(4);
        a.set_active_vertices({0, 2});
        const auto r = tr::ImpTMZeroCopy().prepare(g, parts, a);
        expect(r.mode == tr::ZeroCopyMode::Modeled && !r.host_memory_mapped, "zero-copy: modeled");
        expect(r.active_vertex_count == 2 && r.active_edge_count == 3, "zero-copy: active counts");
        expect(r.memory_request_count == 2 && r.alignment_overhead_count == 1 && r.modeled_tlp_count == 1,
               "zero-copy: 2 requests, 1 alignment, 1 TLP");
        std::size_t seen = 0;
        for (const auto &p : r.partitions)
        {
            expect(p.active_vertices.size() == p.vertex_metrics.size() && p.modeled_tlp_count == 0,
                   "zero-copy: per-partition metrics");
            seen += p.active_vertices.size();
        }
        expect(seen == 2, "zero-copy: each active vertex once");
        CSRGraph wide(2, {0, 40, 40}, Ids(40, 1));
        ActivityTracker aw(2);
        aw.set_active(0);
        tr::ZeroCopyOptions o;
        o.request_payload_bytes = 16U;
        o.max_requests_per_tlp = 2U;
        const auto rw = tr::ImpTMZeroCopy(o).prepare(wide, LogicalPartitioner(1024U).partition(wide), aw);
        expect(rw.memory_request_count == 10 && rw.modeled_tlp_count == 5, "zero-copy: m=16 -> 10 requests, 5 TLPs");
        CSRGraph zd(3, {0, 1, 1, 1}, {1});
        ActivityTracker az(3);
        az.set_active(1);
        const auto rz = tr::ImpTMZeroCopy().prepare(zd, LogicalPartitioner(1024U).partition(zd), az);
        expect(rz.active_vertex_count == 1 && rz.active_edge_count == 0 && rz.memory_request_count == 0 &&
                   rz.alignment_overhead_count == 0 && rz.modeled_tlp_count == 0,
               "zero-copy: zero-degree vertex costs nothing");
        tr::ZeroCopyOptions zero;
        zero.request_payload_bytes = 0;
        expect_throw<std::invalid_argument>([&] { (void)tr::ImpTMZeroCopy(zero).prepare(g, parts, a); },
                                            "zero-copy: options checked");
    }

    void test_scheduler_known_answers()
    {
        using P = sc::ContributionPriority;
        sc::ContributionScheduler s;
        auto plan = s.schedule({{0, 0.10}, {1, 0.90}, {2, 0.30}, {3, 0.70}});
        expect(plan.ordered_item_indices == std::vector<std::size_t>({1, 3, 2, 0}), "scheduler: descending");
        expect(plan.size() == 4 && plan.metrics.input_item_count == 4 && plan.metrics.scheduled_item_count == 4,
               "scheduler: counts");
        expect(std::fabs(plan.metrics.scheduling_overhead - 2.0) < 1e-12, "scheduler: log2 overhead");
        plan = s.schedule({{3, 0.5}, {1, 0.5}, {2, 0.9}, {0, 0.5}});
        expect(plan.ordered_item_indices == std::vector<std::size_t>({2, 0, 1, 3}), "scheduler: ties by index");
        sc::ContributionSchedulerOptions eo;
        eo.zero_contribution_epsilon = 1.0e-6;
        plan = sc::ContributionScheduler(eo).schedule({{0, 0.0}, {1, 1e-8}, {2, 0.25}, {3, -1e-8}});
        expect(plan.metrics.zero_contribution_count == 3 &&
                   plan.ordered_item_indices == std::vector<std::size_t>({2, 1, 0, 3}),
               "scheduler: zero epsilon");
        plan = s.schedule({{0, 0.10}, {1, 0.90}, {2, 0.30}});
        expect(plan.metrics.reordered_item_count == 3 && plan.reordered(), "scheduler: reordered count");
        plan = sc::ContributionScheduler::synchronous_reference({{5, 0.1}, {2, 0.9}, {8, 0.3}});
        expect(plan.ordered_item_indices == std::vector<std::size_t>({5, 2, 8}), "scheduler: synchronous order");
        plan = s.schedule({{4, 1.0}, {4, 2.0}, {1, 0.0}});
        expect(plan.metrics.redundant_work_count == 1, "scheduler: duplicate items");
        expect_throw<std::invalid_argument>([&] { (void)s.schedule({P{0, NAN}}); }, "scheduler: non-finite");
        expect_throw<std::invalid_argument>([] { (void)sc::make_contribution_priorities({1, 2}, {0.5}); },
                                            "priorities: sizes");
        sc::ContributionSchedulerOptions nd;
        nd.deterministic_tie_break = false;
        plan = sc::ContributionScheduler(nd).schedule({{3, 0.5}, {1, 0.5}, {0, 0.5}});
        expect(plan.ordered_item_indices == std::vector<std::size_t>({3, 1, 0}), "scheduler: stable without tie break");
        const auto m = sc::ContributionScheduler::measure_work_metrics({{1, 0.0}, {1, 1.0}}, {{1, true}, {2, false}});
        expect(m.zero_contribution_count == 1 && m.redundant_work_count == 1 && m.stale_work_count == 1,
               "scheduler: work metrics");
    }

    // Seeded random graphs: every planner against a direct recomputation.
    void test_random_graphs()
    {
        std::mt19937_64 rng(12345);
        for (int trial = 0; trial < 40; ++trial)
        {
            const std::uint32_t V = 1 + static_cast<std::uint32_t>(rng() % 300);
            Offs off(V + 1, 0);
            Ids col;
            for (std::uint32_t v = 0; v < V; ++v)
            {
                const std::uint32_t d = (rng() % 7 == 0) ? static_cast<std::uint32_t>(rng() % 90)
                                                         : static_cast<std::uint32_t>(rng() % 6);
                for (std::uint32_t k = 0; k < d; ++k)
                    col.push_back(static_cast<CSRGraph::vertex_id>(rng() % V));
                off[v + 1] = col.size();
            }
            CSRGraph g(V, off, col);
            const auto parts = LogicalPartitioner(4U * (1 + rng() % 64)).partition(g);
            ActivityTracker a(V);
            for (std::uint32_t v = 0; v < V; ++v)
                if (rng() % 3 == 0)
                    a.set_active(v);
            tr::ZeroCopyOptions zo;
            zo.request_payload_bytes = 4U * (1 + rng() % 40);
            zo.max_requests_per_tlp = 1 + rng() % 8;
            zo.alignment_bytes = 4U * (1 + rng() % 40);
            const auto fp = tr::ExpTMFilter().plan(g, parts, a);
            const auto cp = tr::ExpTMCompaction().compact(g, parts, a);
            const auto zp = tr::ImpTMZeroCopy(zo).prepare(g, parts, a);
            std::size_t ci = 0;
            std::uint64_t all_req = 0;
            for (std::size_t i = 0; i < parts.size(); ++i)
            {
                const auto &p = parts[i];
                Ids act, nb;
                Offs idx{0};
                std::uint64_t aedges = 0, req = 0, al = 0;
                for (auto v = p.vertex_begin(); v < p.vertex_end(); ++v)
                {
                    if (!a.is_active(v))
                        continue;
                    act.push_back(v);
                    const std::uint64_t d = off[v + 1] - off[v];
                    for (auto e = off[v]; e < off[v + 1]; ++e)
                        nb.push_back(col[e]);
                    idx.push_back(nb.size());
                    aedges += d;
                    req += (d * 4 + zo.request_payload_bytes - 1) / zo.request_payload_bytes;
                    al += (d != 0 && (off[v] * 4) % zo.alignment_bytes != 0) ? 1 : 0;
                }
                expect(fp.partitions[i].active == (aedges != 0) && fp.partitions[i].active_edges == aedges,
                       "random: filter activity");
                if (!act.empty())
                {
                    expect(ci < cp.partitions.size() && cp.partitions[ci].partition_index == i, "random: compacted id");
                    expect(cp.partitions[ci].active_vertices == act && cp.partitions[ci].neighbors == nb &&
                               cp.partitions[ci].neighbor_index == idx,
                           "random: compacted payload");
                    ++ci;
                }
                expect(zp.partitions[i].active_vertices == act && zp.partitions[i].active_edge_count == aedges &&
                           zp.partitions[i].memory_request_count == req &&
                           zp.partitions[i].alignment_overhead_count == al,
                       "random: zero-copy partition metrics");
                all_req += req;
            }
            expect(ci == cp.partitions.size(), "random: no extra compacted partitions");
            expect(zp.modeled_tlp_count == (all_req + zo.max_requests_per_tlp - 1) / zo.max_requests_per_tlp,
                   "random: modeled TLPs");
        }
    }
    // Writes `text` to a temporary file and returns its path.
    std::string temp_file(const std::string &name, const std::string &text)
    {
        const auto path = std::filesystem::temp_directory_path() / ("hytgraph_amd_" + name);
        std::ofstream(path) << text;
        return path.string();
    }

    void test_graph_loader_known_answers()
    {
        using hytgraph::graph::GraphLoader;
        {
            const std::string p = temp_file("u.el", "# simple directed graph\n0 1\n0 2\n1 2\n2 0\n");
            const CSRGraph g = GraphLoader::load_edge_list(p, 4, false);
            expect(g.num_vertices() == 4 && g.num_edges() == 4 && !g.has_weights(), "loader: unweighted sizes");
            expect(g.out_degree(0) == 2 && g.out_degree(1) == 1 && g.out_degree(2) == 1 && g.out_degree(3) == 0,
                   "loader: degrees");
            expect(g.neighbor_at(0) == 1 && g.neighbor_at(1) == 2 && g.neighbor_at(2) == 2 && g.neighbor_at(3) == 0,
                   "loader: neighbours");
            std::filesystem::remove(p);
        }
        {
            const std::string p = temp_file("w.el", "# weighted graph\n0 1 0.5\n0 2 1.25\n1 2 2.5\n");
            const CSRGraph g = GraphLoader::load_edge_list(p, 3, true);
            expect(g.has_weights() && g.num_edges() == 3, "loader: weighted sizes");
            expect(g.weight_at(0) == 0.5F && g.weight_at(1) == 1.25F && g.weight_at(2) == 2.5F, "loader: weights");
            // binary round trip keeps every array and the float weights
            const std::string b = temp_file("w.hytcsr", "");
            GraphLoader::save_binary(g, b);
            const CSRGraph h = GraphLoader::load_binary(b);
            expect(h.row_offsets() == g.row_offsets() && h.column_indices() == g.column_indices() &&
                       h.edge_weights() == g.edge_weights(),
                   "binary CSR round trip");
            std::filesystem::remove(p);
            std::filesystem::remove(b);
        }
        {
            const std::string p = temp_file("par.el", "0 1\n0 1\n0 2\n");
            const CSRGraph g = GraphLoader::load_edge_list(p, 3, false);
            expect(g.num_edges() == 3 && g.out_degree(0) == 3, "loader: parallel edges kept");
            expect(g.neighbor_at(0) == 1 && g.neighbor_at(1) == 1 && g.neighbor_at(2) == 2, "loader: edge order");
            std::filesystem::remove(p);
        }
        const std::string bad_id = temp_file("bad_id.el", "0 3\n");
        expect_throw<std::invalid_argument>([&] { (void)GraphLoader::load_edge_list(bad_id, 3, false); },
                                            "loader: destination out of range");
        const std::string no_w = temp_file("no_w.el", "0 1\n");
        expect_throw<std::invalid_argument>([&] { (void)GraphLoader::load_edge_list(no_w, 2, true); },
                                            "loader: missing weight");
        const std::string junk = temp_file("junk.el", "this is not an edge\n");
        expect_throw<std::invalid_argument>([&] { (void)GraphLoader::load_edge_list(junk, 2, false); },
                                            "loader: malformed record");
        expect_throw<std::runtime_error>(
            [&] { (void)GraphLoader::load_edge_list("/nonexistent/hytgraph_amd.el", 2, false); },
            "loader: unreadable file");
        try
        {
            (void)GraphLoader::load_edge_list(junk, 2, false);
        }
        catch (const std::invalid_argument &e)
        {
            expect(std::string(e.what()) == "GraphLoader::load_edge_list: malformed edge at line 1",
                   std::string("loader: message: ") + e.what());
        }
        for (const auto &p : {bad_id, no_w, junk})
            std::filesystem::remove(p);
    }
    // the original's hub-sort known answers (tests/unit_tests.cpp:2468-2749),
    // and on seeded random graphs the ranked prefix equals a full ranking of
    // every vertex (recomputed here from the degrees)
    void test_hub_sort_known_answers()
    {
        using sc::HubSorter;
        using sc::HubSortOptions;
        {
            const CSRGraph g(3, Offs{0, 2, 3, 4}, Ids{1, 2, 2, 0});
            const auto r = HubSorter(HubSortOptions{2.0 / 3.0}).sort(g);
            expect(r.scores.size() == 3 && std::abs(r.scores[0] - 0.5) < 1e-12 && std::abs(r.scores[1] - 0.25) < 1e-12 &&
                       std::abs(r.scores[2] - 0.5) < 1e-12,
                   "hub scores 0.5 / 0.25 / 0.5");
        }
        {
            CSRGraph g(5, Offs{0, 3, 4, 5, 6, 7}, Ids{1, 2, 3, 0, 0, 0, 0});
            const auto r = HubSorter(HubSortOptions{0.4}).sort(g);
            expect(r.hub_count == 2 && r.vertex_order == Ids({0, 1, 2, 3, 4}), "hubs first, others in order");
            g.reorder_vertices(r.vertex_order);
            expect(g.row_offsets() == Offs({0, 3, 4, 5, 6, 7}) && g.column_indices() == Ids({1, 2, 3, 0, 0, 0, 0}),
                   "hub order applied to the CSR");
        }
        {
            const CSRGraph g(6, Offs{0, 2, 3, 4, 5, 6, 6}, Ids{1, 2, 2, 3, 0, 0});
            const auto r = HubSorter().sort(g);
            std::vector<bool> seen(6, false);
            for (const auto v : r.vertex_order)
            {
                expect(v < 6 && !seen[v], "hub order is a permutation");
                seen[v] = true;
            }
            expect(r.vertex_order.size() == 6, "hub order covers every vertex");
        }
        {
            const CSRGraph g(4, Offs{0, 0, 0, 0, 0}, Ids{});
            const auto r = HubSorter().sort(g);
            expect(r.hub_count == 1 && r.vertex_order == Ids({0, 1, 2, 3}), "edgeless graph: one hub, natural order");
            for (const double x : r.scores)
                expect(x == 0.0, "edgeless graph: zero scores");
        }
        expect_throw<std::invalid_argument>([] { HubSorter(HubSortOptions{1.5}); }, "hub fraction above 1");
        expect_throw<std::invalid_argument>([] { HubSorter(HubSortOptions{std::nan("")}); }, "hub fraction NaN");
        std::mt19937_64 rng(11);
        for (int t = 0; t < 20; ++t)
        {
            const std::uint32_t V = 50 + static_cast<std::uint32_t>(rng() % 400);
            Offs off(V + 1, 0);
            Ids col;
            for (std::uint32_t v = 0; v < V; ++v)
            {
                const std::uint32_t d = static_cast<std::uint32_t>(rng() % 9 == 0 ? rng() % 40 : rng() % 4);
                for (std::uint32_t k = 0; k < d; ++k)
                    col.push_back(static_cast<std::uint32_t>(rng() % 7 == 0 ? rng() % 5 : rng() % V));
                off[v + 1] = col.size();
            }
            const CSRGraph g(V, off, col);
            const double frac = static_cast<double>(rng() % 101) / 100.0;
            const auto r = HubSorter(HubSortOptions{frac}).sort(g);
            std::vector<std::uint64_t> in(V, 0);
            for (auto c : col)
                ++in[c];
            std::vector<std::uint32_t> all(V);
            for (std::uint32_t v = 0; v < V; ++v)
                all[v] = v;
            const auto key = [&](std::uint32_t v) {
                return static_cast<long double>(off[v + 1] - off[v]) * static_cast<long double>(in[v]);
            };
            std::stable_sort(all.begin(), all.end(), [&](std::uint32_t a, std::uint32_t b) { return key(a) > key(b); });
            const double want = V * frac;
            const std::size_t hubs = std::min<std::size_t>(
                V, static_cast<std::size_t>(std::fabs(want - std::round(want)) < 1e-12 ? std::round(want) : std::ceil(want)));
            expect(r.hub_count == hubs, "random: hub count");
            for (std::size_t i = 0; i < hubs; ++i)
                expect(r.vertex_order[i] == all[i], "random: hub prefix equals the full ranking");
            std::size_t at = hubs;
            std::vector<bool> hub(V, false);
            for (std::size_t i = 0; i < hubs; ++i)
                hub[all[i]] = true;
            for (std::uint32_t v = 0; v < V; ++v)
                if (!hub[v])
                    expect(r.vertex_order[at++] == v, "random: others in natural order");
        }
    }

    // the original's runtime known answers (tests/unit_tests.cpp:154-204)
    void test_runtime_known_answers()
    {
        namespace rt = hytgraph::runtime;
        rt::Config config;
        config.set("algorithm", "pagerank");
        config.set("partition.size_bytes", "33554432");
        config.set("cache.enabled", "true");
        config.set("transfer.gamma", "0.625");
        expect(config.contains("algorithm") && config.get("algorithm") == "pagerank", "config: string lookup");
        expect(config.get_int("partition.size_bytes") == 33554432, "config: integer");
        expect(config.get_bool("cache.enabled"), "config: boolean");
        expect(config.get_double("transfer.gamma") == 0.625, "config: double");
        expect(config.get_or("missing", "fallback") == "fallback", "config: fallback");
        expect_throw<std::out_of_range>([&] { (void)config.get("missing"); }, "config: missing key");
        expect_throw<std::invalid_argument>([&] { config.set("", "x"); }, "config: empty key");
        config.set("flag", "Maybe");
        expect_throw<std::invalid_argument>([&] { (void)config.get_bool("flag"); }, "config: bad boolean");
        config.set("flag", "OFF");
        expect(!config.get_bool("flag"), "config: OFF");

        rt::Result result;
        result.experiment_id = "unit-test";
        result.timestamp_utc = "2026-01-01T00:00:00Z";
        result.config["algorithm"] = "pagerank";
        result.measurements["runtime_seconds"] = 1.25;
        result.measurements["gteps"] = 0.1;
        const std::string json = rt::result_to_json(result);
        for (const char *want : {"\"schema_version\": \"0.1\"", "\"experiment_id\": \"unit-test\"",
                                 "\"runtime_seconds\": 1.25", "\"gteps\": 0.1", "\"status\": \"success\"",
                                 "\"algorithm\": \"pagerank\""})
            expect(json.find(want) != std::string::npos, std::string("result json: ") + want);
        const auto path = std::filesystem::temp_directory_path() / "hytgraph_amd_result.json";
        expect(rt::write_result_json(result, path.string()), "result file writes");
        std::ifstream file(path);
        std::string back((std::istreambuf_iterator<char>(file)), std::istreambuf_iterator<char>());
        expect(back == json, "result file content");
        std::filesystem::remove(path);
        expect(!rt::write_result_json(result, "/nonexistent/dir/r.json"), "result file: unwritable path");
        rt::set_log_level(rt::LogLevel::Error);
        rt::log(rt::LogLevel::Info, "filtered");
        rt::set_log_level(rt::LogLevel::Info);
    }
} // namespace

int main()
{
    try
    {
        test_filter_known_answers();
        test_compaction_known_answers();
        test_zero_copy_known_answers();
        test_scheduler_known_answers();
        test_random_graphs();
        test_graph_loader_known_answers();
        test_runtime_known_answers();
        test_hub_sort_known_answers();
        std::cout << "planning API tests passed.\n";
        return 0;
    }
    catch (const std::exception &e)
    {
        std::cerr << e.what() << '\n';
        return 1;
    }
}
