// Synthetic graph generator: RMAT (Chakrabarti, Zhan, Faloutsos 2004) and
// Erdos-Renyi G(V, E), deterministic for any thread count.
//
// Edge i is a pure function of (seed, i): for RMAT each of the `scale` bit
// levels draws one 32-bit uniform and picks a quadrant with probabilities
// a, b, c, d (top-left = (0,0)). Self-loops are skipped, duplicates removed,
// rows sorted by destination. Weight of the edge at final CSR position p is
// a pure function of (seed, p). Vertex labels are not scrambled.
//
// Layout of the work (the peak is the CSR plus a u32 per vertex: no edge
// list is held beside the CSR, so an RMAT scale-30 graph, 17 B generated
// edges, fits in about 150 GB):
//   1. count each source's generated edges (atomic u32 per vertex);
//   2. prefix sums give every row a range of the column array, sized for its
//      generated edges; the edges are generated again and each scattered into
//      its row (a per-vertex cursor);
//   3. each row is sorted and deduplicated in place, in parallel;
//   4. the rows are packed to the left in ascending order (each row only
//      moves to lower positions), the offsets rewritten;
//   5. weights are drawn for the final positions into their own array.
// Generating each edge twice costs less than storing it twice.
#include "common.hpp"
#include "host_graph.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace hyt
{
    namespace
    {
        int resolve_threads(int requested)
        {
            if (requested > 0)
                return requested;
            unsigned hw = std::thread::hardware_concurrency();
            if (hw == 0)
                hw = 1;
            return static_cast<int>(std::min(hw, 16U));
        }

        // Progress lines on stderr for long generations (HYT_GEN_PROGRESS set):
        // thread 0 reports each sixteenth of its share of a pass.
        struct Progress
        {
            bool on = false;
            std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
            void say(const char *what, double frac) const
            {
                if (!on)
                    return;
                const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                std::fprintf(stderr, "[generator] %s %3.0f%% at %.1f s\n", what, 100.0 * frac, s);
                std::fflush(stderr);
            }
        };

        template <class Fn>
        void parallel_for(int threads, std::uint64_t n, Fn &&fn)
        {
            // Static contiguous split: thread t gets [t*n/T, (t+1)*n/T).
            std::vector<std::thread> pool;
            for (int t = 0; t < threads; ++t)
            {
                std::uint64_t lo = n * static_cast<std::uint64_t>(t) / threads;
                std::uint64_t hi = n * static_cast<std::uint64_t>(t + 1) / threads;
                pool.emplace_back([&fn, t, lo, hi] { fn(t, lo, hi); });
            }
            for (auto &th : pool)
                th.join();
        }

        template <class Fn>
        void parallel_dynamic(int threads, std::uint64_t n, Fn &&fn)
        {
            std::atomic<std::uint64_t> next{0};
            std::vector<std::thread> pool;
            for (int t = 0; t < threads; ++t)
                pool.emplace_back([&] {
                    for (;;)
                    {
                        std::uint64_t i = next.fetch_add(1);
                        if (i >= n)
                            break;
                        fn(i);
                    }
                });
            for (auto &th : pool)
                th.join();
        }

        struct EdgeSource
        {
            const hyt_gen_params &p;
            std::uint32_t thr_a, thr_ab, thr_abc;
            std::uint64_t key; // the edge stream's key: draw(seed, edges, i, w) = mix(key + i c1 + w c2)

            explicit EdgeSource(const hyt_gen_params &params)
                : p(params), key(mix64(params.seed ^ mix64(kStreamEdges + 0x632BE59BD9B4E019ULL)))
            {
                auto t = [](double x) {
                    double v = std::floor(x * 4294967296.0);
                    if (v < 0)
                        v = 0;
                    if (v > 4294967295.0)
                        v = 4294967295.0;
                    return static_cast<std::uint32_t>(v);
                };
                thr_a = t(p.a);
                thr_ab = t(p.a + p.b);
                thr_abc = t(p.a + p.b + p.c);
            }

            // Returns false for a self-loop. The words are those of
            // draw(seed, kStreamEdges, i, w), with the stream key hoisted.
            bool edge(std::uint64_t i, std::uint32_t &src, std::uint32_t &dst) const
            {
                const std::uint64_t base = key + i * 0xD1B54A32D192ED03ULL;
                if (p.kind == HYT_GEN_ER)
                {
                    std::uint64_t r = mix64(base);
                    std::uint64_t V = 1ULL << p.scale;
                    src = static_cast<std::uint32_t>(bounded32(static_cast<std::uint32_t>(r), V));
                    dst = static_cast<std::uint32_t>(bounded32(static_cast<std::uint32_t>(r >> 32), V));
                }
                else
                {
                    // quadrant per level without branches: u < a -> (0,0),
                    // a <= u < a+b -> (0,1), a+b <= u < a+b+c -> (1,0), else (1,1)
                    std::uint32_t s = 0, d = 0;
                    auto level = [&](std::uint32_t u, std::uint32_t bit) {
                        const std::uint32_t ga = u >= thr_a, gab = u >= thr_ab, gabc = u >= thr_abc;
                        s |= gab << bit;
                        d |= (ga ^ gab ^ gabc) << bit;
                    };
                    std::uint32_t bit = p.scale - 1U;
                    for (std::uint32_t w = 0; 2 * w < p.scale; ++w)
                    {
                        const std::uint64_t r = mix64(base + w * 0x8CB92BA72F3D8DD7ULL);
                        level(static_cast<std::uint32_t>(r), bit);
                        if (2 * w + 1 < p.scale)
                            level(static_cast<std::uint32_t>(r >> 32), bit - 1U);
                        bit -= 2U;
                    }
                    src = s;
                    dst = d;
                }
                return src != dst;
            }
        };

        void validate_params(const hyt_gen_params &p)
        {
            require(p.kind == HYT_GEN_RMAT || p.kind == HYT_GEN_ER, "unknown generator kind");
            require(p.scale >= 1 && p.scale <= 32, "scale must be in [1, 32]");
            require(p.edge_factor >= 1 && p.edge_factor <= 1024, "edge_factor must be in [1, 1024]");
            if (p.kind == HYT_GEN_RMAT)
            {
                require(p.a >= 0 && p.b >= 0 && p.c >= 0 && p.a + p.b + p.c <= 1.0 + 1e-12,
                        "RMAT probabilities must be non-negative and a+b+c <= 1");
            }
            require(p.weight_min <= p.weight_max, "weight_min must not exceed weight_max");
            require(p.memory == HYT_HOST_PAGEABLE || p.memory == HYT_HOST_PINNED, "unknown memory kind");
        }
    } // namespace

    std::unique_ptr<HostGraph> generate(const hyt_gen_params &p)
    {
        validate_params(p);
        const int T = resolve_threads(p.threads);
        const std::uint64_t V = 1ULL << p.scale;
        const std::uint64_t n_gen = static_cast<std::uint64_t>(p.edge_factor) * V;
        const EdgeSource es(p);

        Progress pg;
        pg.on = std::getenv("HYT_GEN_PROGRESS") != nullptr;
        // thread 0 walks [lo, hi) in sixteenths, reporting after each
        auto in_steps = [&](int t, std::uint64_t lo, std::uint64_t hi, const char *what, auto body) {
            const std::uint64_t step = (hi - lo + 15) / 16;
            for (std::uint64_t a = lo; a < hi; a += step)
            {
                const std::uint64_t b = std::min(hi, a + step);
                body(a, b);
                if (t == 0)
                    pg.say(what, static_cast<double>(b - lo) / static_cast<double>(hi - lo));
            }
        };

        // 1. generated edges per source (no self-loops)
        std::unique_ptr<std::uint32_t[]> cnt(new std::uint32_t[V]());
        parallel_for(T, n_gen, [&](int t, std::uint64_t lo, std::uint64_t hi) {
            in_steps(t, lo, hi, "count", [&](std::uint64_t a, std::uint64_t b) {
                for (std::uint64_t i = a; i < b; ++i)
                {
                    std::uint32_t s, d;
                    if (es.edge(i, s, d))
                        __atomic_fetch_add(&cnt[s], 1U, __ATOMIC_RELAXED);
                }
            });
        });

        // 2. row ranges for the generated edges, then the scatter
        auto g = std::make_unique<HostGraph>(V, 0, false, p.memory);
        std::uint64_t run = 0;
        for (std::uint64_t v = 0; v < V; ++v)
        {
            g->off[v] = run;
            run += cnt[v];
        }
        g->off[V] = run;
        const std::uint64_t n_raw = run;
        host_free(g->col, p.memory);
        g->col = nullptr;
        g->col = static_cast<std::uint32_t *>(host_alloc(n_raw * sizeof(std::uint32_t), p.memory));
        parallel_for(T, n_gen, [&](int t, std::uint64_t lo, std::uint64_t hi) {
            in_steps(t, lo, hi, "scatter", [&](std::uint64_t a, std::uint64_t b) {
                for (std::uint64_t i = a; i < b; ++i)
                {
                    std::uint32_t s, d;
                    if (es.edge(i, s, d))
                    {
                        const std::uint32_t left = __atomic_sub_fetch(&cnt[s], 1U, __ATOMIC_RELAXED);
                        g->col[g->off[s] + left] = d;
                    }
                }
            });
        });

        // 3. sort + dedupe each row in place; cnt[v] = its unique edges
        constexpr std::uint64_t kRowsPerTask = 4096;
        parallel_dynamic(T, (V + kRowsPerTask - 1) / kRowsPerTask, [&](std::uint64_t task) {
            const std::uint64_t v0 = task * kRowsPerTask, v1 = std::min(V, v0 + kRowsPerTask);
            for (std::uint64_t v = v0; v < v1; ++v)
            {
                std::uint32_t *first = g->col + g->off[v];
                std::uint32_t *last = g->col + g->off[v + 1];
                std::sort(first, last);
                cnt[v] = static_cast<std::uint32_t>(std::unique(first, last) - first);
            }
        });
        pg.say("sort", 1.0);

        // 4. pack the rows to the left in ascending order (a row's final start
        // never exceeds its generated start)
        std::uint64_t pos = 0;
        for (std::uint64_t v = 0; v < V; ++v)
        {
            const std::uint64_t from = g->off[v];
            const std::uint64_t n = cnt[v];
            if (pos != from && n != 0)
                std::memmove(g->col + pos, g->col + from, n * sizeof(std::uint32_t));
            g->off[v] = pos;
            pos += n;
        }
        g->off[V] = pos;
        g->E = pos;
        cnt.reset();
        pg.say("pack", 1.0);

        // 5. weights of the final positions
        if (p.weight_max != 0)
        {
            g->w = static_cast<std::uint32_t *>(host_alloc(g->E * sizeof(std::uint32_t), p.memory));
            g->weighted = true;
            const std::uint64_t range = static_cast<std::uint64_t>(p.weight_max) - p.weight_min + 1;
            parallel_for(T, g->E, [&](int, std::uint64_t lo, std::uint64_t hi) {
                for (std::uint64_t e = lo; e < hi; ++e)
                {
                    std::uint64_t r = draw(p.seed, kStreamWeights, e, 0);
                    g->w[e] = p.weight_min + static_cast<std::uint32_t>(
                                                 bounded32(static_cast<std::uint32_t>(r >> 32), range));
                }
            });
        }
        return g;
    }

    void sample_sources(const std::uint64_t *off, std::uint64_t V, std::uint64_t seed,
                        std::uint32_t count, std::uint32_t *out)
    {
        require(off != nullptr && out != nullptr, "null pointer");
        require(V > 0, "graph has no vertices");
        const std::uint64_t E = off[V];
        require(E > 0, "graph has no edges to sample sources from");
        for (std::uint32_t i = 0; i < count; ++i)
        {
            std::uint64_t e = bounded64(draw(seed, kStreamSources, i, 0), E);
            // Row containing edge e: last v with off[v] <= e.
            const std::uint64_t *it = std::upper_bound(off, off + V + 1, e);
            out[i] = static_cast<std::uint32_t>((it - off) - 1);
        }
    }

} // namespace hyt
