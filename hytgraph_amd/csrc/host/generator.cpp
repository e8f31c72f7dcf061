// Synthetic graph generator: RMAT (Chakrabarti, Zhan, Faloutsos 2004) and
// Erdos-Renyi G(V, E), deterministic for any thread count.
//
// Edge i is a pure function of (seed, i): for RMAT each of the `scale` bit
// levels draws one 32-bit uniform and picks a quadrant with probabilities
// a, b, c, d (top-left = (0,0)). Self-loops are skipped, duplicates removed,
// rows sorted by destination. Weight of the edge at final CSR position p is
// a pure function of (seed, p). Vertex labels are not scrambled.
//
// Layout of the work: edges are bucketed by the top bits of the source
// (counted, then scattered, generating each edge twice instead of storing it
// twice), each bucket is sorted and deduplicated independently, and the
// buckets write disjoint ranges of the CSR. Every phase is parallel over
// buckets or edge ranges.
#include "common.hpp"
#include "host_graph.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
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

            explicit EdgeSource(const hyt_gen_params &params) : p(params)
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

            // Returns false for a self-loop.
            bool edge(std::uint64_t i, std::uint32_t &src, std::uint32_t &dst) const
            {
                if (p.kind == HYT_GEN_ER)
                {
                    std::uint64_t r = draw(p.seed, kStreamEdges, i, 0);
                    std::uint64_t V = 1ULL << p.scale;
                    src = static_cast<std::uint32_t>(bounded32(static_cast<std::uint32_t>(r), V));
                    dst = static_cast<std::uint32_t>(bounded32(static_cast<std::uint32_t>(r >> 32), V));
                }
                else
                {
                    std::uint32_t s = 0, d = 0;
                    std::uint64_t r = 0;
                    for (std::uint32_t level = 0; level < p.scale; ++level)
                    {
                        if ((level & 1U) == 0U)
                            r = draw(p.seed, kStreamEdges, i, level >> 1);
                        std::uint32_t u = (level & 1U) ? static_cast<std::uint32_t>(r >> 32)
                                                       : static_cast<std::uint32_t>(r);
                        std::uint32_t sb, db;
                        if (u < thr_a)
                            sb = 0, db = 0;
                        else if (u < thr_ab)
                            sb = 0, db = 1;
                        else if (u < thr_abc)
                            sb = 1, db = 0;
                        else
                            sb = 1, db = 1;
                        std::uint32_t bit = p.scale - 1U - level;
                        s |= sb << bit;
                        d |= db << bit;
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

        // Buckets by the top bits of the source vertex.
        const std::uint32_t bbits = std::min<std::uint32_t>(p.scale, 16);
        const std::uint32_t bshift = p.scale - bbits;
        const std::uint64_t nb = 1ULL << bbits;

        // Pass 1: per-thread bucket counts.
        std::vector<std::vector<std::uint64_t>> cnt(T, std::vector<std::uint64_t>(nb, 0));
        parallel_for(T, n_gen, [&](int t, std::uint64_t lo, std::uint64_t hi) {
            auto &c = cnt[t];
            for (std::uint64_t i = lo; i < hi; ++i)
            {
                std::uint32_t s, d;
                if (es.edge(i, s, d))
                    ++c[s >> bshift];
            }
        });
        // Bucket starts and per-thread cursors (bucket-major, thread-minor:
        // thread t's edges of bucket b precede thread t+1's).
        std::vector<std::uint64_t> bstart(nb + 1, 0);
        std::vector<std::vector<std::uint64_t>> cur(T, std::vector<std::uint64_t>(nb, 0));
        {
            std::uint64_t run = 0;
            for (std::uint64_t b = 0; b < nb; ++b)
            {
                bstart[b] = run;
                for (int t = 0; t < T; ++t)
                {
                    cur[t][b] = run;
                    run += cnt[t][b];
                }
            }
            bstart[nb] = run;
        }
        cnt.clear();
        const std::uint64_t n_kept = bstart[nb];
        std::vector<std::uint64_t> keys(n_kept);

        // Pass 2: regenerate and scatter keys (src << 32 | dst).
        parallel_for(T, n_gen, [&](int t, std::uint64_t lo, std::uint64_t hi) {
            auto &c = cur[t];
            for (std::uint64_t i = lo; i < hi; ++i)
            {
                std::uint32_t s, d;
                if (es.edge(i, s, d))
                    keys[c[s >> bshift]++] = (static_cast<std::uint64_t>(s) << 32) | d;
            }
        });
        cur.clear();

        // Sort + dedupe each bucket in place.
        std::vector<std::uint64_t> bunique(nb, 0);
        parallel_dynamic(T, nb, [&](std::uint64_t b) {
            auto first = keys.begin() + static_cast<std::ptrdiff_t>(bstart[b]);
            auto last = keys.begin() + static_cast<std::ptrdiff_t>(bstart[b + 1]);
            std::sort(first, last);
            bunique[b] = static_cast<std::uint64_t>(std::unique(first, last) - first);
        });
        std::vector<std::uint64_t> bout(nb + 1, 0);
        for (std::uint64_t b = 0; b < nb; ++b)
            bout[b + 1] = bout[b] + bunique[b];
        const std::uint64_t E = bout[nb];

        auto g = std::make_unique<HostGraph>(V, E, p.weight_max != 0, p.memory);

        // Each bucket writes its rows' offsets and columns.
        parallel_dynamic(T, nb, [&](std::uint64_t b) {
            const std::uint64_t vlo = b << bshift;
            const std::uint64_t vhi = (b + 1) << bshift;
            std::uint64_t pos = bout[b];
            const std::uint64_t *k = keys.data() + bstart[b];
            const std::uint64_t n = bunique[b];
            std::uint64_t j = 0;
            for (std::uint64_t v = vlo; v < vhi; ++v)
            {
                g->off[v] = pos;
                while (j < n && (k[j] >> 32) == v)
                {
                    g->col[pos++] = static_cast<std::uint32_t>(k[j]);
                    ++j;
                }
            }
        });
        g->off[V] = E;
        keys.clear();
        keys.shrink_to_fit();

        if (g->weighted)
        {
            const std::uint64_t range = static_cast<std::uint64_t>(p.weight_max) - p.weight_min + 1;
            parallel_for(T, E, [&](int, std::uint64_t lo, std::uint64_t hi) {
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
