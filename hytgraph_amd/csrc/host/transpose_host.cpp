// Incoming CSR of a host graph, built on host threads.
//
// The original builds the transpose its pull kernels read on the host
// (build_incoming_csr, src/algorithms/sssp.cpp:21-80; cuda/pagerank.cu): a
// counting sort of the edges by destination, sources in ascending order.
// Here each thread owns a range of destinations and scans every edge, so no
// counter is shared and the order within a destination is the scan order
// (ascending source, edge order within a source) whatever the thread count:
//   pass 1: in-degrees of the thread's destinations (vertex ranges of V/T);
//   prefix sum into in_off;
//   pass 2: ranges balanced by in-edges; every source row in order, each edge
//           whose destination is the thread's appends the source at that
//           destination's cursor (in_off[v], advanced in place);
//   the cursors end at the next row's start, so a shift by one restores
//   in_off.
#include "common.hpp"
#include "host_graph.hpp"

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

namespace hyt
{
    namespace
    {
        template <class Fn>
        void run_threads(int T, Fn fn)
        {
            if (T == 1)
            {
                fn(0);
                return;
            }
            std::vector<std::thread> pool;
            pool.reserve(static_cast<std::size_t>(T));
            for (int t = 0; t < T; ++t)
                pool.emplace_back([&fn, t] { fn(t); });
            for (auto &th : pool)
                th.join();
        }
    } // namespace

    void ensure_host_transpose(HostGraph &g, int threads)
    {
        std::lock_guard<std::recursive_mutex> lock(g.transpose_mu);
        if (g.in_off != nullptr && g.in_col != nullptr)
            return;
        const std::uint64_t V = g.V, E = g.E;
        const int hw = static_cast<int>(std::max(1U, std::thread::hardware_concurrency()));
        int T = threads > 0 ? threads : std::min(16, hw);
        T = static_cast<int>(std::max<std::uint64_t>(1, std::min<std::uint64_t>(static_cast<std::uint64_t>(T),
                                                                              std::max<std::uint64_t>(V, 1))));
        T = std::min(T, 64);
        std::uint64_t *in_off = nullptr;
        std::uint32_t *in_col = nullptr;
        try
        {
            in_off = static_cast<std::uint64_t *>(host_alloc((V + 1) * sizeof(std::uint64_t), g.memory));
            in_col = static_cast<std::uint32_t *>(host_alloc(E * sizeof(std::uint32_t), g.memory));
        }
        catch (...)
        {
            host_free(in_off, g.memory);
            throw;
        }
        std::memset(in_off, 0, (V + 1) * sizeof(std::uint64_t));
        const std::uint64_t *off = g.off;
        const std::uint32_t *col = g.col;

        // pass 1: in-degree of v into in_off[v + 1]
        run_threads(T, [&](int t) {
            const std::uint64_t lo = V * static_cast<std::uint64_t>(t) / static_cast<std::uint64_t>(T);
            const std::uint64_t hi = V * static_cast<std::uint64_t>(t + 1) / static_cast<std::uint64_t>(T);
            const std::uint64_t n = hi - lo;
            std::uint64_t *cnt = in_off + 1 + lo;
            for (std::uint64_t e = 0; e < E; ++e)
            {
                const std::uint64_t d = static_cast<std::uint64_t>(col[e]) - lo;
                if (d < n)
                    ++cnt[d];
            }
        });
        for (std::uint64_t v = 0; v < V; ++v)
            in_off[v + 1] += in_off[v];

        // pass 2: destination ranges of about E / T in-edges each
        std::vector<std::uint64_t> cut(static_cast<std::size_t>(T) + 1, V);
        cut[0] = 0;
        for (int t = 1; t < T; ++t)
        {
            const std::uint64_t want = E * static_cast<std::uint64_t>(t) / static_cast<std::uint64_t>(T);
            // first v whose in-edges start at or after `want`
            cut[static_cast<std::size_t>(t)] =
                static_cast<std::uint64_t>(std::lower_bound(in_off, in_off + V, want) - in_off);
        }
        for (int t = 1; t <= T; ++t)
            cut[static_cast<std::size_t>(t)] = std::max(cut[static_cast<std::size_t>(t)], cut[static_cast<std::size_t>(t) - 1]);
        run_threads(T, [&](int t) {
            const std::uint64_t lo = cut[static_cast<std::size_t>(t)];
            const std::uint64_t n = cut[static_cast<std::size_t>(t) + 1] - lo;
            if (n == 0)
                return;
            std::uint64_t *cur = in_off + lo;
            for (std::uint64_t u = 0; u < V; ++u)
            {
                const std::uint64_t b = off[u], e1 = off[u + 1];
                for (std::uint64_t e = b; e < e1; ++e)
                {
                    const std::uint64_t d = static_cast<std::uint64_t>(col[e]) - lo;
                    if (d < n)
                        in_col[cur[d]++] = static_cast<std::uint32_t>(u);
                }
            }
        });
        // in_off[v] now holds the end of v's sources (the start of v + 1)
        std::memmove(in_off + 1, in_off, V * sizeof(std::uint64_t));
        in_off[0] = 0;
        g.in_off = in_off;
        g.in_col = in_col;
    }
} // namespace hyt
