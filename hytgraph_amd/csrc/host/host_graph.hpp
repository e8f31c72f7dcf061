// Host-resident CSR owned by the library (generator output, or the
// host-side graph of the hybrid transfer engine). Arrays are either pageable
// (std::malloc) or pinned and device-mapped (hipHostMalloc), the latter so the
// transfer engine can DMA from them and zero-copy kernels can read them.
#pragma once

#include "common.hpp"

#include <cstdint>
#include <memory>
#include <mutex>

namespace hyt
{
    struct HostGraph
    {
        std::uint64_t V = 0, E = 0;
        bool weighted = false;
        int memory = HYT_HOST_PAGEABLE;
        std::uint64_t *off = nullptr;
        std::uint32_t *col = nullptr;
        std::uint32_t *w = nullptr;
        // incoming CSR (the transpose, sources of each destination in
        // ascending order), same memory kind; built on first use, under
        // transpose_mu (engines on several threads may share a host graph)
        std::uint64_t *in_off = nullptr;
        std::uint32_t *in_col = nullptr;
        std::recursive_mutex transpose_mu;

        HostGraph(std::uint64_t v, std::uint64_t e, bool weights, int mem);
        ~HostGraph();
        HostGraph(const HostGraph &) = delete;
        HostGraph &operator=(const HostGraph &) = delete;
    };

    // Allocation helpers for host memory of the given kind.
    void *host_alloc(std::size_t bytes, int memory);
    void host_free(void *p, int memory) noexcept;

    std::unique_ptr<HostGraph> generate(const hyt_gen_params &p);
    // The same graph built on the current HIP device, the key buckets within
    // budget bytes of HBM (0: 40 GiB); params.threads is not used.
    std::unique_ptr<HostGraph> generate_device(const hyt_gen_params &p, std::uint64_t budget);
    HostGraph *host_graph_of(hyt_host_graph_t h);
    // Builds g.in_off / g.in_col if absent, on up to `threads` host threads.
    void ensure_host_transpose(HostGraph &g, int threads);
    void sample_sources(const std::uint64_t *off, std::uint64_t V, std::uint64_t seed,
                        std::uint32_t count, std::uint32_t *out);
} // namespace hyt

// Handle behind hyt_host_graph_t.
struct hyt_host_graph
{
    std::unique_ptr<hyt::HostGraph> g;
};
