// C-ABI entry points that need no device: error reporting, version,
// generator, host graph views and source sampling.
#include "common.hpp"
#include "host_graph.hpp"

#include <hip/hip_runtime_api.h>

#include <cstdlib>
#include <string>

namespace hyt
{
    namespace
    {
        thread_local std::string g_last_error;
    }

    void set_last_error(const std::string &msg) { g_last_error = msg; }

    void *host_alloc(std::size_t bytes, int memory)
    {
        if (bytes == 0)
            bytes = 1;
        if (memory == HYT_HOST_PINNED)
        {
            void *p = nullptr;
            hipError_t st = hipHostMalloc(&p, bytes, hipHostMallocMapped | hipHostMallocPortable);
            if (st != hipSuccess || p == nullptr)
                fail(HYT_ERR_OUT_OF_MEMORY, std::string("hipHostMalloc(") + std::to_string(bytes) +
                                                " bytes): " + hipGetErrorString(st));
            return p;
        }
        void *p = std::malloc(bytes);
        if (p == nullptr)
            fail(HYT_ERR_OUT_OF_MEMORY, "host malloc of " + std::to_string(bytes) + " bytes failed");
        return p;
    }

    void host_free(void *p, int memory) noexcept
    {
        if (p == nullptr)
            return;
        if (memory == HYT_HOST_PINNED)
            (void)hipHostFree(p);
        else
            std::free(p);
    }

    HostGraph::HostGraph(std::uint64_t v, std::uint64_t e, bool weights, int mem)
        : V(v), E(e), weighted(weights), memory(mem)
    {
        try
        {
            off = static_cast<std::uint64_t *>(host_alloc((V + 1) * sizeof(std::uint64_t), memory));
            col = static_cast<std::uint32_t *>(host_alloc(E * sizeof(std::uint32_t), memory));
            if (weighted)
                w = static_cast<std::uint32_t *>(host_alloc(E * sizeof(std::uint32_t), memory));
        }
        catch (...)
        {
            host_free(off, memory);
            host_free(col, memory);
            host_free(w, memory);
            throw;
        }
    }

    HostGraph::~HostGraph()
    {
        host_free(off, memory);
        host_free(col, memory);
        host_free(w, memory);
        host_free(in_off, memory);
        host_free(in_col, memory);
    }
} // namespace hyt

namespace hyt
{
    HostGraph *host_graph_of(hyt_host_graph_t h)
    {
        require(h != nullptr && h->g, "null host graph");
        return h->g.get();
    }
} // namespace hyt

extern "C" {

const char *hyt_last_error(void) { return hyt::g_last_error.c_str(); }

const char *hyt_version(void) { return "hytgraph_amd 0.1.0 (gfx950)"; }

int hyt_device_count(int *count)
{
    return hyt::guarded("hyt_device_count", [&] {
        hyt::require(count != nullptr, "count is null");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess)
            n = 0;
        *count = n;
    });
}

int hyt_generate(const hyt_gen_params *params, hyt_host_graph_t *out)
{
    return hyt::guarded("hyt_generate", [&] {
        hyt::require(params != nullptr && out != nullptr, "null pointer");
        *out = nullptr;
        auto h = new hyt_host_graph{hyt::generate(*params)};
        *out = h;
    });
}

int hyt_host_graph_view(hyt_host_graph_t g, uint64_t *num_vertices, uint64_t *num_edges,
                        const uint64_t **row_offsets, const uint32_t **columns,
                        const uint32_t **weights)
{
    return hyt::guarded("hyt_host_graph_view", [&] {
        hyt::require(g != nullptr && g->g, "null host graph");
        if (num_vertices)
            *num_vertices = g->g->V;
        if (num_edges)
            *num_edges = g->g->E;
        if (row_offsets)
            *row_offsets = g->g->off;
        if (columns)
            *columns = g->g->col;
        if (weights)
            *weights = g->g->w;
    });
}

int hyt_host_graph_transpose(hyt_host_graph_t g, uint32_t threads, const uint64_t **in_offsets,
                             const uint32_t **in_sources)
{
    return hyt::guarded("hyt_host_graph_transpose", [&] {
        hyt::require(g != nullptr && g->g, "null host graph");
        hyt::ensure_host_transpose(*g->g, static_cast<int>(threads));
        if (in_offsets)
            *in_offsets = g->g->in_off;
        if (in_sources)
            *in_sources = g->g->in_col;
    });
}

int hyt_host_graph_free(hyt_host_graph_t g)
{
    return hyt::guarded("hyt_host_graph_free", [&] { delete g; });
}

int hyt_sample_sources(const uint64_t *row_offsets, uint64_t num_vertices, uint64_t seed,
                       uint32_t count, uint32_t *out_sources)
{
    return hyt::guarded("hyt_sample_sources", [&] {
        hyt::sample_sources(row_offsets, num_vertices, seed, count, out_sources);
    });
}

} // extern "C"
