// Device graph upload / teardown and workspace management.
#include "dgraph.hpp"

#include <algorithm>
#include <string>
#include <vector>

namespace hyt
{
    DeviceGraph::~DeviceGraph()
    {
        if (stream)
            (void)hipStreamSynchronize(stream);
        if (ev0)
            (void)hipEventDestroy(ev0);
        if (ev1)
            (void)hipEventDestroy(ev1);
        if (stream)
            (void)hipStreamDestroy(stream);
    }

    namespace
    {
        // CSR invariants of the original CSRGraph::validate: offsets start at
        // 0, are monotone and end at E; every column id is < V.
        void validate_csr(std::uint64_t V, std::uint64_t E, const std::uint64_t *off, const std::uint32_t *col)
        {
            require(V > 0, "graph must have at least one vertex");
            require(V <= 0xFFFFFFFFULL, "vertex ids must fit in 32 bits (V <= 2^32-1)");
            require(off != nullptr, "row_offsets is null");
            require(E == 0 || col != nullptr, "columns is null");
            require(off[0] == 0, "row_offsets[0] must be 0");
            require(off[V] == E, "row_offsets[V] must equal the edge count");
            for (std::uint64_t v = 0; v < V; ++v)
                if (off[v + 1] < off[v])
                    fail(HYT_ERR_INVALID_ARGUMENT, "row_offsets must be non-decreasing (vertex " +
                                                       std::to_string(v) + ")");
            for (std::uint64_t e = 0; e < E; ++e)
                if (col[e] >= V)
                    fail(HYT_ERR_INVALID_ARGUMENT, "column index out of range at edge " + std::to_string(e));
        }

    } // namespace

    void ensure_workspace(DeviceGraph &g, std::uint32_t light_max, std::uint32_t chunk)
    {
        Workspace &ws = g.ws;
        const std::uint64_t Vp = g.V + kPad;
        const std::uint64_t cap = heavy_capacity(g.E, light_max, chunk);
        if (!ws.ready)
        {
            ws.dist.alloc(Vp * 4, "hipMalloc(dist)");
            ws.prev.alloc(Vp * 4, "hipMalloc(prev)");
            ws.touch.alloc(Vp, "hipMalloc(touch)");
            ws.q.alloc(Vp * 4, "hipMalloc(queue)");
            ws.qd.alloc(Vp * 4, "hipMalloc(queue values)");
            ws.ctr.alloc(sizeof(Counters), "hipMalloc(counters)");
            ws.partial.alloc(4096 * sizeof(double), "hipMalloc(partials)");
            ws.active.alloc(Vp / 8 + 64, "hipMalloc(frontier bitmap)");
            ws.share.alloc(Vp * 8, "hipMalloc(push shares)");
            ws.rc.alloc(sizeof(RoundCtl), "hipMalloc(round control)");
            // touch must start (and is kept) all-zero: compaction clears what it reads
            hip_check(hipMemsetAsync(ws.touch.ptr, 0, Vp, g.stream), "clear touch");
            ws.ready = true;
        }
        if (ws.heavy.bytes < cap * sizeof(HeavyChunk))
        {
            ws.heavy.alloc(cap * sizeof(HeavyChunk), "hipMalloc(heavy chunks)");
        }
        ws.heavy_cap = ws.heavy.bytes / sizeof(HeavyChunk);
    }

    void ensure_mass_state(DeviceGraph &g)
    {
        const std::uint64_t bytes = (g.V + kPad) * 8;
        g.ws.res.ensure(bytes, "hipMalloc(PageRank residuals)");
        g.ws.rank.ensure(bytes, "hipMalloc(PageRank ranks)");
    }
} // namespace hyt

extern "C" {

int hyt_graph_upload(uint64_t num_vertices, uint64_t num_edges, const uint64_t *row_offsets,
                     const uint32_t *columns, const uint32_t *weights, uint32_t flags,
                     hyt_graph_t *out)
{
    return hyt::guarded("hyt_graph_upload", [&] {
        using namespace hyt;
        require(out != nullptr, "out is null");
        *out = nullptr;
        require((flags & ~static_cast<uint32_t>(HYT_GRAPH_WEIGHTS_F32 | HYT_GRAPH_OFF64)) == 0, "unknown graph flags");
        const bool f32 = (flags & HYT_GRAPH_WEIGHTS_F32) != 0;
        validate_csr(num_vertices, num_edges, row_offsets, columns);
        if (weights != nullptr && f32)
        {
            // Float weights must be finite and >= 0 (original sssp_cpu check);
            // -0.0 is allowed and behaves as +0 in the sums.
            for (std::uint64_t e = 0; e < num_edges; ++e)
            {
                std::uint32_t bits = weights[e];
                bool neg = (bits >> 31) != 0 && (bits & 0x7FFFFFFFU) != 0;
                bool nonfinite = (bits & 0x7F800000U) == 0x7F800000U;
                if (neg || nonfinite)
                    fail(HYT_ERR_INVALID_ARGUMENT, "SSSP requires finite non-negative edge weights");
            }
        }
        const int dev = require_device();
        auto h = std::make_unique<hyt_graph>();
        DeviceGraph &g = h->g;
        g.V = num_vertices;
        g.E = num_edges;
        g.device = dev;
        g.off64 = num_edges >= 0xFFFFFFFFULL || (flags & HYT_GRAPH_OFF64) != 0;
        g.weighted = weights != nullptr;
        g.weights_f32 = f32;
        hip_check(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreate(&g.ev0), "hipEventCreate");
        hip_check(hipEventCreate(&g.ev1), "hipEventCreate");
        g.host_ctr = std::make_unique<PinnedBuffer>(sizeof(Counters) + sizeof(RoundCtl));

        const std::uint64_t V = num_vertices, E = num_edges;
        for (std::uint64_t v = 0; v < V; ++v)
            g.max_degree = std::max<std::uint64_t>(g.max_degree, row_offsets[v + 1] - row_offsets[v]);
        if (g.off64)
        {
            g.off.alloc((V + 1) * 8, "hipMalloc(row offsets)");
            hip_check(hipMemcpy(g.off.ptr, row_offsets, (V + 1) * 8, hipMemcpyHostToDevice), "copy offsets");
        }
        else
        {
            std::vector<std::uint32_t> off32(V + 1);
            for (std::uint64_t v = 0; v <= V; ++v)
                off32[v] = static_cast<std::uint32_t>(row_offsets[v]);
            g.off.alloc((V + 1) * 4, "hipMalloc(row offsets)");
            hip_check(hipMemcpy(g.off.ptr, off32.data(), (V + 1) * 4, hipMemcpyHostToDevice), "copy offsets");
        }
        g.col.alloc(E * 4, "hipMalloc(columns)");
        if (E)
            hip_check(hipMemcpy(g.col.ptr, columns, E * 4, hipMemcpyHostToDevice), "copy columns");
        if (g.weighted)
        {
            g.w.alloc(E * 4, "hipMalloc(weights)");
            if (E)
                hip_check(hipMemcpy(g.w.ptr, weights, E * 4, hipMemcpyHostToDevice), "copy weights");
        }
        // the traversal workspace is allocated by the first single-source or
        // batched call (lane sessions bring their own)
        hip_check(hipStreamSynchronize(g.stream), "upload synchronize");
        *out = h.release();
    });
}

int hyt_graph_trim(hyt_graph_t g)
{
    return hyt::guarded("hyt_graph_trim", [&] {
        hyt::require(g != nullptr, "null graph");
        hyt::DeviceGraph &d = g->g;
        hyt::hip_check(hipStreamSynchronize(d.stream), "hyt_graph_trim");
        d.round_graph.reset(); // its nodes name workspace arrays
        d.pws.reset();
        d.ws = hyt::Workspace{};
        d.kws.reset();
        d.in_off.release();
        d.in_col.release();
        d.in_w.release();
        d.in_heavy.release();
        d.in_heavy_count = 0;
        d.in_heavy_light_max = 0;
        d.in_heavy_chunk = 0;
    });
}

int hyt_graph_free(hyt_graph_t g)
{
    return hyt::guarded("hyt_graph_free", [&] { delete g; });
}

int hyt_graph_info(hyt_graph_t g, uint64_t *num_vertices, uint64_t *num_edges, uint64_t *device_bytes)
{
    return hyt::guarded("hyt_graph_info", [&] {
        hyt::require(g != nullptr, "null graph");
        if (num_vertices)
            *num_vertices = g->g.V;
        if (num_edges)
            *num_edges = g->g.E;
        if (device_bytes)
            *device_bytes = g->g.device_bytes();
    });
}

} // extern "C"
