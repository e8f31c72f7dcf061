// Device preprocessing of an uploaded graph: the hub-first vertex order of
// HubSorter::sort (original src/scheduling/hub_sort.cpp:28-183), the CSR
// permutation of CSRGraph::reorder_vertices (src/graph/csr_graph.cpp:171-283)
// and the download of the device CSR.
//
// Hub order: in-degrees by one atomic per edge, the score a*b of every vertex
// (out-degree times in-degree) over max_out*max_in with the original's
// extended-precision division (extended_quotient.hpp), then one stable radix
// sort of (score descending, vertex ascending) pairs: the first hub_count
// vertices in sorted order, then the others in vertex order (a scan of the
// non-hub flags places them). Reorder: the permutation is checked on the
// device (first offending position decides between the original's two
// errors), new row starts are a scan of the permuted degrees, and one
// wavefront per new row copies the row with its columns renamed; the new
// arrays replace the graph's (the transpose and workspaces are dropped).
#include "dgraph.hpp"
#include "extended_quotient.hpp"
#include "kern_common.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace hyt
{
    namespace
    {
        unsigned grid_for(std::uint64_t n)
        {
            std::uint64_t b = (n + kBlock - 1) / kBlock;
            if (b > 65536)
                b = 65536;
            return static_cast<unsigned>(b == 0 ? 1 : b);
        }

        __global__ void count_in_degrees(const std::uint32_t *col, std::uint64_t E, unsigned long long *indeg)
        {
            std::uint64_t e = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; e < E; e += stride)
                atomicAdd(&indeg[col[e]], 1ULL);
        }

        __global__ __launch_bounds__(kBlock) void max_of(const unsigned long long *x, std::uint64_t n,
                                                         unsigned long long *out)
        {
            __shared__ unsigned long long s[kWavesPerBlock];
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            unsigned long long m = 0;
            for (; i < n; i += stride)
                m = x[i] > m ? x[i] : m;
            for (int d = kWave / 2; d > 0; d >>= 1)
            {
                const unsigned long long o = __shfl_down(m, d, kWave);
                m = o > m ? o : m;
            }
            if (lane_id() == 0)
                s[threadIdx.x / kWave] = m;
            __syncthreads();
            if (threadIdx.x == 0)
            {
                for (int w = 1; w < kWavesPerBlock; ++w)
                    m = s[w] > m ? s[w] : m;
                atomicMax(out, m);
            }
        }

        // score[v] = out(v) * in(v) / den (0 when den = 0), and the sort pair:
        // key = ~bits(score) (descending score; scores are >= 0), value = v
        template <class OffT>
        __global__ void hub_scores(const OffT *off, const unsigned long long *indeg, std::uint64_t V,
                                   unsigned long long den, double *score, unsigned long long *key,
                                   std::uint32_t *id)
        {
            std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; v < V; v += stride)
            {
                const unsigned long long a = static_cast<unsigned long long>(off[v + 1] - off[v]) * indeg[v];
                const double s = den ? extended_quotient(a, den) : 0.0;
                score[v] = s;
                key[v] = ~static_cast<unsigned long long>(__double_as_longlong(s));
                id[v] = static_cast<std::uint32_t>(v);
            }
        }

        // order[r] = ranked[r] and hub[ranked[r]] = 1 for r < hubs
        __global__ void place_hubs(const std::uint32_t *ranked, std::uint64_t hubs, std::uint32_t *order,
                                   std::uint32_t *hub)
        {
            std::uint64_t r = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; r < hubs; r += stride)
            {
                order[r] = ranked[r];
                hub[ranked[r]] = 1U;
            }
        }

        __global__ void not_hub(const std::uint32_t *hub, std::uint64_t V, std::uint32_t *flag)
        {
            std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; v < V; v += stride)
                flag[v] = 1U - hub[v];
        }

        // the non-hubs in vertex order after the hubs (pos = exclusive scan of
        // the non-hub flags)
        __global__ void place_rest(const std::uint32_t *hub, const std::uint32_t *pos, std::uint64_t V,
                                   std::uint64_t hubs, std::uint32_t *order)
        {
            std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; v < V; v += stride)
                if (!hub[v])
                    order[hubs + pos[v]] = static_cast<std::uint32_t>(v);
        }

        // Permutation check: bad[0] = first position holding an id >= V,
        // first[o] = first position holding o
        __global__ void order_first(const std::uint32_t *order, std::uint64_t V, unsigned long long *first,
                                    unsigned long long *bad)
        {
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; i < V; i += stride)
            {
                const std::uint32_t o = order[i];
                if (o >= V)
                    atomicMin(&bad[0], static_cast<unsigned long long>(i));
                else
                    atomicMin(&first[o], static_cast<unsigned long long>(i));
            }
        }

        // bad[1] = first position repeating an earlier id; new_of_old and the
        // permuted degrees
        template <class OffT>
        __global__ void order_repeat(const std::uint32_t *order, const OffT *off, std::uint64_t V,
                                     const unsigned long long *first, unsigned long long *bad,
                                     std::uint32_t *new_of_old, unsigned long long *deg)
        {
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; i < V; i += stride)
            {
                const std::uint32_t o = order[i];
                deg[i] = 0;
                if (o >= V)
                    continue;
                if (first[o] != i)
                    atomicMin(&bad[1], static_cast<unsigned long long>(i));
                new_of_old[o] = static_cast<std::uint32_t>(i);
                deg[i] = static_cast<unsigned long long>(off[o + 1] - off[o]);
            }
        }

        // One wavefront per new row: the old row's edges, columns renamed.
        template <class OffT>
        __global__ __launch_bounds__(kBlock) void move_rows(const OffT *off, const std::uint32_t *col,
                                                            const std::uint32_t *w, const std::uint32_t *order,
                                                            const std::uint32_t *new_of_old,
                                                            const unsigned long long *start, std::uint64_t V,
                                                            std::uint64_t E, OffT *off2, std::uint32_t *col2,
                                                            std::uint32_t *w2)
        {
            const int lane = lane_id();
            const std::uint64_t waves = static_cast<std::uint64_t>(gridDim.x) * kWavesPerBlock;
            for (std::uint64_t n = (static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave; n < V;
                 n += waves)
            {
                const std::uint32_t u = order[n];
                const std::uint64_t b = off[u], cnt = static_cast<std::uint64_t>(off[u + 1]) - b;
                const std::uint64_t nb = start[n];
                for (std::uint64_t k = static_cast<std::uint64_t>(lane); k < cnt; k += kWave)
                {
                    col2[nb + k] = new_of_old[col[b + k]];
                    if (w2 != nullptr)
                        w2[nb + k] = w[b + k];
                }
                if (lane == 0)
                {
                    off2[n] = static_cast<OffT>(nb);
                    if (n == V - 1)
                        off2[V] = static_cast<OffT>(E);
                }
            }
        }

        template <class OffT>
        __global__ void widen_offsets(const OffT *off, std::uint64_t n, unsigned long long *out)
        {
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; i < n; i += stride)
                out[i] = static_cast<unsigned long long>(off[i]);
        }

        // hub_count of the original: V * fraction, rounded to the nearest
        // integer when within 1e-12 of it, else up; at most V
        std::uint64_t hub_count_of(std::uint64_t V, double fraction)
        {
            const long double want = static_cast<long double>(V) * static_cast<long double>(fraction);
            const long double nearest = std::round(want);
            const long double count = std::fabs(want - nearest) < 1e-12L ? nearest : std::ceil(want);
            return std::min<std::uint64_t>(V, static_cast<std::uint64_t>(count));
        }

        template <class OffT>
        void hub_sort_typed(DeviceGraph &g, double fraction, std::uint32_t *order_out, double *scores_out,
                            bool out_on_device, std::uint64_t &hubs)
        {
            hipStream_t s = g.stream;
            const std::uint64_t V = g.V, E = g.E;
            require(V < (1ULL << 31), "device hub sort needs fewer than 2^31 vertices (radix sort limit)");
            DeviceBuffer indeg((V + 1) * 8, "hipMalloc(in-degrees)"), mx(8, "hipMalloc(max)");
            hip_check(hipMemsetAsync(indeg.ptr, 0, V * 8, s), "clear in-degrees");
            hip_check(hipMemsetAsync(mx.ptr, 0, 8, s), "clear max");
            if (E)
            {
                count_in_degrees<<<grid_for(E), kBlock, 0, s>>>(g.col.as<std::uint32_t>(), E,
                                                                indeg.as<unsigned long long>());
                hip_check(hipGetLastError(), "launch count_in_degrees");
            }
            max_of<<<std::min<unsigned>(grid_for(V), 1024), kBlock, 0, s>>>(indeg.as<unsigned long long>(), V,
                                                                            mx.as<unsigned long long>());
            hip_check(hipGetLastError(), "launch max_of");
            unsigned long long max_in = 0;
            hip_check(hipMemcpyAsync(&max_in, mx.ptr, 8, hipMemcpyDeviceToHost, s), "read max in-degree");
            hip_check(hipStreamSynchronize(s), "hub sort degrees");
            const unsigned long long max_out = g.max_degree;
            require(max_in == 0 || max_out <= ~0ULL / max_in,
                    "device hub sort needs max out-degree * max in-degree < 2^64");
            const unsigned long long den = max_out * max_in;

            DeviceBuffer score(V * 8, "hipMalloc(hub scores)");
            DeviceBuffer keys(V * 8, "hipMalloc(hub keys)"), keys2(V * 8, "hipMalloc(hub keys)");
            DeviceBuffer ids(V * 4, "hipMalloc(hub ids)"), ranked(V * 4, "hipMalloc(hub ranks)");
            hub_scores<OffT><<<grid_for(V), kBlock, 0, s>>>(g.off.as<OffT>(), indeg.as<unsigned long long>(), V, den,
                                                           score.as<double>(), keys.as<unsigned long long>(),
                                                           ids.as<std::uint32_t>());
            hip_check(hipGetLastError(), "launch hub_scores");
            const int n = static_cast<int>(V);
            std::size_t temp = 0;
            hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, keys.as<unsigned long long>(),
                                                         keys2.as<unsigned long long>(), ids.as<std::uint32_t>(),
                                                         ranked.as<std::uint32_t>(), n, 0, 64, s),
                      "hub sort size");
            DeviceBuffer tmp(temp + 16, "hipMalloc(sort temp)");
            hip_check(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, temp, keys.as<unsigned long long>(),
                                                         keys2.as<unsigned long long>(), ids.as<std::uint32_t>(),
                                                         ranked.as<std::uint32_t>(), n, 0, 64, s),
                      "hub sort");

            hubs = hub_count_of(V, fraction);
            // hub flags, the scan of the non-hub flags, the order (reusing
            // the key arrays: 8 bytes per vertex each)
            std::uint32_t *hub = reinterpret_cast<std::uint32_t *>(keys.ptr);
            std::uint32_t *flag = hub + V;
            std::uint32_t *pos = reinterpret_cast<std::uint32_t *>(keys2.ptr);
            DeviceBuffer order_dev;
            std::uint32_t *order = order_out;
            if (!out_on_device)
            {
                order_dev.alloc(V * 4, "hipMalloc(vertex order)");
                order = order_dev.as<std::uint32_t>();
            }
            hip_check(hipMemsetAsync(hub, 0, V * 4, s), "clear hub flags");
            if (hubs)
            {
                place_hubs<<<grid_for(hubs), kBlock, 0, s>>>(ranked.as<std::uint32_t>(), hubs, order, hub);
                hip_check(hipGetLastError(), "launch place_hubs");
            }
            not_hub<<<grid_for(V), kBlock, 0, s>>>(hub, V, flag);
            hip_check(hipGetLastError(), "launch not_hub");
            temp = 0;
            hip_check(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, flag, pos, n, s), "scan size");
            if (temp + 16 > tmp.bytes)
                tmp.alloc(temp + 16, "hipMalloc(scan temp)");
            hip_check(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, temp, flag, pos, n, s), "scan non-hubs");
            place_rest<<<grid_for(V), kBlock, 0, s>>>(hub, pos, V, hubs, order);
            hip_check(hipGetLastError(), "launch place_rest");
            const hipMemcpyKind k = out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
            if (!out_on_device)
                hip_check(hipMemcpyAsync(order_out, order, V * 4, k, s), "copy vertex order");
            if (scores_out)
                hip_check(hipMemcpyAsync(scores_out, score.ptr, V * 8, k, s), "copy hub scores");
            hip_check(hipStreamSynchronize(s), "hub sort");
        }

        template <class OffT>
        void reorder_typed(DeviceGraph &g, const std::uint32_t *order_in, bool on_device)
        {
            hipStream_t s = g.stream;
            const std::uint64_t V = g.V, E = g.E;
            require(V < (1ULL << 31), "device reorder needs fewer than 2^31 vertices (scan limit)");
            DeviceBuffer order_dev;
            const std::uint32_t *order = order_in;
            if (!on_device)
            {
                order_dev.alloc(V * 4, "hipMalloc(vertex order)");
                hip_check(hipMemcpyAsync(order_dev.ptr, order_in, V * 4, hipMemcpyHostToDevice, s), "copy order");
                order = order_dev.as<std::uint32_t>();
            }
            DeviceBuffer first(V * 8, "hipMalloc(first positions)"), bad(16, "hipMalloc(order check)");
            DeviceBuffer new_of_old(V * 4, "hipMalloc(new ids)"), deg((V + 1) * 8, "hipMalloc(new degrees)");
            DeviceBuffer start((V + 1) * 8, "hipMalloc(new row starts)");
            hip_check(hipMemsetAsync(first.ptr, 0xFF, V * 8, s), "init first positions");
            hip_check(hipMemsetAsync(bad.ptr, 0xFF, 16, s), "init order check");
            order_first<<<grid_for(V), kBlock, 0, s>>>(order, V, first.as<unsigned long long>(),
                                                       bad.as<unsigned long long>());
            hip_check(hipGetLastError(), "launch order_first");
            order_repeat<OffT><<<grid_for(V), kBlock, 0, s>>>(order, g.off.as<OffT>(), V,
                                                             first.as<unsigned long long>(),
                                                             bad.as<unsigned long long>(),
                                                             new_of_old.as<std::uint32_t>(),
                                                             deg.as<unsigned long long>());
            hip_check(hipGetLastError(), "launch order_repeat");
            unsigned long long hb[2] = {0, 0};
            hip_check(hipMemcpyAsync(hb, bad.ptr, 16, hipMemcpyDeviceToHost, s), "read order check");
            hip_check(hipStreamSynchronize(s), "order check");
            if (hb[0] != ~0ULL || hb[1] != ~0ULL)
            {
                if (hb[0] < hb[1])
                    fail(HYT_ERR_INVALID_ARGUMENT,
                         "CSRGraph::reorder_vertices: vertex_order contains an out-of-range vertex ID");
                fail(HYT_ERR_INVALID_ARGUMENT, "CSRGraph::reorder_vertices: vertex_order contains duplicate vertex IDs");
            }
            std::size_t temp = 0;
            const int n = static_cast<int>(V);
            hip_check(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, deg.as<unsigned long long>(),
                                                       start.as<unsigned long long>(), n, s),
                      "scan size");
            DeviceBuffer tmp(temp + 16, "hipMalloc(scan temp)");
            hip_check(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, temp, deg.as<unsigned long long>(),
                                                       start.as<unsigned long long>(), n, s),
                      "scan degrees");
            DeviceBuffer off2((V + 1) * sizeof(OffT), "hipMalloc(row offsets)");
            DeviceBuffer col2(E * 4, "hipMalloc(columns)");
            DeviceBuffer w2;
            if (g.weighted)
                w2.alloc(E * 4, "hipMalloc(weights)");
            move_rows<OffT><<<std::min<unsigned>(grid_for(V * kWave), 65536), kBlock, 0, s>>>(
                g.off.as<OffT>(), g.col.as<std::uint32_t>(), g.weighted ? g.w.as<std::uint32_t>() : nullptr, order,
                new_of_old.as<std::uint32_t>(), start.as<unsigned long long>(), V, E, off2.as<OffT>(),
                col2.as<std::uint32_t>(), g.weighted ? w2.as<std::uint32_t>() : nullptr);
            hip_check(hipGetLastError(), "launch move_rows");
            hip_check(hipStreamSynchronize(s), "reorder");
            // the transpose, the workspaces and the captured rounds name the
            // old arrays or depend on the vertex numbering
            g.round_graph.reset();
            g.ws = Workspace{};
            g.kws.reset();
            g.in_off.release();
            g.in_col.release();
            g.in_w.release();
            g.in_heavy.release();
            g.in_heavy_count = 0;
            g.in_heavy_light_max = 0;
            g.in_heavy_chunk = 0;
            g.off = std::move(off2);
            g.col = std::move(col2);
            if (g.weighted)
                g.w = std::move(w2);
        }
    } // namespace
} // namespace hyt

extern "C" {

int hyt_hub_sort(hyt_graph_t g, double hub_fraction, uint32_t *vertex_order, double *scores, uint64_t *hub_count,
                 uint32_t out_on_device)
{
    return hyt::guarded("hyt_hub_sort", [&] {
        using namespace hyt;
        require(g != nullptr && vertex_order != nullptr, "null graph or vertex_order");
        if (!std::isfinite(hub_fraction) || hub_fraction < 0.0 || hub_fraction > 1.0)
            fail(HYT_ERR_INVALID_ARGUMENT, "HubSortOptions.hub_fraction must be finite and in [0, 1]");
        DeviceGraph &dg = g->g;
        hip_check(hipSetDevice(dg.device), "hipSetDevice");
        std::uint64_t hubs = 0;
        if (dg.off64)
            hub_sort_typed<std::uint64_t>(dg, hub_fraction, vertex_order, scores, out_on_device != 0, hubs);
        else
            hub_sort_typed<std::uint32_t>(dg, hub_fraction, vertex_order, scores, out_on_device != 0, hubs);
        if (hub_count)
            *hub_count = hubs;
    });
}

int hyt_graph_reorder(hyt_graph_t g, const uint32_t *vertex_order, uint32_t order_on_device)
{
    return hyt::guarded("hyt_graph_reorder", [&] {
        using namespace hyt;
        require(g != nullptr && vertex_order != nullptr, "null graph or vertex_order");
        DeviceGraph &dg = g->g;
        hip_check(hipSetDevice(dg.device), "hipSetDevice");
        hip_check(hipStreamSynchronize(dg.stream), "hyt_graph_reorder");
        if (dg.off64)
            reorder_typed<std::uint64_t>(dg, vertex_order, order_on_device != 0);
        else
            reorder_typed<std::uint32_t>(dg, vertex_order, order_on_device != 0);
    });
}

int hyt_graph_download(hyt_graph_t g, uint64_t *row_offsets, uint32_t *columns, uint32_t *weights)
{
    return hyt::guarded("hyt_graph_download", [&] {
        using namespace hyt;
        require(g != nullptr, "null graph");
        DeviceGraph &dg = g->g;
        require(weights == nullptr || dg.weighted, "the graph has no weights");
        hip_check(hipSetDevice(dg.device), "hipSetDevice");
        hipStream_t s = dg.stream;
        if (row_offsets)
        {
            DeviceBuffer wide((dg.V + 1) * 8, "hipMalloc(offsets)");
            if (dg.off64)
                widen_offsets<std::uint64_t><<<grid_for(dg.V + 1), kBlock, 0, s>>>(
                    dg.off.as<std::uint64_t>(), dg.V + 1, wide.as<unsigned long long>());
            else
                widen_offsets<std::uint32_t><<<grid_for(dg.V + 1), kBlock, 0, s>>>(
                    dg.off.as<std::uint32_t>(), dg.V + 1, wide.as<unsigned long long>());
            hip_check(hipGetLastError(), "launch widen_offsets");
            hip_check(hipMemcpyAsync(row_offsets, wide.ptr, (dg.V + 1) * 8, hipMemcpyDeviceToHost, s), "copy offsets");
            hip_check(hipStreamSynchronize(s), "download offsets");
        }
        if (columns && dg.E)
            hip_check(hipMemcpyAsync(columns, dg.col.ptr, dg.E * 4, hipMemcpyDeviceToHost, s), "copy columns");
        if (weights && dg.E)
            hip_check(hipMemcpyAsync(weights, dg.w.ptr, dg.E * 4, hipMemcpyDeviceToHost, s), "copy weights");
        hip_check(hipStreamSynchronize(s), "download");
    });
}

} // extern "C"
