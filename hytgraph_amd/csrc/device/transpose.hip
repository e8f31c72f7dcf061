// Device-built transpose (incoming CSR) for pull rounds.
//
// keys[e] = (dst << 32) | src for every edge, radix-sorted (rocPRIM via
// hipCUB) together with the weights, gives rows ordered by destination and,
// within a row, sources in ascending order. Row starts come from a lower_bound
// per vertex. Built once per graph, on first use of a pull round, in HBM.
#include "dgraph.hpp"
#include "kern_common.hpp"

#include <hipcub/hipcub.hpp>

namespace hyt
{
    namespace
    {
        template <class OffT>
        __global__ void edge_keys(const OffT *off, std::uint64_t V, const std::uint32_t *col, std::uint64_t E,
                                  unsigned long long *keys)
        {
            std::uint64_t e = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; e < E; e += stride)
            {
                // row of edge e: last u with off[u] <= e
                std::uint64_t lo = 0, hi = V; // off[lo] <= e < off[hi]
                while (hi - lo > 1)
                {
                    const std::uint64_t mid = (lo + hi) >> 1;
                    if (static_cast<std::uint64_t>(off[mid]) <= e)
                        lo = mid;
                    else
                        hi = mid;
                }
                keys[e] = (static_cast<unsigned long long>(col[e]) << 32) | lo;
            }
        }

        template <class OffT>
        __global__ void row_starts(const unsigned long long *keys, std::uint64_t E, std::uint64_t V, OffT *in_off,
                                   std::uint32_t *in_col)
        {
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (std::uint64_t v = i; v <= V; v += stride)
            {
                // first key >= (v << 32)
                const unsigned long long target = static_cast<unsigned long long>(v) << 32;
                std::uint64_t lo = 0, hi = E;
                while (lo < hi)
                {
                    const std::uint64_t mid = (lo + hi) >> 1;
                    if (keys[mid] < target)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                in_off[v] = static_cast<OffT>(lo);
            }
            for (std::uint64_t e = i; e < E; e += stride)
                in_col[e] = static_cast<std::uint32_t>(keys[e]);
        }

        // Static chunk list of the rows whose in-degree exceeds light_max.
        template <class OffT>
        __global__ __launch_bounds__(kBlock) void heavy_rows(const OffT *in_off, std::uint64_t V,
                                                             std::uint32_t light_max, std::uint32_t chunk,
                                                             HeavyChunk *out, std::uint64_t cap, Counters *ctr)
        {
            const std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            std::uint32_t nch = 0;
            std::uint64_t deg = 0, begin = 0;
            if (v < V)
            {
                begin = in_off[v];
                deg = static_cast<std::uint64_t>(in_off[v + 1]) - begin;
                if (deg > light_max)
                    nch = static_cast<std::uint32_t>((deg + chunk - 1) / chunk);
            }
            const int lane = lane_id();
            const std::uint32_t incl = wave_incl_scan(nch);
            const std::uint32_t total = __shfl(incl, kWave - 1, kWave);
            if (total == 0)
                return;
            std::uint32_t base = 0;
            if (lane == 0)
                base = atomicAdd(&ctr->heavy_count, total);
            base = __shfl(base, 0, kWave);
            if (static_cast<std::uint64_t>(base) + total > cap)
            {
                if (lane == 0)
                    atomicOr(&ctr->heavy_overflow, 1U);
                return;
            }
            for (std::uint32_t c = 0; c < nch; ++c)
            {
                HeavyChunk h;
                h.value = static_cast<std::uint32_t>(v);
                const std::uint64_t lo = static_cast<std::uint64_t>(c) * chunk;
                h.count = static_cast<std::uint32_t>(deg - lo < chunk ? deg - lo : chunk);
                h.begin = begin + lo;
                h.row = static_cast<std::uint32_t>(v);
                h.record = 0;
                out[base + incl - nch + c] = h;
            }
        }

        unsigned grid_for(std::uint64_t n)
        {
            std::uint64_t b = (n + kBlock - 1) / kBlock;
            if (b > 65536)
                b = 65536;
            return static_cast<unsigned>(b == 0 ? 1 : b);
        }

        template <class OffT>
        void build_typed(DeviceGraph &g)
        {
            hipStream_t s = g.stream;
            const std::uint64_t V = g.V, E = g.E;
            require(E < (1ULL << 31), "pull rounds need fewer than 2^31 edges (transpose sort limit)");
            int vbits = 1;
            while ((1ULL << vbits) < V)
                ++vbits;
            DeviceBuffer keys_in(E * 8 + 8, "hipMalloc(transpose keys)");
            DeviceBuffer keys_out(E * 8 + 8, "hipMalloc(transpose keys)");
            edge_keys<OffT><<<grid_for(E), kBlock, 0, s>>>(g.off.as<OffT>(), V, g.col.as<std::uint32_t>(), E,
                                                          keys_in.as<unsigned long long>());
            hip_check(hipGetLastError(), "launch edge_keys");
            g.in_off.alloc((V + 1) * sizeof(OffT), "hipMalloc(in offsets)");
            g.in_col.alloc(E * 4 + 4, "hipMalloc(in columns)");
            std::size_t temp = 0;
            if (g.weighted)
            {
                g.in_w.alloc(E * 4 + 4, "hipMalloc(in weights)");
                hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, keys_in.as<unsigned long long>(),
                                                             keys_out.as<unsigned long long>(),
                                                             g.w.as<std::uint32_t>(), g.in_w.as<std::uint32_t>(), E,
                                                             0, 32 + vbits, s),
                          "radix sort size");
                DeviceBuffer tmp(temp + 16, "hipMalloc(sort temp)");
                hip_check(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, temp, keys_in.as<unsigned long long>(),
                                                             keys_out.as<unsigned long long>(),
                                                             g.w.as<std::uint32_t>(), g.in_w.as<std::uint32_t>(), E,
                                                             0, 32 + vbits, s),
                          "radix sort");
                hip_check(hipStreamSynchronize(s), "transpose sort");
            }
            else
            {
                hip_check(hipcub::DeviceRadixSort::SortKeys(nullptr, temp, keys_in.as<unsigned long long>(),
                                                            keys_out.as<unsigned long long>(), E, 0, 32 + vbits, s),
                          "radix sort size");
                DeviceBuffer tmp(temp + 16, "hipMalloc(sort temp)");
                hip_check(hipcub::DeviceRadixSort::SortKeys(tmp.ptr, temp, keys_in.as<unsigned long long>(),
                                                            keys_out.as<unsigned long long>(), E, 0, 32 + vbits, s),
                          "radix sort");
                hip_check(hipStreamSynchronize(s), "transpose sort");
            }
            row_starts<OffT><<<grid_for(E > V ? E : V + 1), kBlock, 0, s>>>(keys_out.as<unsigned long long>(), E, V,
                                                                          g.in_off.as<OffT>(),
                                                                          g.in_col.as<std::uint32_t>());
            hip_check(hipGetLastError(), "launch row_starts");
            hip_check(hipStreamSynchronize(s), "transpose");
        }

        template <class OffT>
        void heavy_list_typed(DeviceGraph &g, std::uint32_t light_max, std::uint32_t chunk)
        {
            const std::uint64_t cap = g.E / chunk + g.E / (static_cast<std::uint64_t>(light_max) + 1) + 64;
            g.in_heavy.alloc(cap * sizeof(HeavyChunk), "hipMalloc(in heavy chunks)");
            Counters *d_ctr = g.ws.ctr.as<Counters>();
            hip_check(hipMemsetAsync(d_ctr, 0, sizeof(Counters), g.stream), "reset counters");
            heavy_rows<OffT><<<blocks_for(g.V), kBlock, 0, g.stream>>>(g.in_off.as<OffT>(), g.V, light_max, chunk,
                                                                        g.in_heavy.as<HeavyChunk>(), cap, d_ctr);
            hip_check(hipGetLastError(), "launch heavy_rows");
            Counters h{};
            hip_check(hipMemcpyAsync(&h, d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, g.stream), "read counters");
            hip_check(hipStreamSynchronize(g.stream), "heavy rows");
            if (h.heavy_overflow)
                fail(HYT_ERR_INTERNAL, "in-heavy chunk list overflow");
            g.in_heavy_count = h.heavy_count;
            g.in_heavy_light_max = light_max;
            g.in_heavy_chunk = chunk;
        }
    } // namespace

    void build_incoming_csr(DeviceGraph &g)
    {
        if (g.off64)
            build_typed<std::uint64_t>(g);
        else
            build_typed<std::uint32_t>(g);
    }

    void ensure_transpose(DeviceGraph &g, std::uint32_t light_max, std::uint32_t chunk)
    {
        if (!g.in_col.ptr)
        {
            if (g.off64)
                build_typed<std::uint64_t>(g);
            else
                build_typed<std::uint32_t>(g);
        }
        if (!g.in_heavy.ptr || g.in_heavy_light_max != light_max || g.in_heavy_chunk != chunk)
        {
            if (g.off64)
                heavy_list_typed<std::uint64_t>(g, light_max, chunk);
            else
                heavy_list_typed<std::uint32_t>(g, light_max, chunk);
        }
    }
} // namespace hyt
