// Per-vertex contributions on the device: pagerank_contributions (original
// src/algorithms/pagerank.cpp:165-243) and sssp_contributions
// (src/algorithms/sssp.cpp:236-290), with the original's float arithmetic
// and summation orders, so the values equal the original's bit for bit.
//
// PageRank: contribution[v] = |base + dangling_share + d * in_sum(v) - x[v]|
// with base = (1 - d) / n, dangling_share = d * dangling / n, dangling the
// float sum of x over the vertices without out-edges in vertex order, and
// in_sum(v) the float sum of x[u] / outdeg(u) over v's in-edges in the
// original's incoming-CSR order (sources ascending, parallel edges in edge
// order) - the order of the device transpose. Each operation is a separately
// rounded IEEE float op (contraction into fused multiply-adds is switched
// off). One thread per destination adds its in-edges in that order; the dangling sum runs in one wavefront, which
// adds the dangling values of each 64-vertex block in vertex order.
// SSSP: elementwise, unit priority for a vertex that becomes reachable.
#include "dgraph.hpp"
#include "kern_common.hpp"

#include <cmath>

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

        // x[u] / (float)outdeg(u) for rows with out-edges
        template <class OffT>
        __global__ void source_shares(const OffT *off, const float *x, std::uint64_t V, float *share)
        {
            std::uint64_t u = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; u < V; u += stride)
            {
                const std::uint64_t deg = static_cast<std::uint64_t>(off[u + 1] - off[u]);
                share[u] = deg ? x[u] / __ull2float_rn(deg) : 0.0F;
            }
        }

        // Launched as one wavefront: the float sum of x over the dangling
        // vertices, added one at a time in vertex order (every lane carries
        // the same running sum).
        template <class OffT>
        __global__ __launch_bounds__(kWave) void dangling_sum(const OffT *off, const float *x, std::uint64_t V,
                                                              float *out)
        {
            const int lane = lane_id();
            float sum = 0.0F;
            for (std::uint64_t base = 0; base < V; base += kWave)
            {
                const std::uint64_t v = base + static_cast<std::uint64_t>(lane);
                const bool dangling = v < V && off[v + 1] == off[v];
                const float val = dangling ? x[v] : 0.0F;
                std::uint64_t mask = __ballot(dangling);
                while (mask)
                {
                    const int j = __ffsll(static_cast<unsigned long long>(mask)) - 1;
                    sum += __shfl(val, j, kWave);
                    mask &= mask - 1;
                }
            }
            if (lane == 0)
                *out = sum;
        }

        template <class OffT>
        __global__ void pagerank_contrib(const OffT *in_off, const std::uint32_t *in_col, const float *share,
                                         const float *x, const float *dangling, std::uint64_t V, float damping,
                                         double *out)
        {
            // separately rounded operations: the pragma keeps the compiler
            // from fusing d * sum into the following add (it contracts by
            // default, also inside the __f*_rn helpers, so plain operators
            // are used under the pragma; float division is correctly rounded)
#pragma clang fp contract(off)
            std::uint64_t v = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            const float n = __ull2float_rn(V);
            const float base = (1.0F - damping) / n;
            const float dshare = damping * *dangling / n;
            const float head = base + dshare;
            for (; v < V; v += stride)
            {
                float sum = 0.0F;
                const std::uint64_t e = in_off[v + 1];
                for (std::uint64_t p = in_off[v]; p < e; ++p)
                    sum += share[in_col[p]];
                const float prod = damping * sum;
                const float next = head + prod;
                out[v] = static_cast<double>(fabsf(next - x[v]));
            }
        }

        __global__ void sssp_contrib(const float *cur, const float *cand, std::uint64_t n, float tol, double *out)
        {
            std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (; i < n; i += stride)
            {
                const float c = cur[i], d = cand[i];
                double value = 0.0;
                if (isfinite(c) && isfinite(d))
                {
                    const float gain = c - d;
                    if (gain > tol)
                        value = static_cast<double>(gain);
                }
                else if (!isfinite(c) && isfinite(d))
                    value = 1.0;
                out[i] = value;
            }
        }

        template <class OffT>
        void pagerank_contrib_typed(DeviceGraph &g, const float *x, float damping, double *out)
        {
            hipStream_t s = g.stream;
            const std::uint64_t V = g.V;
            DeviceBuffer share((V + 1) * 4, "hipMalloc(shares)"), dang(4, "hipMalloc(dangling sum)");
            source_shares<OffT><<<grid_for(V), kBlock, 0, s>>>(g.off.as<OffT>(), x, V, share.as<float>());
            hip_check(hipGetLastError(), "launch source_shares");
            dangling_sum<OffT><<<1, kWave, 0, s>>>(g.off.as<OffT>(), x, V, dang.as<float>());
            hip_check(hipGetLastError(), "launch dangling_sum");
            pagerank_contrib<OffT><<<grid_for(V), kBlock, 0, s>>>(g.in_off.as<OffT>(), g.in_col.as<std::uint32_t>(),
                                                                 share.as<float>(), x, dang.as<float>(), V, damping,
                                                                 out);
            hip_check(hipGetLastError(), "launch pagerank_contrib");
            hip_check(hipStreamSynchronize(s), "pagerank contributions");
        }
    } // namespace
} // namespace hyt

extern "C" {

int hyt_pagerank_contributions(hyt_graph_t g, const float *current_ranks, float damping, double *contributions,
                               uint32_t on_device)
{
    return hyt::guarded("hyt_pagerank_contributions", [&] {
        using namespace hyt;
        require(g != nullptr && current_ranks != nullptr && contributions != nullptr, "null argument");
        if (!(damping >= 0.0F && damping < 1.0F))
            fail(HYT_ERR_INVALID_ARGUMENT, "PageRank damping_factor must be in [0, 1)");
        DeviceGraph &dg = g->g;
        hip_check(hipSetDevice(dg.device), "hipSetDevice");
        // the transpose with PageRank's own split (no rebuild for later runs)
        ensure_workspace(dg, kPagerankLightMaxDegree, kPagerankHeavyChunk);
        ensure_transpose(dg, kPagerankLightMaxDegree, kPagerankHeavyChunk);
        const std::uint64_t V = dg.V;
        DeviceBuffer x_dev, out_dev;
        const float *x = current_ranks;
        double *out = contributions;
        if (!on_device)
        {
            x_dev.alloc(V * 4, "hipMalloc(ranks)");
            out_dev.alloc(V * 8, "hipMalloc(contributions)");
            hip_check(hipMemcpyAsync(x_dev.ptr, current_ranks, V * 4, hipMemcpyHostToDevice, dg.stream), "copy ranks");
            x = x_dev.as<float>();
            out = out_dev.as<double>();
        }
        if (dg.off64)
            pagerank_contrib_typed<std::uint64_t>(dg, x, damping, out);
        else
            pagerank_contrib_typed<std::uint32_t>(dg, x, damping, out);
        if (!on_device)
        {
            hip_check(hipMemcpyAsync(contributions, out, V * 8, hipMemcpyDeviceToHost, dg.stream), "copy out");
            hip_check(hipStreamSynchronize(dg.stream), "pagerank contributions");
        }
    });
}

int hyt_sssp_contributions(const float *current, const float *candidate, uint64_t n, float tolerance,
                           double *contributions, uint32_t on_device)
{
    return hyt::guarded("hyt_sssp_contributions", [&] {
        using namespace hyt;
        if (!(tolerance >= 0.0F) || !std::isfinite(tolerance))
            fail(HYT_ERR_INVALID_ARGUMENT, "SSSP contribution tolerance must be finite and non-negative");
        if (n == 0)
            return;
        require(current != nullptr && candidate != nullptr && contributions != nullptr, "null argument");
        require_device();
        hipStream_t s = nullptr;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        struct StreamGuard
        {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard{s};
        DeviceBuffer a, b, o;
        const float *cur = current, *cand = candidate;
        double *out = contributions;
        if (!on_device)
        {
            a.alloc(n * 4, "hipMalloc(current)");
            b.alloc(n * 4, "hipMalloc(candidate)");
            o.alloc(n * 8, "hipMalloc(contributions)");
            hip_check(hipMemcpyAsync(a.ptr, current, n * 4, hipMemcpyHostToDevice, s), "copy current");
            hip_check(hipMemcpyAsync(b.ptr, candidate, n * 4, hipMemcpyHostToDevice, s), "copy candidate");
            cur = a.as<float>();
            cand = b.as<float>();
            out = o.as<double>();
        }
        sssp_contrib<<<grid_for(n), kBlock, 0, s>>>(cur, cand, n, tolerance, out);
        hip_check(hipGetLastError(), "launch sssp_contrib");
        if (!on_device)
            hip_check(hipMemcpyAsync(contributions, out, n * 8, hipMemcpyDeviceToHost, s), "copy out");
        hip_check(hipStreamSynchronize(s), "sssp contributions");
    });
}

} // extern "C"
