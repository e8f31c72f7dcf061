// The generator's graphs built on the device (hyt_generate_device): the same
// CSR, bit for bit, as the host generator (generator.cpp) with the same
// parameters - row v holds the sorted distinct destinations of the edges
// generated from v (self-loops dropped), and the weight at final position p
// is a pure function of (seed, p) - in seconds instead of minutes at RMAT
// scale 30 (17 B generated edges), with the result in host memory of the
// requested kind.
//
// 1. histogram: every generated edge counts in the bin of its source's top
//    B bits (B = min(scale, 14)); a block keeps its bins in LDS. Self-loops
//    are counted too, so bin sums bound the edges of any source range.
// 2. the bins are cut into source-range buckets of at most `cap` edges
//    (cap from the HBM budget, 16 bytes per edge);
// 3. per bucket, in source order: every edge is generated again - RMAT
//    draws its quadrants from the top bit down, so an edge stops as soon as
//    its source prefix falls outside the bucket - and the edges of the bucket
//    are appended as 64-bit keys (row in bucket << scale | destination), one
//    counter atomic per workgroup and step; a radix sort over the key bits
//    and a unique pass leave each row's sorted distinct destinations; row
//    offsets come from a binary search per row, columns and weights from the
//    keys and final positions, all copied to the host arrays.
#include "common.hpp"
#include "host_graph.hpp"
#include "kern_common.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace hyt
{
    namespace
    {
        constexpr std::uint32_t kMaxBinBits = 14;
        constexpr int kPerThread = 8; // edges a thread draws per step
        constexpr std::uint64_t kMaxBucket = 1ULL << 30;

        struct GenDevice
        {
            std::uint64_t key;  // the edge stream's key (generator.cpp EdgeSource)
            std::uint32_t thr_a, thr_ab, thr_abc;
            std::uint32_t scale;
            bool er;
        };

        // Edge i's source and destination, or false when it is a self-loop
        // or its source falls outside [lo, hi) (RMAT stops drawing as soon as
        // the source bits drawn so far rule the range out). Same draws as the
        // host generator: word w of edge i is mix64(key + i c1 + w c2), its
        // low and high halves pick the quadrants of levels 2w and 2w + 1 from
        // the top bit down.
        __device__ __forceinline__ bool edge_in(const GenDevice &g, std::uint64_t i, std::uint64_t lo,
                                                std::uint64_t hi, std::uint32_t &src, std::uint32_t &dst)
        {
            const std::uint64_t base = g.key + i * 0xD1B54A32D192ED03ULL;
            if (g.er)
            {
                const std::uint64_t r = mix64(base);
                const std::uint64_t V = 1ULL << g.scale;
                src = static_cast<std::uint32_t>(bounded32(static_cast<std::uint32_t>(r), V));
                dst = static_cast<std::uint32_t>(bounded32(static_cast<std::uint32_t>(r >> 32), V));
                return src != dst && src >= lo && src < hi;
            }
            std::uint32_t s = 0, d = 0;
            int bit = static_cast<int>(g.scale) - 1;
            for (std::uint32_t w = 0; 2 * w < g.scale; ++w)
            {
                const std::uint64_t r = mix64(base + w * 0x8CB92BA72F3D8DD7ULL);
#pragma unroll
                for (int h = 0; h < 2; ++h)
                {
                    if (bit < 0)
                        break;
                    const std::uint32_t u = static_cast<std::uint32_t>(h == 0 ? r : (r >> 32));
                    const std::uint32_t ga = u >= g.thr_a, gab = u >= g.thr_ab, gabc = u >= g.thr_abc;
                    s |= gab << bit;
                    d |= (ga ^ gab ^ gabc) << bit;
                    --bit;
                }
                // sources still possible: [s, s + 2^(bit+1))
                const std::uint64_t smin = s;
                const std::uint64_t smax = static_cast<std::uint64_t>(s) + ((1ULL << (bit + 1)) - 1ULL);
                if (smax < lo || smin >= hi)
                    return false;
            }
            src = s;
            dst = d;
            return s != d;
        }

        // Source of edge i, self-loops included (histogram pass: only the top
        // bin bits are drawn).
        __device__ __forceinline__ std::uint32_t source_bin(const GenDevice &g, std::uint64_t i, std::uint32_t bin_bits)
        {
            const std::uint64_t base = g.key + i * 0xD1B54A32D192ED03ULL;
            if (g.er)
            {
                const std::uint64_t r = mix64(base);
                const std::uint64_t s = bounded32(static_cast<std::uint32_t>(r), 1ULL << g.scale);
                return static_cast<std::uint32_t>(s >> (g.scale - bin_bits));
            }
            std::uint32_t b = 0;
            std::uint32_t drawn = 0;
            for (std::uint32_t w = 0; drawn < bin_bits; ++w)
            {
                const std::uint64_t r = mix64(base + w * 0x8CB92BA72F3D8DD7ULL);
                for (int h = 0; h < 2 && drawn < bin_bits; ++h)
                {
                    const std::uint32_t u = static_cast<std::uint32_t>(h == 0 ? r : (r >> 32));
                    b = (b << 1) | (u >= g.thr_ab ? 1U : 0U);
                    ++drawn;
                }
            }
            return b;
        }

        __global__ __launch_bounds__(kBlock) void gen_histogram(GenDevice g, std::uint64_t n, std::uint32_t bin_bits,
                                                                unsigned long long *bins)
        {
            extern __shared__ std::uint32_t s_bins[];
            const std::uint32_t nb = 1U << bin_bits;
            for (std::uint32_t b = threadIdx.x; b < nb; b += kBlock)
                s_bins[b] = 0;
            __syncthreads();
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * kBlock;
            for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
                atomicAdd(&s_bins[source_bin(g, i, bin_bits)], 1U);
            __syncthreads();
            for (std::uint32_t b = threadIdx.x; b < nb; b += kBlock)
                if (s_bins[b])
                    atomicAdd(&bins[b], static_cast<unsigned long long>(s_bins[b]));
        }

        // The keys of the edges with a source in [lo, hi): each thread draws
        // kPerThread consecutive edges per step; one atomic per workgroup
        // and step reserves the step's slots.
        __global__ __launch_bounds__(kBlock) void gen_bucket(GenDevice g, std::uint64_t n, std::uint64_t lo,
                                                             std::uint64_t hi, std::uint64_t *keys,
                                                             std::uint64_t cap, unsigned long long *count,
                                                             std::uint32_t *overflow)
        {
            __shared__ std::uint32_t s_w[kWavesPerBlock];
            __shared__ unsigned long long s_base;
            const int lane = lane_id();
            const int wid = static_cast<int>(threadIdx.x) / kWave;
            const std::uint64_t per_step = static_cast<std::uint64_t>(kBlock) * kPerThread;
            const std::uint64_t steps = (n + per_step - 1) / per_step;
            for (std::uint64_t st = blockIdx.x; st < steps; st += gridDim.x)
            {
                const std::uint64_t i0 = st * per_step + static_cast<std::uint64_t>(threadIdx.x) * kPerThread;
                std::uint64_t k[kPerThread];
                bool in[kPerThread];
                std::uint32_t c = 0;
#pragma unroll
                for (int j = 0; j < kPerThread; ++j)
                {
                    std::uint32_t s = 0, d = 0;
                    in[j] = i0 + j < n && edge_in(g, i0 + j, lo, hi, s, d);
                    k[j] = (static_cast<std::uint64_t>(s - static_cast<std::uint32_t>(lo)) << g.scale) | d;
                    c += in[j] ? 1U : 0U;
                }
                const std::uint32_t incl = wave_incl_scan(c);
                if (lane == kWave - 1)
                    s_w[wid] = incl;
                __syncthreads();
                std::uint32_t before = 0, total = 0;
#pragma unroll
                for (int w = 0; w < kWavesPerBlock; ++w)
                {
                    if (w < wid)
                        before += s_w[w];
                    total += s_w[w];
                }
                if (threadIdx.x == 0)
                    s_base = total ? atomicAdd(count, static_cast<unsigned long long>(total)) : 0ULL;
                __syncthreads();
                std::uint64_t at = s_base + before + incl - c;
#pragma unroll
                for (int j = 0; j < kPerThread; ++j)
                    if (in[j])
                    {
                        if (at < cap)
                            keys[at] = k[j];
                        else
                            atomicOr(overflow, 1U);
                        ++at;
                    }
                __syncthreads(); // s_w and s_base are rewritten by the next step
            }
        }

        // off[r] = first position of a key of row r in the sorted distinct
        // keys (binary search), plus the edges of earlier buckets.
        __global__ void gen_offsets(const std::uint64_t *keys, std::uint64_t n, std::uint64_t rows,
                                    std::uint32_t scale, std::uint64_t before, std::uint64_t *off)
        {
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (std::uint64_t r = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < rows;
                 r += stride)
            {
                const std::uint64_t target = r << scale;
                std::uint64_t lo = 0, hi = n; // first index with keys[i] >= target
                while (lo < hi)
                {
                    const std::uint64_t mid = (lo + hi) >> 1;
                    if (keys[mid] < target)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                off[r] = before + lo;
            }
        }

        __global__ void gen_columns(const std::uint64_t *keys, std::uint64_t n, std::uint64_t mask, std::uint32_t *col)
        {
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
                col[i] = static_cast<std::uint32_t>(keys[i] & mask);
        }

        // Weight of final position before + i: weight_min + the high half of
        // draw(seed, weights, position, 0) mapped to the range (generator.cpp).
        __global__ void gen_weights(std::uint64_t wkey, std::uint64_t n, std::uint64_t before, std::uint32_t wmin,
                                    std::uint64_t range, std::uint32_t *w)
        {
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
            {
                const std::uint64_t r = mix64(wkey + (before + i) * 0xD1B54A32D192ED03ULL);
                w[i] = wmin + static_cast<std::uint32_t>(bounded32(static_cast<std::uint32_t>(r >> 32), range));
            }
        }

        unsigned grid_for(std::uint64_t n)
        {
            const std::uint64_t b = (n + kBlock - 1) / kBlock;
            return static_cast<unsigned>(std::max<std::uint64_t>(1, std::min<std::uint64_t>(b, 65536)));
        }

        std::uint32_t threshold(double x)
        {
            double v = std::floor(x * 4294967296.0);
            v = std::max(0.0, std::min(4294967295.0, v));
            return static_cast<std::uint32_t>(v);
        }
    } // namespace

    std::unique_ptr<HostGraph> generate_device(const hyt_gen_params &p, std::uint64_t budget)
    {
        require(p.kind == HYT_GEN_RMAT || p.kind == HYT_GEN_ER, "unknown generator kind");
        require(p.scale >= 1 && p.scale <= 32, "scale must be in [1, 32]");
        require(p.edge_factor >= 1 && p.edge_factor <= 1024, "edge_factor must be in [1, 1024]");
        if (p.kind == HYT_GEN_RMAT)
            require(p.a >= 0 && p.b >= 0 && p.c >= 0 && p.a + p.b + p.c <= 1.0 + 1e-12,
                    "RMAT probabilities must be non-negative and a+b+c <= 1");
        require(p.weight_min <= p.weight_max, "weight_min must not exceed weight_max");
        require(p.memory == HYT_HOST_PAGEABLE || p.memory == HYT_HOST_PINNED, "unknown memory kind");
        const int dev = require_device();
        (void)dev;
        const bool progress = std::getenv("HYT_GEN_PROGRESS") != nullptr;
        const auto t0 = std::chrono::steady_clock::now();
        auto say = [&](const std::string &what) {
            if (!progress)
                return;
            const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::fprintf(stderr, "[generator/device] %s at %.1f s\n", what.c_str(), s);
            std::fflush(stderr);
        };

        GenDevice g{};
        g.key = mix64(p.seed ^ mix64(kStreamEdges + 0x632BE59BD9B4E019ULL));
        g.thr_a = threshold(p.a);
        g.thr_ab = threshold(p.a + p.b);
        g.thr_abc = threshold(p.a + p.b + p.c);
        g.scale = p.scale;
        g.er = p.kind == HYT_GEN_ER;
        const std::uint64_t V = 1ULL << p.scale;
        const std::uint64_t n_gen = static_cast<std::uint64_t>(p.edge_factor) * V;
        const std::uint32_t bin_bits = std::min<std::uint32_t>(p.scale, kMaxBinBits);
        const std::uint64_t nbins = 1ULL << bin_bits;

        hipStream_t s = nullptr;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        struct StreamGuard
        {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } sg{s};

        // 1. histogram of generated edges by source bin
        DeviceBuffer bins_d(nbins * 8, "hipMalloc(generator bins)");
        hip_check(hipMemsetAsync(bins_d.ptr, 0, nbins * 8, s), "clear bins");
        gen_histogram<<<2048, kBlock, nbins * 4, s>>>(g, n_gen, bin_bits, bins_d.as<unsigned long long>());
        hip_check(hipGetLastError(), "launch gen_histogram");
        std::vector<unsigned long long> bins(nbins);
        hip_check(hipMemcpyAsync(bins.data(), bins_d.ptr, nbins * 8, hipMemcpyDeviceToHost, s), "read bins");
        hip_check(hipStreamSynchronize(s), "generator histogram");
        std::uint64_t n_raw = 0;
        for (auto b : bins)
            n_raw += b;
        say("histogram: " + std::to_string(n_raw) + " edges");

        // 2. buckets of whole bins, at most cap edges each
        const std::uint64_t cap_budget = (budget ? budget : (40ULL << 30)) / 17;
        const std::uint64_t cap = std::max<std::uint64_t>(1, std::min<std::uint64_t>(cap_budget, kMaxBucket));
        std::vector<std::uint64_t> cut{0}; // first bin of each bucket
        std::vector<std::uint64_t> bucket_edges;
        std::uint64_t run = 0;
        for (std::uint64_t b = 0; b < nbins; ++b)
        {
            if (bins[b] > cap)
                fail(HYT_ERR_OUT_OF_MEMORY, "generator: one source bin holds " + std::to_string(bins[b]) +
                                                " edges, above the bucket capacity " + std::to_string(cap) +
                                                " of the device budget");
            if (run + bins[b] > cap)
            {
                cut.push_back(b);
                bucket_edges.push_back(run);
                run = 0;
            }
            run += bins[b];
        }
        cut.push_back(nbins);
        bucket_edges.push_back(run);
        std::uint64_t largest = 0;
        for (auto e : bucket_edges)
            largest = std::max(largest, e);
        const std::uint64_t bcap = std::max<std::uint64_t>(largest, 1);

        // 3. the host CSR (columns and weights sized for the generated edges;
        // the final count is at most that)
        auto out = std::make_unique<HostGraph>(V, n_raw, p.weight_max != 0, p.memory);
        DeviceBuffer keys_a(bcap * 8 + 64, "hipMalloc(generator keys)");
        DeviceBuffer keys_b(bcap * 8 + 64, "hipMalloc(generator keys)");
        DeviceBuffer ctl(256, "hipMalloc(generator counters)");
        std::uint64_t rows_cap = 0; // offsets staging: the rows of the largest bucket
        for (std::size_t k = 0; k + 1 < cut.size(); ++k)
            rows_cap = std::max(rows_cap, (cut[k + 1] - cut[k]) << (p.scale - bin_bits));
        DeviceBuffer off_d(std::max<std::uint64_t>(rows_cap, 1) * 8, "hipMalloc(generator offsets)");
        std::size_t sort_bytes = 0, uniq_bytes = 0;
        hip_check(hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, keys_a.as<std::uint64_t>(),
                                                    keys_b.as<std::uint64_t>(), static_cast<int>(bcap), 0, 64, s),
                  "radix sort temp size");
        hip_check(hipcub::DeviceSelect::Unique(nullptr, uniq_bytes, keys_b.as<std::uint64_t>(),
                                               keys_a.as<std::uint64_t>(),
                                               reinterpret_cast<unsigned long long *>(ctl.ptr),
                                               static_cast<std::int64_t>(bcap), s),
                  "unique temp size");
        DeviceBuffer temp(std::max(sort_bytes, uniq_bytes) + 256, "hipMalloc(generator sort temp)");
        auto *count = reinterpret_cast<unsigned long long *>(ctl.ptr);
        auto *uniq_n = count + 1;
        auto *overflow = reinterpret_cast<std::uint32_t *>(count + 2);
        const std::uint64_t wkey = mix64(p.seed ^ mix64(kStreamWeights + 0x632BE59BD9B4E019ULL));
        const std::uint64_t wrange = static_cast<std::uint64_t>(p.weight_max) - p.weight_min + 1;
        const std::uint64_t mask = p.scale >= 32 ? 0xFFFFFFFFULL : ((1ULL << p.scale) - 1ULL);

        std::uint64_t E = 0;
        for (std::size_t k = 0; k + 1 < cut.size(); ++k)
        {
            const std::uint64_t lo = cut[k] << (p.scale - bin_bits);
            const std::uint64_t hi = cut[k + 1] << (p.scale - bin_bits);
            const std::uint64_t rows = hi - lo;
            hip_check(hipMemsetAsync(ctl.ptr, 0, 256, s), "clear generator counters");
            gen_bucket<<<4096, kBlock, 0, s>>>(g, n_gen, lo, hi, keys_a.as<std::uint64_t>(), bcap, count, overflow);
            hip_check(hipGetLastError(), "launch gen_bucket");
            unsigned long long h[3] = {0, 0, 0};
            hip_check(hipMemcpyAsync(h, ctl.ptr, sizeof(h), hipMemcpyDeviceToHost, s), "read bucket count");
            hip_check(hipStreamSynchronize(s), "generator bucket");
            require(static_cast<std::uint32_t>(h[2]) == 0 && h[0] <= bcap, "generator bucket overflow");
            const std::uint64_t n = h[0];
            // sort the keys over their bits: row-in-bucket above the destination
            int end_bit = static_cast<int>(p.scale);
            while (end_bit < 64 && (rows - 1) >> (end_bit - p.scale) != 0)
                ++end_bit;
            std::size_t tb = temp.bytes;
            if (n > 0)
            {
                hip_check(hipcub::DeviceRadixSort::SortKeys(temp.ptr, tb, keys_a.as<std::uint64_t>(),
                                                            keys_b.as<std::uint64_t>(), static_cast<int>(n), 0,
                                                            end_bit, s),
                          "radix sort bucket");
                tb = temp.bytes;
                hip_check(hipcub::DeviceSelect::Unique(temp.ptr, tb, keys_b.as<std::uint64_t>(),
                                                       keys_a.as<std::uint64_t>(), uniq_n,
                                                       static_cast<std::int64_t>(n), s),
                          "unique bucket");
            }
            hip_check(hipMemcpyAsync(h, ctl.ptr, sizeof(h), hipMemcpyDeviceToHost, s), "read unique count");
            hip_check(hipStreamSynchronize(s), "generator sort");
            const std::uint64_t u = n > 0 ? h[1] : 0;
            // offsets, columns and weights of the bucket
            gen_offsets<<<grid_for(rows), kBlock, 0, s>>>(keys_a.as<std::uint64_t>(), u, rows, p.scale, E,
                                                          off_d.as<std::uint64_t>());
            hip_check(hipGetLastError(), "launch gen_offsets");
            hip_check(hipMemcpyAsync(out->off + lo, off_d.ptr, rows * 8, hipMemcpyDeviceToHost, s), "copy offsets");
            auto *col_d = keys_b.as<std::uint32_t>();
            auto *w_d = col_d + u;
            if (u > 0)
            {
                gen_columns<<<grid_for(u), kBlock, 0, s>>>(keys_a.as<std::uint64_t>(), u, mask, col_d);
                hip_check(hipGetLastError(), "launch gen_columns");
                hip_check(hipMemcpyAsync(out->col + E, col_d, u * 4, hipMemcpyDeviceToHost, s), "copy columns");
                if (out->weighted)
                {
                    gen_weights<<<grid_for(u), kBlock, 0, s>>>(wkey, u, E, p.weight_min, wrange, w_d);
                    hip_check(hipGetLastError(), "launch gen_weights");
                    hip_check(hipMemcpyAsync(out->w + E, w_d, u * 4, hipMemcpyDeviceToHost, s), "copy weights");
                }
            }
            hip_check(hipStreamSynchronize(s), "generator copy");
            E += u;
            say("bucket " + std::to_string(k + 1) + " / " + std::to_string(cut.size() - 1) + ": " +
                std::to_string(n) + " generated, " + std::to_string(u) + " distinct");
        }
        out->off[V] = E;
        out->E = E;
        return out;
    }
} // namespace hyt

extern "C" int hyt_generate_device(const hyt_gen_params *params, uint64_t device_budget_bytes,
                                   hyt_host_graph_t *out)
{
    return hyt::guarded("hyt_generate_device", [&] {
        using namespace hyt;
        require(params != nullptr && out != nullptr, "null params or out");
        *out = nullptr;
        auto h = std::make_unique<hyt_host_graph>();
        h->g = generate_device(*params, device_budget_bytes);
        *out = h.release();
    });
}
