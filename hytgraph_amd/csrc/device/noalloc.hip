// hyt_sssp_noalloc: single-source SSSP that allocates nothing.
//
// For callers that cannot call hipMalloc (code under HIP graph capture,
// tight serving loops, harnesses that forbid allocation). The graph is
// uploaded beforehand (hyt_graph_upload) and the caller passes the device
// output; the run's scratch - distances, the expansion record, the queue and
// its values, the heavy-chunk queue and the counters - lives in `scratch`, a
// file-scope device array of kScratchBytes that every GPU the library runs on
// holds until the process ends. Host code never uses the array's name as a
// device address: the first call on a device (that device current) takes the
// array's address with hipGetSymbolAddress and keeps it per device, and the
// kernels get that pointer as an argument.
//
// Every call on one GPU shares that GPU's array and leaves the last call's
// contents in it: each call clears what must start at zero (the counters)
// and fills the distance and record arrays, and calls that overlap on one
// GPU race on the array (callers serialise them, e.g. on one stream).
// Rounds are push-only (no transpose), SYNC or ASYNC, any weight mode.
#include "dgraph.hpp"
#include "kern_common.hpp"

#include <chrono>
#include <mutex>

namespace hyt
{
    namespace
    {
        constexpr std::size_t kScratchBytes = std::size_t(128) << 20; // 128 MiB per GPU
        constexpr int kMaxDevices = 64;

        static __device__ __align__(256) unsigned char scratch[kScratchBytes];

        std::mutex g_scratch_mu;
        void *g_scratch_ptr[kMaxDevices] = {};

        // The device address of this device's `scratch` (device current).
        void *scratch_ptr(int device)
        {
            require(device >= 0 && device < kMaxDevices, "device id out of range");
            std::lock_guard<std::mutex> lock(g_scratch_mu);
            if (g_scratch_ptr[device] == nullptr)
            {
                void *p = nullptr;
                hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
                require(p != nullptr, "hipGetSymbolAddress returned no address");
                g_scratch_ptr[device] = p;
            }
            return g_scratch_ptr[device];
        }

        constexpr std::uint64_t align256(std::uint64_t n) { return (n + 255) & ~std::uint64_t(255); }

        // Scratch layout for V vertices and a heavy-chunk capacity:
        // counters | dist | prev | q | qd | heavy chunks; returns the bytes used.
        std::uint64_t layout(std::uint64_t V, std::uint64_t heavy_cap, unsigned char *base, SsspBuffers *b)
        {
            const std::uint64_t vb = align256((V + kPad) * 4);
            std::uint64_t at = 0;
            auto take = [&](std::uint64_t n) {
                unsigned char *p = base ? base + at : nullptr;
                at += align256(n);
                return p;
            };
            unsigned char *ctr = take(sizeof(Counters));
            unsigned char *dist = take(vb), *prev = take(vb), *q = take(vb), *qd = take(vb);
            unsigned char *heavy = take(heavy_cap * sizeof(HeavyChunk));
            if (b)
                *b = SsspBuffers{reinterpret_cast<std::uint32_t *>(dist),
                                 reinterpret_cast<std::uint32_t *>(prev),
                                 reinterpret_cast<std::uint32_t *>(q),
                                 reinterpret_cast<std::uint32_t *>(qd),
                                 nullptr,
                                 nullptr,
                                 reinterpret_cast<HeavyChunk *>(heavy),
                                 heavy_cap,
                                 reinterpret_cast<Counters *>(ctr)};
            return at;
        }
    } // namespace
} // namespace hyt

extern "C" int hyt_sssp_noalloc(hyt_graph_t g, uint32_t source, uint32_t *dist, const hyt_run_options *opts,
                                hyt_run_stats *stats)
{
    return hyt::guarded("hyt_sssp_noalloc", [&] {
        using namespace hyt;
        require(g != nullptr && dist != nullptr, "null graph or dist");
        DeviceGraph &dg = g->g;
        require(source < dg.V, "SSSP source vertex is out of range");
        hyt_run_options o{};
        if (opts)
            o = *opts;
        require(o.out_on_device == 1, "hyt_sssp_noalloc writes device memory only (out_on_device = 1)");
        require(o.mode == HYT_SSSP_SYNC || o.mode == HYT_SSSP_ASYNC, "unknown SSSP mode");
        require(o.tolerance == 0.0F, "hyt_sssp_noalloc runs tolerance 0 only (tolerance > 0 builds a transpose)");
        require(o.persistent == 0, "hyt_sssp_noalloc runs the captured rounds (persistent runs allocate their queues)");
        const std::uint64_t pct = o.pull_percent ? o.pull_percent : kDefaultPullPercentSssp;
        require(o.direction == HYT_DIR_PUSH || (o.direction == HYT_DIR_AUTO && pct > 100),
                "hyt_sssp_noalloc runs push rounds only (pull rounds build a transpose)");
        std::uint32_t light_max = 0, chunk = 0;
        sssp_split(dg, o, light_max, chunk);
        const std::uint64_t cap = heavy_capacity(dg.E, light_max, chunk);
        const std::uint64_t need = layout(dg.V, cap, nullptr, nullptr);
        if (need > kScratchBytes)
            fail(HYT_ERR_INVALID_ARGUMENT,
                 "graph too large for hyt_sssp_noalloc: needs " + std::to_string(need) + " bytes of scratch, " +
                     std::to_string(kScratchBytes) + " available (about V <= 2^22 at E <= 2^28); use hyt_sssp");
        hip_check(hipSetDevice(dg.device), "hipSetDevice");
        SsspBuffers b{};
        layout(dg.V, cap, static_cast<unsigned char *>(scratch_ptr(dg.device)), &b);
        hyt_run_stats st{};
        const auto t0 = std::chrono::steady_clock::now();
        try
        {
            sssp_rounds(dg, source, b, o, light_max, chunk, st);
        }
        catch (...)
        {
            if (stats)
                *stats = st;
            throw;
        }
        hip_check(hipMemcpyAsync(dist, b.dist, dg.V * 4, hipMemcpyDeviceToDevice, dg.stream), "copy distances");
        hip_check(hipStreamSynchronize(dg.stream), "sssp output");
        st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (stats)
            *stats = st;
    });
}
