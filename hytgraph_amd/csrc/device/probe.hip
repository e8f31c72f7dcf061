// hyt_stream_probe: a read-only streaming kernel over a device buffer of
// known size - the reference for the rocprofv3 byte counters (FETCH_SIZE of
// a launch that reads exactly N bytes) and for the HBM read rate a plain
// coalesced stream reaches on this device.
#include "dgraph.hpp"
#include "kern_common.hpp"

namespace hyt
{
    namespace
    {
        // 16-byte loads, grid-stride; the XOR is kept live by a store that
        // never happens for the fill value used.
        __global__ __launch_bounds__(kBlock) void stream_read(const uint4 *src, std::uint64_t n, unsigned *sink)
        {
            unsigned acc = 0;
            const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
            for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
                 i += stride)
            {
                const uint4 v = src[i];
                acc ^= v.x ^ v.y ^ v.z ^ v.w;
            }
            if (acc == 0x9E3779B9U)
                sink[0] = acc;
        }
    } // namespace
} // namespace hyt

extern "C" int hyt_stream_probe(uint64_t bytes, uint32_t reps, double *gbs)
{
    return hyt::guarded("hyt_stream_probe", [&] {
        using namespace hyt;
        require(bytes >= 16 && bytes % 16 == 0, "bytes must be a positive multiple of 16");
        require(reps >= 1, "reps must be >= 1");
        require_device();
        hipStream_t s = nullptr;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        struct StreamGuard
        {
            hipStream_t s;
            hipEvent_t a = nullptr, b = nullptr;
            ~StreamGuard()
            {
                if (a)
                    (void)hipEventDestroy(a);
                if (b)
                    (void)hipEventDestroy(b);
                (void)hipStreamDestroy(s);
            }
        } guard{s};
        DeviceBuffer buf(bytes, "hipMalloc(stream probe)");
        DeviceBuffer sink(64, "hipMalloc(stream sink)");
        launch_fill<std::uint32_t>(buf.as<std::uint32_t>(), bytes / 4, 0x01010101U, s);
        hip_check(hipEventCreate(&guard.a), "hipEventCreate");
        hip_check(hipEventCreate(&guard.b), "hipEventCreate");
        const std::uint64_t n = bytes / 16;
        const unsigned grid = 8192; // 32 workgroups per CU
        stream_read<<<grid, kBlock, 0, s>>>(buf.as<uint4>(), n, sink.as<unsigned>()); // warm-up
        hip_check(hipGetLastError(), "launch stream_read");
        hip_check(hipEventRecord(guard.a, s), "hipEventRecord");
        for (std::uint32_t r = 0; r < reps; ++r)
            stream_read<<<grid, kBlock, 0, s>>>(buf.as<uint4>(), n, sink.as<unsigned>());
        hip_check(hipGetLastError(), "launch stream_read");
        hip_check(hipEventRecord(guard.b, s), "hipEventRecord");
        hip_check(hipEventSynchronize(guard.b), "stream probe");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, guard.a, guard.b), "hipEventElapsedTime");
        if (gbs)
            *gbs = static_cast<double>(bytes) * reps / (ms * 1e-3) / 1e9;
    });
}
