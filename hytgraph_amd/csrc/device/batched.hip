// Multi-source [V][K] SSSP and personalized PageRank.
#include "dgraph.hpp"

namespace hyt
{
    void sssp_batched(DeviceGraph &, const std::uint32_t *, std::uint32_t, std::uint32_t *, const hyt_run_options &,
                      hyt_run_stats &)
    {
        fail(HYT_ERR_UNSUPPORTED, "batched SSSP (K > 1) is not available in this build");
    }

    void ppr_batched(DeviceGraph &, const std::uint32_t *, std::uint32_t, float, float, float *,
                     const hyt_run_options &, hyt_run_stats &)
    {
        fail(HYT_ERR_UNSUPPORTED, "batched PPR (K > 1) is not available in this build");
    }
} // namespace hyt
