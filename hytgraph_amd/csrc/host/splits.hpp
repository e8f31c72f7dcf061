// Default degree-bin split of single-source SSSP push rounds: rows up to
// light_max edges are walked by the light kernel (64 rows per wavefront),
// longer rows in chunk-edge pieces by the chunk kernel, each chunk relaxing
// with its row's live distance, chunks before light rows. Autotuned on one
// MI355X with 8 degree-weighted sources per graph, async rounds (ms per
// query):
//   RMAT 24 (E = 263 M): 256 / 512 8.05, 128 / 256 8.17, 512 / 1024 8.42,
//     128 / 512 8.40, 64 / 256 8.55;
//   RMAT 26 (E = 1.07 B): 512 / 1024 35.05, 256 / 512 36.32,
//     128 / 512 36.67, 64 / 256 39.48;
//   ER 24: every split within 0.2%.
// 256 / 512: the best at RMAT 24, within 4% at RMAT 26.
#pragma once

#include <cstdint>

namespace hyt
{
    inline void default_sssp_split(std::uint64_t /*num_edges*/, std::uint32_t &light_max, std::uint32_t &chunk)
    {
        light_max = 256;
        chunk = 512;
    }
} // namespace hyt
