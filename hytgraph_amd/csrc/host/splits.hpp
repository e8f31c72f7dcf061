// Default degree-bin split of single-source SSSP push rounds: rows up to
// light_max edges are walked by the light kernel (64 rows per wavefront),
// longer rows in chunk-edge pieces by the chunk kernel, each chunk relaxing
// with its row's live distance. Autotuned on one MI355X with 8
// degree-weighted sources per graph, async rounds (ms per query):
//   RMAT 24 (E = 263 M): 256 / 512 8.15, 128 / 256 8.23, 64 / 256 8.25,
//     512 / 1024 9.26, 1024 / 2048 9.44;
//   RMAT 26 (E = 1.07 B): 64 / 256 37.28, 128 / 512 37.30, 256 / 512 37.74,
//     512 / 1024 39.00, 1024 / 2048 41.59;
//   ER 24: every split within 0.2%.
// 256 / 512 is within 1.3% of the best everywhere measured.
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
