// Default degree-bin split of single-source SSSP push rounds: rows up to
// light_max edges are walked by the light kernel (64 rows per wavefront),
// longer rows in chunk-edge pieces by the chunk kernel. Autotuned on one
// MI355X with 8 degree-weighted sources per graph, async rounds:
//   RMAT 24 (E = 263 M): 256 / 512 12.03 ms per query, 1024 / 2048 12.45,
//     2048 / 4096 13.98;
//   RMAT 26 (E = 1.07 B): 2048 / 4096 58.3, 1024 / 2048 58.4, 256 / 512 61.2;
//   ER 24: every split within 0.2%.
// Graphs below 2^29 edges take 256 / 512, larger ones 1024 / 2048.
#pragma once

#include <cstdint>

namespace hyt
{
    inline void default_sssp_split(std::uint64_t num_edges, std::uint32_t &light_max, std::uint32_t &chunk)
    {
        if (num_edges < (1ULL << 29))
        {
            light_max = 256;
            chunk = 512;
        }
        else
        {
            light_max = 1024;
            chunk = 2048;
        }
    }
} // namespace hyt
