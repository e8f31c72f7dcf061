// PageRank / PPR mass in unsigned 64-bit fixed point.
//
// Residuals and pushed mass (rank) are integer words, word = mass * 2^frac.
// Integer atomic adds are exact and associative, so the residual a vertex
// holds after a round does not depend on the order in which its incoming
// pushes land: a query gives the same bits alone, inside a batch or inside a
// served mix. What is left of rounding is
//   * each pushing row's per-edge share: its scale d / outdeg(u) rounded to
//     float, times the residual word, rounded to the float that carries it
//     through the expand kernels (relative error <= 2^-23 with both roundings),
//     then to the nearest word, and
//   * the final normalisation x / |x|_1 to float (relative 2^-24 per entry).
// Every form (push or pull rounds, one query or K lanes, in HBM or through
// the hybrid engine) computes a share word with the same operations, so a
// PPR query gives the same words in all of them.
// With unnormalised delta-push the pushed mass of a whole run is at most
// |x|_1 * d / (1 - d), so the share rounding moves the normalised L1 result
// by at most d / (1 - d) * 2^-23 (6.8e-7 at d = 0.85), against 1e-6 allowed;
// the roundings are unbiased, and the measured error is far below the bound.
//
// frac: a run seeded with n vertices of residual (1 - d) never holds more
// than n in total (each push of r adds at most d * r to the residuals and r
// to the rank, so rank + residual stays below n * (1 - d) / (1 - d)); with
// frac = 62 - ceil(log2 n) no word, and no sum of all words, passes 2^62.
// Global PageRank seeds every vertex (n = V), PPR one vertex (n = 1).
#pragma once

#include <cmath>
#include <cstdint>

namespace hyt
{
    struct MassScale
    {
        int frac = 62;               // fraction bits of a word
        double one = 0.0;            // the word of mass 1 (2^frac)
        unsigned long long eps = 0;  // a residual word r is above the threshold when r > eps
        unsigned long long r0 = 0;   // initial residual (1 - d) of a seeded vertex
    };

    inline MassScale mass_scale(std::uint64_t seeded, float damping, float eps)
    {
        int lg = 0;
        while (lg < 62 && (1ULL << lg) < seeded)
            ++lg;
        MassScale m;
        m.frac = 62 - lg;
        m.one = std::ldexp(1.0, m.frac);
        // r > eps (real) <=> word > floor(eps * 2^frac) for an integer word;
        // the product is exact (a power-of-two scaling of a float)
        const double e = static_cast<double>(eps) * m.one;
        m.eps = e >= 18446744073709551615.0 ? ~0ULL : static_cast<unsigned long long>(std::floor(e));
        m.r0 = static_cast<unsigned long long>(std::llround((1.0 - static_cast<double>(damping)) * m.one));
        return m;
    }

    // d / outdeg(u) as a float (the same in every form).
    __device__ __forceinline__ float row_scale_of(float damping, std::uint64_t deg)
    {
        return deg ? __fdiv_rn(damping, static_cast<float>(deg)) : 0.0F;
    }

    // Share per out-edge, in words, of a pushing residual word r with the
    // row's scale (float: what the expand kernels carry per row).
    __device__ __forceinline__ float share_of(unsigned long long r, float scale)
    {
        return static_cast<float>(static_cast<double>(r) * static_cast<double>(scale));
    }

    // Word of a share, rounded to nearest.
    __device__ __forceinline__ unsigned long long share_word(float share) { return __float2ull_rn(share); }
} // namespace hyt
