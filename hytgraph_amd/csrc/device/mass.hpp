// PageRank / PPR mass in unsigned 64-bit fixed point.
//
// Residuals and pushed mass (rank) are integer words, word = mass * 2^frac.
// Integer atomic adds are exact and associative, so the residual a vertex
// holds after a round does not depend on the order in which its incoming
// pushes land: a query gives the same bits alone, inside a batch or inside a
// served mix. What is left of rounding is
//   * each pushing row's per-edge share: its scale d / outdeg(u) in double,
//     times the residual word in double (relative error <= 2^-52 with both
//     roundings), rounded to the nearest word - and, for global PageRank,
//     that word rounded to float and back (2^-24): its pull rounds gather
//     4-byte shares, and a row that crosses eps in one form but not in the
//     other moves eps of mass V, nothing in the normalised ranks;
//   * the final normalisation x / |x|_1 to float (relative 2^-24 per entry).
// Every form (push or pull rounds, one query or K lanes, in HBM or through
// the hybrid engine) computes a share word with the same operations, so a
// query gives the same words in all of them.
// PPR keeps the words: its shares are then as precise as the FP64
// restatement's, so the rows that cross eps in a round are the FP64
// procedure's. With float shares a residual within that margin of eps could
// be pushed in one and not in the other, and each such flip moved about eps
// of one seed's mass (seen: L1 1.7e-4 at eps 1e-6 on a scale-13 graph). With
// unnormalised delta-push the pushed mass of a whole run is at most
// |x|_1 * d / (1 - d), so the share rounding moves the normalised L1 result
// by at most d / (1 - d) * 2^-52 (words) or d / (1 - d) * 2^-24 (float
// shares, 3.4e-7 at d = 0.85), plus half a word per push.
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

    // d / outdeg(u) in double (the same in every form; 0 for no out-edges).
    __device__ __forceinline__ double row_scale_of(float damping, std::uint64_t deg)
    {
        return deg ? __ddiv_rn(static_cast<double>(damping), static_cast<double>(deg)) : 0.0;
    }

    // Word of the share per out-edge of a pushing residual word r at the
    // row's scale: r * scale in double, rounded to the nearest word.
    __device__ __forceinline__ unsigned long long share_word(unsigned long long r, double scale)
    {
        return __double2ull_rn(__dmul_rn(static_cast<double>(r), scale));
    }
} // namespace hyt
