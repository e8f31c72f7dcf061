// The hub score of the original HubSorter (src/scheduling/hub_sort.cpp:88-110)
// is static_cast<double>(static_cast<long double>(a) / static_cast<long
// double>(b)) for integers a <= b < 2^64, evaluated on x86-64 where long
// double is the x87 extended format (64-bit significand). Both integers are
// exact in that format; the quotient is rounded to 64 significant bits
// (nearest, ties to even), and the cast rounds that value to 53 bits - a
// double rounding, which a plain double division does not reproduce (about
// one quotient in a few thousand differs in the last bit).
//
// extended_quotient computes the same two roundings with integer arithmetic
// only (restoring division, one quotient bit per step), so device code gives
// the original's scores bit for bit. Host code includes this header too (its
// CPU test compares it with the compiler's long double on random and
// adversarial pairs).
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define HYT_HOST_DEVICE __host__ __device__
#else
#define HYT_HOST_DEVICE
#endif

namespace hyt
{
    // (double)((long double)a / (long double)b) with long double = x87
    // extended, for 0 <= a <= b, 0 < b < 2^64.
    HYT_HOST_DEVICE inline double extended_quotient(std::uint64_t a, std::uint64_t b)
    {
        if (a == 0)
            return 0.0;
        if (a == b)
            return 1.0;
        // a / b < 1: its binary digits after the point, one per step; the
        // remainder r < b stays below 2^64, 2r may carry out of 64 bits
        std::uint64_t r = a;
        auto digit = [&r, b]() -> std::uint64_t {
            const bool carry = (r >> 63) != 0;
            r <<= 1;
            if (carry || r >= b)
            {
                r -= b; // modulo 2^64: the true value 2^64 + r - b is < b
                return 1;
            }
            return 0;
        };
        int zeros = 0; // digits before the leading 1 (a >= 1, b < 2^64: at most 63)
        while (digit() == 0)
            ++zeros;
        std::uint64_t sig = 1;
        for (int i = 0; i < 63; ++i)
            sig = (sig << 1) | digit();
        const std::uint64_t guard = digit();
        const bool sticky = r != 0;
        int exp = -(zeros + 1) - 63; // value = sig * 2^exp, 2^63 <= sig < 2^64
        if (guard && (sticky || (sig & 1)))
        {
            sig += 1;
            if (sig == 0)
            {
                sig = 1ULL << 63;
                exp += 1;
            }
        }
        // second rounding, of the 64-bit value alone, to 53 bits
        const std::uint64_t low = sig & 0x7FF;
        std::uint64_t m = sig >> 11;
        exp += 11;
        if (low > 0x400 || (low == 0x400 && (m & 1)))
        {
            m += 1;
            if (m == (1ULL << 53))
            {
                m = 1ULL << 52;
                exp += 1;
            }
        }
        return std::ldexp(static_cast<double>(m), exp);
    }
} // namespace hyt
