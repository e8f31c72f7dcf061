// CPU check of hyt::extended_quotient (the device hub score) against the
// compiler's own x87 long double division and double cast, on random and
// adversarial integer pairs a <= b < 2^64. Exit status 0 = all equal.
#include "extended_quotient.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

static bool same(double x, double y)
{
    return std::memcmp(&x, &y, sizeof(double)) == 0;
}

static std::uint64_t fails = 0, checked = 0, double_rounded = 0;

static void check(std::uint64_t a, std::uint64_t b)
{
    if (b == 0)
        return;
    if (a > b)
    {
        const std::uint64_t t = a;
        a = b;
        b = t;
    }
    const double want = static_cast<double>(static_cast<long double>(a) / static_cast<long double>(b));
    const double got = hyt::extended_quotient(a, b);
    ++checked;
    // how often the one-step double division differs (the case the helper exists for)
    if (!same(want, static_cast<double>(a) / static_cast<double>(b)) && a < (1ULL << 53) && b < (1ULL << 53))
        ++double_rounded;
    if (!same(want, got))
    {
        if (++fails <= 10)
            std::printf("mismatch a=%llu b=%llu want=%a got=%a\n", (unsigned long long)a, (unsigned long long)b,
                        want, got);
    }
}

int main()
{
    static_assert(sizeof(long double) > sizeof(double), "needs the x87 extended long double");
    std::mt19937_64 rng(12345);
    // degree products of graphs: a = outdeg * indeg, b = max_out * max_in
    for (int i = 0; i < 2000000; ++i)
    {
        const std::uint64_t mo = 1 + rng() % (1ULL << (1 + rng() % 32));
        const std::uint64_t mi = 1 + rng() % (1ULL << (1 + rng() % 32));
        check((rng() % (mo + 1)) * (rng() % (mi + 1)), mo * mi);
    }
    // full-range pairs and ones near the ends
    for (int i = 0; i < 2000000; ++i)
    {
        const std::uint64_t b = rng() >> (rng() % 64);
        check(rng() % (b + 1), b);
        check(b - (rng() % 16), b);
        check(1 + rng() % 16, b);
    }
    for (std::uint64_t b = 1; b < 2000; ++b)
        for (std::uint64_t a = 0; a <= b; ++a)
            check(a, b);
    check(1, ~0ULL);
    check(~0ULL - 1, ~0ULL);
    check(1ULL << 63, ~0ULL);
    std::printf("extended_quotient: %llu pairs, %llu mismatches, %llu where one double division differs\n",
                (unsigned long long)checked, (unsigned long long)fails, (unsigned long long)double_rounded);
    return fails == 0 && double_rounded > 0 ? 0 : 1;
}
