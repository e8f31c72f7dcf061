// Internal helpers shared by the host generator and the HIP library:
// thread-local error reporting and the counter-based hash used for every
// synthetic draw (graph edges, weights, source picks).
#pragma once

#include "hytgraph_amd.h"

#include <cstdint>
#include <exception>
#include <stdexcept>
#include <string>

namespace hyt
{

    // An error carrying a C-ABI status code.
    struct Error : std::runtime_error
    {
        int status;
        Error(int s, const std::string &msg) : std::runtime_error(msg), status(s) {}
    };

    void set_last_error(const std::string &msg);

    [[noreturn]] inline void fail(int status, const std::string &msg)
    {
        throw Error(status, msg);
    }

    inline void require(bool cond, const std::string &msg)
    {
        if (!cond)
            fail(HYT_ERR_INVALID_ARGUMENT, msg);
    }

    // Runs fn() and converts any exception into a status + last-error message.
    template <class Fn>
    int guarded(const char *where, Fn &&fn) noexcept
    {
        try
        {
            fn();
            return HYT_OK;
        }
        catch (const Error &e)
        {
            set_last_error(std::string(where) + ": " + e.what());
            return e.status;
        }
        catch (const std::invalid_argument &e)
        {
            set_last_error(std::string(where) + ": " + e.what());
            return HYT_ERR_INVALID_ARGUMENT;
        }
        catch (const std::overflow_error &e)
        {
            set_last_error(std::string(where) + ": " + e.what());
            return HYT_ERR_OVERFLOW;
        }
        catch (const std::bad_alloc &)
        {
            set_last_error(std::string(where) + ": host allocation failed");
            return HYT_ERR_OUT_OF_MEMORY;
        }
        catch (const std::exception &e)
        {
            set_last_error(std::string(where) + ": " + e.what());
            return HYT_ERR_INTERNAL;
        }
        catch (...)
        {
            set_last_error(std::string(where) + ": unknown exception");
            return HYT_ERR_INTERNAL;
        }
    }

    // SplitMix64 finaliser (Steele, Lea, Flood 2014): a bijective 64-bit mixer.
    constexpr inline std::uint64_t mix64(std::uint64_t x)
    {
        x += 0x9E3779B97F4A7C15ULL;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
        return x ^ (x >> 31);
    }

    // Counter-based draw: a pure function of (seed, stream, counter, word).
    // Distinct streams separate edges, weights and source picks.
    constexpr inline std::uint64_t draw(std::uint64_t seed, std::uint64_t stream,
                                        std::uint64_t counter, std::uint64_t word)
    {
        std::uint64_t k = mix64(seed ^ mix64(stream + 0x632BE59BD9B4E019ULL));
        return mix64(k + counter * 0xD1B54A32D192ED03ULL + word * 0x8CB92BA72F3D8DD7ULL);
    }

    // Uniform integer in [0, range) from a 32-bit draw (multiply-shift).
    constexpr inline std::uint64_t bounded32(std::uint32_t r, std::uint64_t range)
    {
        return (static_cast<std::uint64_t>(r) * range) >> 32;
    }

    // Uniform integer in [0, range) from a 64-bit draw, any 64-bit range.
    inline std::uint64_t bounded64(std::uint64_t r, std::uint64_t range)
    {
        return static_cast<std::uint64_t>(
            (static_cast<unsigned __int128>(r) * range) >> 64);
    }

    enum Stream : std::uint64_t
    {
        kStreamEdges = 1,
        kStreamWeights = 2,
        kStreamSources = 3
    };

} // namespace hyt
