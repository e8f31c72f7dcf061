// Per-algorithm parts of the traversal kernels (row values, edge loads,
// destination updates, compaction selectors), shared by the in-HBM drivers
// and the hybrid transfer engine.
#pragma once

#include <type_traits>

#include "compact.hpp"
#include "expand.hpp"
#include "mass.hpp"

namespace hyt
{
    template <int WM>
    struct Weights;

    template <>
    struct Weights<kWU32>
    {
        static constexpr std::uint32_t kInf = 0xFFFFFFFFU;
        __device__ static std::uint32_t load(const std::uint32_t *w, std::uint64_t pos)
        {
            return __builtin_nontemporal_load(w + pos);
        }
        __device__ static bool add(std::uint32_t d, std::uint32_t w, std::uint32_t &out)
        {
            const std::uint64_t s = static_cast<std::uint64_t>(d) + w;
            out = static_cast<std::uint32_t>(s);
            return s < 0xFFFFFFFFULL;
        }
    };
    template <>
    struct Weights<kWUnitU32> : Weights<kWU32>
    {
        __device__ static std::uint32_t load(const std::uint32_t *, std::uint64_t) { return 1U; }
    };
    template <>
    struct Weights<kWF32>
    {
        static constexpr std::uint32_t kInf = 0x7F800000U;
        __device__ static std::uint32_t load(const std::uint32_t *w, std::uint64_t pos)
        {
            return __builtin_nontemporal_load(w + pos);
        }
        __device__ static bool add(std::uint32_t d, std::uint32_t w, std::uint32_t &out)
        {
            out = __float_as_uint(__fadd_rn(__uint_as_float(d), __uint_as_float(w)));
            return true;
        }
    };
    template <>
    struct Weights<kWUnitF32> : Weights<kWF32>
    {
        __device__ static std::uint32_t load(const std::uint32_t *, std::uint64_t) { return 0x3F800000U; }
    };

    template <int WM, bool SYNC>
    struct SsspOp
    {
        const std::uint32_t *__restrict__ col;
        const std::uint32_t *__restrict__ w;
        std::uint32_t *dist;
        std::uint8_t *touch;
        const std::uint32_t *qd; // snapshot values (SYNC)
        std::uint32_t *overflow;
        // ASYNC, optional: the compaction's last-seen distances. A row read
        // its live distance here, so recording that value keeps the next
        // compaction from queuing the row again when it was improved in this
        // round before its own expansion (its edges already saw the value).
        std::uint32_t *expanded_at = nullptr;

        __device__ std::uint32_t row_value(std::uint32_t i, std::uint32_t u, std::uint64_t) const
        {
            return SYNC ? qd[i] : dist[u];
        }
        // Called once per row with the value its edges use; whole_row = every
        // edge of the row is walked with it in this launch.
        __device__ void expanded(std::uint32_t u, std::uint32_t value, bool whole_row) const
        {
            if (!SYNC && whole_row && expanded_at != nullptr)
                expanded_at[u] = value;
        }
        // A row going out as chunks: each chunk reads the row's live distance
        // when it starts (the light kernel may have lowered it since the row
        // was queued) and raises the record to the largest value any chunk
        // used, so the record starts at 0 (ASYNC, whole rows only).
        __device__ void expanded_chunked(std::uint32_t u, bool whole_row) const
        {
            if (!SYNC && whole_row && expanded_at != nullptr)
                expanded_at[u] = 0U;
        }
        // A row above light_max that waits this round (its distance is above
        // the round's bound): its record becomes "never expanded", so the
        // next compaction queues it again whatever its distance does.
        __device__ void deferred(std::uint32_t u) const
        {
            if (!SYNC && expanded_at != nullptr)
                expanded_at[u] = 0xFFFFFFFFU;
        }
        __device__ std::uint32_t chunk_value(std::uint32_t u, std::uint32_t queued, bool record) const
        {
            if (SYNC)
                return queued;
            std::uint32_t v = 0;
            if (lane_id() == 0)
            {
                const std::uint32_t live = dist[u];
                v = live < queued ? live : queued;
                if (record && expanded_at != nullptr)
                    atomicMax(&expanded_at[u], v);
            }
            return __shfl(v, 0, kWave);
        }
        // columns and weights stream (each edge is read once per expansion):
        // nontemporal loads leave the L2 to the destinations' distances
        __device__ bool load(std::uint64_t pos, std::uint32_t du, std::uint32_t &v, std::uint32_t &c) const
        {
            v = __builtin_nontemporal_load(col + pos);
            if (Weights<WM>::add(du, Weights<WM>::load(w, pos), c))
                return true;
            atomicOr(overflow, 1U);
            return false;
        }
        // The walk reads the destinations of all its in-flight edges first
        // (peek) and only then issues their atomics (apply with the value
        // seen), so the reads overlap instead of each waiting behind the
        // previous edge's atomic to the same array. A value seen may be older
        // and larger than the current one; that only costs an extra atomic,
        // it never skips a needed one. Without touch[] the compaction
        // compares every row's dist with prev.
        __device__ std::uint32_t peek(std::uint32_t v) const { return dist[v]; }
        __device__ void apply(std::uint32_t v, std::uint32_t c, std::uint32_t seen) const
        {
            if (c < seen)
            {
                atomicMin(&dist[v], c);
                if (touch != nullptr)
                    touch[v] = 1;
            }
        }
    };

    // Touched rows that improved join the frontier when below `bound`
    // (delta-stepping bucket; 0xFFFFFFFF = no bucketing); improved rows at
    // or above the bound stay marked and are counted as pending.
    struct SsspSel
    {
        struct Aux
        {
            std::uint32_t pending = 0;
            std::uint32_t inv_max = 0;
            std::uint32_t kept_inv = 0; // max of ~dist over kept rows (track_min)
        };
        std::uint8_t *touch;
        const std::uint32_t *dist;
        std::uint32_t *prev;
        std::uint32_t bound;
        // record the smallest kept distance (for rows above light_max that
        // wait while their distance exceeds it by more than a margin)
        bool track_min = false;

        __device__ void publish(const Aux &aux, Counters *ctr) const
        {
            if (track_min)
            {
                std::uint32_t k = aux.kept_inv;
#pragma unroll
                for (int d = kWave / 2; d > 0; d >>= 1)
                {
                    const std::uint32_t o = __shfl_xor(k, d, kWave);
                    k = o > k ? o : k;
                }
                if (lane_id() == 0 && k)
                    atomicMax(&ctr->slot[blockIdx.x % kEdgeSlots].kinv, k);
            }
            if (bound == 0xFFFFFFFFU)
                return;
            std::uint32_t p = aux.pending, m = aux.inv_max;
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1)
            {
                p += __shfl_xor(p, d, kWave);
                const std::uint32_t o = __shfl_xor(m, d, kWave);
                m = o > m ? o : m;
            }
            if (lane_id() == 0 && p)
            {
                atomicAdd(&ctr->pending, p);
                atomicMax(&ctr->pending_inv_max, m);
            }
        }

        __device__ std::uint32_t select4(std::uint64_t row0, std::uint32_t n, std::uint32_t (&vals)[4],
                                         Aux &aux) const
        {
            if (touch == nullptr)
                return scan4(row0, n, vals, aux);
            // touch has spare bytes past V, so the 4-byte word is in range
            std::uint32_t *tw = reinterpret_cast<std::uint32_t *>(touch + row0);
            const std::uint32_t t = *tw;
            if (t == 0)
                return 0;
            std::uint32_t keep = 0, m = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                if (static_cast<std::uint32_t>(j) < n && ((t >> (8 * j)) & 0xFFU))
                {
                    const std::uint64_t v = row0 + j;
                    const std::uint32_t d = dist[v];
                    if (d < prev[v])
                    {
                        if (d < bound)
                        {
                            prev[v] = d;
                            vals[j] = d;
                            aux.kept_inv = (~d > aux.kept_inv) ? ~d : aux.kept_inv;
                            m |= 1U << j;
                        }
                        else
                        {
                            keep |= 0xFFU << (8 * j);
                            aux.pending += 1;
                            aux.inv_max = (~d > aux.inv_max) ? ~d : aux.inv_max;
                        }
                    }
                }
            }
            if (keep != t)
                *tw = keep;
            return m;
        }

        // Touch-free mode: rows improved since the last compaction are the
        // rows with dist < prev (row0 is a multiple of 4 and both arrays have
        // spare slots past V, so the 16-byte vector accesses stay in range).
        __device__ std::uint32_t scan4(std::uint64_t row0, std::uint32_t n, std::uint32_t (&vals)[4], Aux &aux) const
        {
            const uint4 d4 = *reinterpret_cast<const uint4 *>(dist + row0);
            const uint4 p4 = *reinterpret_cast<const uint4 *>(prev + row0);
            const std::uint32_t d[4] = {d4.x, d4.y, d4.z, d4.w};
            std::uint32_t p[4] = {p4.x, p4.y, p4.z, p4.w};
            std::uint32_t m = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                if (static_cast<std::uint32_t>(j) < n && d[j] < p[j])
                {
                    if (d[j] < bound)
                    {
                        p[j] = d[j];
                        vals[j] = d[j];
                        aux.kept_inv = (~d[j] > aux.kept_inv) ? ~d[j] : aux.kept_inv;
                        m |= 1U << j;
                    }
                    else
                    {
                        aux.pending += 1;
                        aux.inv_max = (~d[j] > aux.inv_max) ? ~d[j] : aux.inv_max;
                    }
                }
            }
            if (m)
                *reinterpret_cast<uint4 *>(prev + row0) = make_uint4(p[0], p[1], p[2], p[3]);
            return m;
        }

        // Touch-free mode reads the thread's 16 rows of dist and prev before
        // any store (row0 is a multiple of 16 below V; the arrays have 16
        // spare slots past V, so the vector reads stay in range, and rows at
        // or past V are never selected).
        __device__ bool whole16() const { return touch == nullptr; }
        __device__ void select16(std::uint64_t row0, std::uint64_t V, std::uint32_t (&vals)[4][4],
                                 std::uint32_t (&masks)[4], Aux &aux) const
        {
            uint4 d4[4], p4[4];
#pragma unroll
            for (int g = 0; g < 4; ++g)
            {
                d4[g] = *reinterpret_cast<const uint4 *>(dist + row0 + 4 * g);
                p4[g] = *reinterpret_cast<const uint4 *>(prev + row0 + 4 * g);
            }
#pragma unroll
            for (int g = 0; g < 4; ++g)
            {
                const std::uint64_t r = row0 + 4 * g;
                const std::uint32_t n = r >= V ? 0U : static_cast<std::uint32_t>(V - r < 4 ? V - r : 4);
                const std::uint32_t d[4] = {d4[g].x, d4[g].y, d4[g].z, d4[g].w};
                std::uint32_t p[4] = {p4[g].x, p4[g].y, p4[g].z, p4[g].w};
                std::uint32_t m = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                {
                    if (static_cast<std::uint32_t>(j) < n && d[j] < p[j])
                    {
                        if (d[j] < bound)
                        {
                            p[j] = d[j];
                            vals[g][j] = d[j];
                            aux.kept_inv = (~d[j] > aux.kept_inv) ? ~d[j] : aux.kept_inv;
                            m |= 1U << j;
                        }
                        else
                        {
                            aux.pending += 1;
                            aux.inv_max = (~d[j] > aux.inv_max) ? ~d[j] : aux.inv_max;
                        }
                    }
                }
                masks[g] = m;
                if (m)
                    *reinterpret_cast<uint4 *>(prev + r) = make_uint4(p[0], p[1], p[2], p[3]);
            }
        }
    };

    // Share words per out-edge, as PrSel stores them by vertex: PPR keeps the
    // words themselves (8 bytes), global PageRank the words rounded to float
    // (4 bytes per gather in pull rounds; mass.hpp).
    template <bool WORDS>
    using ShareT = std::conditional_t<WORDS, unsigned long long, float>;
    __device__ __forceinline__ unsigned long long word_of(unsigned long long w) { return w; }
    __device__ __forceinline__ unsigned long long word_of(float f) { return __float2ull_rn(f); }

    // Delta-push of PageRank / PPR mass (mass.hpp): every edge adds its row's
    // share word to the residual of its destination with an integer atomic
    // (exact, order-independent). PPR (WORDS): a row carries its vertex id
    // and the word is read from share[] (the rows of one wavefront step share
    // few words, so the per-edge read is a cache hit); global PageRank: a row
    // carries its float share as the queue value.
    template <bool WORDS>
    struct PrOp
    {
        const std::uint32_t *__restrict__ col;
        unsigned long long *res;
        const unsigned long long *__restrict__ share; // WORDS: share word per out-edge, by vertex
        const std::uint32_t *qd;                      // !WORDS: float share of each queue row

        __device__ std::uint32_t row_value(std::uint32_t i, std::uint32_t u, std::uint64_t deg) const
        {
            if constexpr (WORDS)
                return u;
            else
                return deg ? qd[i] : 0U;
        }
        __device__ void expanded(std::uint32_t, std::uint32_t, bool) const {}
        __device__ void expanded_chunked(std::uint32_t, bool) const {}
        __device__ void deferred(std::uint32_t) const {}
        __device__ std::uint32_t chunk_value(std::uint32_t, std::uint32_t queued, bool) const { return queued; }
        __device__ bool load(std::uint64_t pos, std::uint32_t share, std::uint32_t &v, std::uint32_t &c) const
        {
            v = col[pos];
            c = share;
            return true;
        }
        __device__ std::uint32_t peek(std::uint32_t) const { return 0U; }
        __device__ void apply(std::uint32_t v, std::uint32_t c, std::uint32_t) const
        {
            unsigned long long w;
            if constexpr (WORDS)
                w = share[c];
            else
                w = word_of(__uint_as_float(c));
            if (w != 0ULL)
                atomicAdd(&res[v], w);
        }
    };

    // Selects the rows whose residual word is above eps, moves the residual
    // into rank and zeroes it, and writes the row's share word per out-edge
    // d * r_u / outdeg(u) (mass.hpp) by vertex into share[] (ShareT<WORDS>;
    // 0 outside the frontier); the queue value is that word as a float (the
    // push rounds of global PageRank, the hybrid engine's pushed-mass
    // priorities).
    template <class OffT, bool WORDS>
    struct PrSel
    {
        using Aux = NoAux;
        unsigned long long *res;
        unsigned long long *rank;
        ShareT<WORDS> *share;
        const OffT *off;
        float damping;
        unsigned long long eps;

        // no 16-row path
        __device__ bool whole16() const { return false; }
        template <class Vals, class A>
        __device__ void select16(std::uint64_t, std::uint64_t, Vals &, std::uint32_t (&)[4], A &) const
        {
        }
        __device__ void publish(const Aux &, Counters *) const {}
        __device__ std::uint32_t select4(std::uint64_t row0, std::uint32_t n, std::uint32_t (&vals)[4],
                                         Aux &) const
        {
            // row0 is a multiple of 4, so a full group is two aligned 16-byte words
            unsigned long long r[4];
            if (n == 4)
            {
                const ulonglong2 a = *reinterpret_cast<const ulonglong2 *>(res + row0);
                const ulonglong2 b = *reinterpret_cast<const ulonglong2 *>(res + row0 + 2);
                r[0] = a.x, r[1] = a.y, r[2] = b.x, r[3] = b.y;
            }
            else
            {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    r[j] = static_cast<std::uint32_t>(j) < n ? res[row0 + j] : 0ULL;
            }
            std::uint32_t m = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (r[j] > eps)
                    m |= 1U << j;
            ShareT<WORDS> sh[4] = {};
            if (m)
            {
                // offsets of the 4 rows (+1); a full group reads them as one
                // vector when the row offsets are 32-bit
                OffT o[5];
                if (n == 4 && sizeof(OffT) == 4)
                {
                    const uint4 o4 = *reinterpret_cast<const uint4 *>(off + row0);
                    o[0] = o4.x, o[1] = o4.y, o[2] = o4.z, o[3] = o4.w;
                    o[4] = off[row0 + 4];
                }
                else
                {
#pragma unroll
                    for (int j = 0; j < 5; ++j)
                        o[j] = static_cast<std::uint32_t>(j) <= n ? off[row0 + j] : o[j - 1];
                }
                unsigned long long k[4] = {0ULL, 0ULL, 0ULL, 0ULL};
                if (n == 4)
                {
                    const ulonglong2 a = *reinterpret_cast<const ulonglong2 *>(rank + row0);
                    const ulonglong2 b = *reinterpret_cast<const ulonglong2 *>(rank + row0 + 2);
                    k[0] = a.x, k[1] = a.y, k[2] = b.x, k[3] = b.y;
                }
                else
                {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (static_cast<std::uint32_t>(j) < n)
                            k[j] = rank[row0 + j];
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                {
                    if ((m >> j) & 1U)
                    {
                        k[j] += r[j];
                        const unsigned long long sw =
                            share_word(r[j], row_scale_of(damping, static_cast<std::uint64_t>(o[j + 1] - o[j])));
                        const float sf = __ull2float_rn(sw);
                        if constexpr (WORDS)
                            sh[j] = sw;
                        else
                            sh[j] = sf;
                        vals[j] = __float_as_uint(sf);
                        r[j] = 0ULL;
                    }
                }
                if (n == 4)
                {
                    *reinterpret_cast<ulonglong2 *>(res + row0) = make_ulonglong2(r[0], r[1]);
                    *reinterpret_cast<ulonglong2 *>(res + row0 + 2) = make_ulonglong2(r[2], r[3]);
                    *reinterpret_cast<ulonglong2 *>(rank + row0) = make_ulonglong2(k[0], k[1]);
                    *reinterpret_cast<ulonglong2 *>(rank + row0 + 2) = make_ulonglong2(k[2], k[3]);
                }
                else
                {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if ((m >> j) & 1U)
                        {
                            res[row0 + j] = 0ULL;
                            rank[row0 + j] = k[j];
                        }
                }
            }
            // every row's share is rewritten (0 outside the frontier), so a
            // pull round needs no frontier test: one gather per in-edge
            if (n == 4)
            {
                if constexpr (WORDS)
                {
                    *reinterpret_cast<ulonglong2 *>(share + row0) = make_ulonglong2(sh[0], sh[1]);
                    *reinterpret_cast<ulonglong2 *>(share + row0 + 2) = make_ulonglong2(sh[2], sh[3]);
                }
                else
                    *reinterpret_cast<float4 *>(share + row0) = make_float4(sh[0], sh[1], sh[2], sh[3]);
            }
            else
            {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (static_cast<std::uint32_t>(j) < n)
                        share[row0 + j] = sh[j];
            }
            return m;
        }
    };

} // namespace hyt
