// Persistent single-source SSSP: every round of a run inside one cooperative
// launch - the data-driven push variants of the original's SEP contract
// (ASYNC_PUSH_DD / SYNC_PUSH_DD, include/sep/sep_execution_variant.hpp:
// 73-315), with the semantics of src/algorithms/sssp.cpp:109-234.
//
// A resident grid (co-residency guaranteed by hipLaunchCooperativeKernel,
// one 1024-thread workgroup per CU) loops over rounds separated by grid
// barriers of its own (two-level counters, below); nothing returns to the host
// between rounds and no launch or O(V) pass runs per round. The next queue
// is built by the relaxations themselves: an edge whose candidate is below
// the distance it read lowers dist[v] (atomicMin, no return value) and
// queues v for the next round unless v is already in it - stamp[v] holds the
// last round v was queued for and is raised by atomicMax, so a vertex enters
// a queue at most once and a queue needs V slots; appends gather in LDS and
// take one counter atomic per flush. Rows up to light_max edges are walked 64
// rows per wavefront tile (degree prefix sums in LDS, one lane per edge);
// longer rows go out as chunk-edge pieces, one wavefront per piece. Tiles
// are taken per wavefront, so no workgroup barrier runs inside a round.
//
// Measured on one MI355X against the captured-round path (ms per query,
// 8 sources, u32 weights [1,255], ASYNC / SYNC): RMAT 16 1.50 / 1.21 vs
// 1.29 / 1.12; RMAT 20 2.73 / 2.62 vs 2.56 / 2.35; RMAT 24 16.0 / 22.5 vs
// 12.9 / 18.6. A round costs ~60 us of barrier, fences and queue handling
// on small graphs, and relaxation runs at ~60 G edges/s against ~88 G/s in
// the captured kernels (the stamp read of every candidate). ASYNC rounds
// wait rows above light_max only when hyt_run_options.heavy_delta is set:
// with heavy_delta 2, RMAT 24 relaxes 457 M edges per query instead of 553 M
// but runs 24.8 rounds instead of 20.3, in the same kernel time (9.4 ms),
// and RMAT 16 / 20 take 1.89 / 3.22 ms instead of 1.37 / 2.50 (captured
// rounds: 345 M relaxations, 16.1 rounds). The captured rounds stay the
// default; this kernel runs when a caller asks for it
// (hyt_run_options.persistent).
//
// ASYNC round r (one barrier): walk the chunks queued in round r-1 (each
// reads its row's live distance), then the rows of queue r with their live
// distances; rows above light_max queue their chunks for round r+1. A row
// whose distance equals the value of its last expansion (rec) is skipped:
// its edges already saw that value (chunks keep rec by atomic max, from 0).
// With waiting rows, a row above the wait degree whose distance exceeds the
// queue's smallest by more than heavy_delta joins queue r+1 again instead;
// the smallest distance of queue r+1 is the smallest candidate of round r or
// of a waiting row, gathered by one atomic per workgroup.
//
// SYNC round r (two barriers, the original's Jacobi rounds and iteration
// count): phase A records each queued row's distance - the value it held at
// the end of round r-1, since no relaxation runs in phase A - and queues the
// chunks of long rows with it; phase B walks chunks and light rows with those
// values. A vertex joins queue r+1 exactly when some edge lowered it in
// round r, the set the original's `changed` flags mark.
//
// Counter slots rotate over three rounds: round r reads slot r%3, appends to
// slot (r+1)%3 and clears slot (r+2)%3, which every workgroup read before the
// last barrier. Queue and chunk buffers alternate over two. Every workgroup
// decides to stop from the same counters, read after a barrier, so all of
// them leave the loop in the same round.
#include "ops.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>

namespace hyt
{
    namespace
    {
        // One 1024-thread workgroup per CU: 16 wavefronts per CU keep loads in
        // flight, and a barrier's cost grows with the workgroups it joins
        // (one MI355X, two-level barrier below: 8.6 us at 256 workgroups,
        // 16 us at 512; ROCm's cooperative-groups grid sync 36 / 52 us).
        // Two workgroups per CU (32 wavefronts) measured the same as one
        // (RMAT 24 SYNC 15.94 vs 15.86 ms of kernel time per query), so one.
        constexpr int kPBlock = 1024;
        constexpr int kPBlocksPerCu = 1;
        constexpr int kPWaves = kPBlock / kWave;
        constexpr int kGroups = 8; // barrier groups (workgroup b joins group b % 8, the XCD it runs on)

        struct alignas(256) BarrierLine
        {
            std::uint32_t n;
        };

        struct alignas(256) PersistCtl
        {
            std::uint32_t rows[3];   // queue sizes by round slot
            std::uint32_t chunks[3]; // chunk-queue sizes by round slot
            std::uint32_t qmin[3];   // ASYNC: smallest distance in the queue, by round slot
            std::uint32_t overflow;  // a u32 distance sum overflowed
            std::uint32_t queue_overflow;
            std::uint32_t truncated; // stopped by the round limit with work left
            unsigned long long rounds, frontier, light_edges, chunk_edges;
            // grid barrier: arrivals per group, groups arrived, generation
            BarrierLine arrive[kGroups];
            BarrierLine groups;
            BarrierLine generation;
            // row tiles taken, per phase (SYNC value pass, walk), round slot
            // and group
            BarrierLine grab[2][3][kGroups];
        };

        // Arrays of one persistent run (every [V] array has kPad spare slots).
        struct PersistWs
        {
            DeviceBuffer stamp, qb, heavy_b, ctl;
            std::uint64_t V = 0, heavy_cap = 0;
        };

        template <class OffT>
        struct PersistArgs
        {
            const OffT *off;
            const std::uint32_t *col;
            const std::uint32_t *w;
            std::uint32_t *dist;
            std::uint32_t *stamp;
            std::uint32_t *rec;  // ASYNC: value of each row's last expansion
            std::uint32_t *qval; // SYNC: the queue rows' values of this round
            std::uint32_t *q[2];
            HeavyChunk *chunks[2];
            std::uint64_t queue_cap, chunk_cap;
            PersistCtl *ctl;
            std::uint32_t light_max, chunk;
            // ASYNC waiting rows: a queued row of more than wait_min edges
            // whose distance exceeds the queue's smallest by more than
            // heavy_delta waits for a later round (heavy_delta 0: none waits)
            std::uint32_t wait_min, heavy_delta;
            std::uint64_t max_rounds;
        };

        __device__ __forceinline__ std::uint32_t ld_agent(const std::uint32_t *p)
        {
            return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __device__ __forceinline__ void st_agent(std::uint32_t *p, std::uint32_t v)
        {
            __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }

        // Grid barrier number `gen` (1, 2, ...) of a co-resident grid, in two
        // levels: a workgroup's first thread releases the workgroup's writes
        // with an agent-scope fence (its XCD's L2 written back), counts itself
        // in its group's line, the last of a group counts the group in the
        // groups line, and the last group raises the generation that every
        // first thread waits for; the acquiring fence after the wait makes
        // every XCD's writes visible. Same-address atomics serialise, so eight
        // group lines of ~32 arrivals each are cheaper than one of 256.
        __device__ __forceinline__ void grid_barrier(PersistCtl *ctl, std::uint32_t gen)
        {
            __syncthreads();
            if (threadIdx.x == 0)
            {
                __threadfence();
                const std::uint32_t ng = gridDim.x < kGroups ? gridDim.x : kGroups;
                const std::uint32_t grp = blockIdx.x % kGroups;
                const std::uint32_t members = (gridDim.x - grp + kGroups - 1) / kGroups;
                const std::uint32_t a = __hip_atomic_fetch_add(&ctl->arrive[grp].n, 1U, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_AGENT);
                if (a == gen * members - 1)
                {
                    const std::uint32_t b = __hip_atomic_fetch_add(&ctl->groups.n, 1U, __ATOMIC_RELAXED,
                                                                   __HIP_MEMORY_SCOPE_AGENT);
                    if (b == gen * ng - 1)
                        st_agent(&ctl->generation.n, gen);
                }
                while (ld_agent(&ctl->generation.n) < gen)
                    __builtin_amdgcn_s_sleep(1);
                __threadfence();
            }
            __syncthreads();
        }

        // Appends the chunks of this wavefront's long rows (nch for a lane's
        // row) to `out` with one atomic per wavefront on *count.
        template <class MakeFn>
        __device__ __forceinline__ void append_chunks(std::uint32_t nch, std::uint32_t *count, HeavyChunk *out,
                                                      std::uint64_t cap, std::uint32_t *overflow, MakeFn make)
        {
            const std::uint32_t incl = wave_incl_scan(nch);
            const std::uint32_t total = __shfl(incl, kWave - 1, kWave);
            if (total == 0)
                return;
            std::uint32_t base = 0;
            if (lane_id() == kWave - 1)
                base = atomicAdd(count, total);
            const std::uint64_t slot = static_cast<std::uint64_t>(__shfl(base, kWave - 1, kWave)) + incl - nch;
            for (std::uint32_t c = 0; c < nch; ++c)
            {
                if (slot + c < cap)
                    out[slot + c] = make(c);
                else
                    atomicOr(overflow, 1U);
            }
        }

        // Row tiles of 64 queue entries: wavefront w takes tile w, and when a
        // round has more tiles than wavefronts the rest are taken kGrabTiles
        // at a time from the wavefront's group counter - group g takes
        // super-tiles g, g + 8, g + 16, ... of the rest, so the groups' shares
        // interleave over the queue, and within a group the wavefronts that
        // finish first take more. Small rounds take no atomic.
        constexpr std::uint32_t kGrabTiles = 2;
        template <class TileFn>
        __device__ __forceinline__ void for_each_tile(std::uint32_t *grab, std::uint32_t F, std::uint32_t gwave,
                                                      std::uint32_t waves, TileFn fn)
        {
            const std::uint64_t tiles = (static_cast<std::uint64_t>(F) + kWave - 1) / kWave;
            if (gwave < tiles)
                fn(gwave * static_cast<std::uint32_t>(kWave));
            if (tiles <= waves)
                return;
            const std::uint32_t grp = blockIdx.x % kGroups;
            for (;;)
            {
                std::uint32_t k = 0;
                if (lane_id() == 0)
                    k = atomicAdd(grab, 1U);
                k = __shfl(k, 0, kWave);
                const std::uint64_t first = waves + (static_cast<std::uint64_t>(k) * kGroups + grp) * kGrabTiles;
                if (first >= tiles)
                    break;
                for (std::uint32_t t = 0; t < kGrabTiles; ++t)
                    if (first + t < tiles)
                        fn(static_cast<std::uint32_t>((first + t) * kWave));
            }
        }

        // Queue appends of one wavefront, gathered in its LDS slice and
        // written with one counter atomic per flush (per-append atomics on
        // the one queue counter serialise across the grid).
        constexpr std::uint32_t kPushCap = 512; // >= kUnroll * kWave, the most one step adds
        struct PushBuf
        {
            std::uint32_t *buf; // LDS, kPushCap entries
            std::uint32_t n;    // entries held (lane-uniform)
            std::uint32_t *qout;
            std::uint32_t *qcount;
            std::uint64_t qcap;
            std::uint32_t *overflow;

            __device__ __forceinline__ void flush()
            {
                if (n == 0)
                    return;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                std::uint32_t base = 0;
                if (lane_id() == 0)
                    base = atomicAdd(qcount, n);
                base = __shfl(base, 0, kWave);
                for (std::uint32_t i = static_cast<std::uint32_t>(lane_id()); i < n; i += kWave)
                {
                    if (static_cast<std::uint64_t>(base) + i < qcap)
                        qout[base + i] = buf[i];
                    else
                        atomicOr(overflow, 1U);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                n = 0;
            }
            // v from every lane whose flag is set (called by the whole wavefront)
            __device__ __forceinline__ void append(bool flag, std::uint32_t v)
            {
                const std::uint32_t incl = wave_incl_scan(flag ? 1U : 0U);
                const std::uint32_t total = __shfl(incl, kWave - 1, kWave);
                if (total == 0)
                    return;
                if (n + total > kPushCap)
                    flush();
                if (flag)
                    buf[n + incl - 1] = v;
                n += total;
            }
        };

        // The relaxation of kUnroll x 64 edges per iteration: all edge loads,
        // then all destination reads, then the min atomics of the candidates
        // below what was read (no return value: nothing waits on them), then
        // the queue test. A candidate c below the distance read lowers v in
        // this round: every read after the round's barrier sees at most the
        // distance v held at the barrier, so the distance ends the round at
        // most c, below its start. v is queued for round `next` unless its
        // stamp already says so: a plain read filters, and only the first
        // candidates of v in a round issue the returning atomicMax that
        // decides. Appends take one atomic per wavefront. qmin keeps the
        // lane's smallest candidate: the smallest distance of the next queue
        // is the smallest candidate of the round (every vertex lowered in the
        // round is queued, and ends it at its smallest candidate).
        template <int WM, class PosFn>
        __device__ __forceinline__ void relax_edges(const std::uint32_t *__restrict__ col,
                                                    const std::uint32_t *__restrict__ w, std::uint32_t *dist,
                                                    std::uint32_t *stamp, std::uint32_t next, PushBuf &pb,
                                                    PersistCtl *ctl, std::uint32_t total, std::uint32_t &qmin,
                                                    PosFn pos_of)
        {
            const int lane = lane_id();
            for (std::uint32_t base = 0; base < total; base += kWave * kUnroll)
            {
                std::uint32_t vv[kUnroll], cc[kUnroll];
                bool ok[kUnroll];
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                {
                    const std::uint32_t e = base + static_cast<std::uint32_t>(k * kWave + lane);
                    ok[k] = false;
                    vv[k] = 0;
                    cc[k] = 0;
                    if (e < total)
                    {
                        std::uint64_t pos;
                        std::uint32_t value;
                        pos_of(e, pos, value);
                        vv[k] = col[pos];
                        ok[k] = Weights<WM>::add(value, Weights<WM>::load(w, pos), cc[k]);
                        if (!ok[k])
                            atomicOr(&ctl->overflow, 1U);
                    }
                }
                std::uint32_t seen[kUnroll];
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                    seen[k] = ok[k] ? dist[vv[k]] : 0U;
                bool cand[kUnroll];
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                {
                    cand[k] = ok[k] && cc[k] < seen[k];
                    if (cand[k])
                    {
                        atomicMin(&dist[vv[k]], cc[k]);
                        qmin = cc[k] < qmin ? cc[k] : qmin;
                    }
                }
                std::uint32_t prior[kUnroll];
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                    prior[k] = cand[k] ? stamp[vv[k]] : next;
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                    if (prior[k] < next)
                        prior[k] = atomicMax(&stamp[vv[k]], next);
                std::uint32_t npush = 0;
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                    npush += prior[k] < next ? 1U : 0U;
                const std::uint32_t incl = wave_incl_scan(npush);
                const std::uint32_t wtotal = __shfl(incl, kWave - 1, kWave);
                if (wtotal == 0)
                    continue;
                if (pb.n + wtotal > kPushCap)
                    pb.flush();
                std::uint32_t slot = pb.n + incl - npush;
#pragma unroll
                for (int k = 0; k < kUnroll; ++k)
                    if (prior[k] < next)
                        pb.buf[slot++] = vv[k];
                pb.n += wtotal;
            }
        }

        template <class OffT, int WM, bool SYNC>
        __global__ __launch_bounds__(kPBlock) void sssp_persistent(PersistArgs<OffT> a)
        {
            // per-wavefront LDS slices: a wavefront's lanes write and read only
            // their own slice, so no workgroup barrier is needed between tiles
            __shared__ std::uint32_t s_ex[kPWaves][kWave];
            __shared__ std::uint64_t s_begin[kPWaves][kWave];
            __shared__ std::uint32_t s_val[kPWaves][kWave];
            __shared__ std::uint32_t s_push[kPWaves][kPushCap];
            __shared__ std::uint32_t s_round[3];
            __shared__ std::uint32_t s_min[kPWaves];
            constexpr bool kFloatDist = WM == kWF32 || WM == kWUnitF32;

            const int lane = lane_id();
            const int wid = static_cast<int>(threadIdx.x) / kWave;
            const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
            const std::uint32_t waves = gridDim.x * kPWaves;
            const std::uint32_t gwave = (blockIdx.x * kPBlock + threadIdx.x) / kWave;
            const std::uint32_t grp = blockIdx.x % kGroups;
            std::uint32_t barriers = 0;
            PersistCtl *ctl = a.ctl;
            unsigned long long light_e = 0, chunk_e = 0; // this wavefront's edges (lane-uniform)

            for (std::uint32_t r = 0;; ++r)
            {
                const std::uint32_t s0 = r % 3, s1 = (r + 1) % 3, s2 = (r + 2) % 3;
                // the round's counters, read once per workgroup
                if (threadIdx.x == 0)
                {
                    s_round[0] = ld_agent(&ctl->rows[s0]);
                    s_round[1] = SYNC ? 0U : ld_agent(&ctl->chunks[s0]);
                    s_round[2] = SYNC ? 0xFFFFFFFFU : ld_agent(&ctl->qmin[s0]);
                }
                __syncthreads();
                const std::uint32_t F = s_round[0];
                const std::uint32_t H = s_round[1];
                const std::uint32_t hbound = SYNC ? 0xFFFFFFFFU : heavy_bound_of(s_round[2], a.heavy_delta, kFloatDist);
                __syncthreads(); // s_round is rewritten after the next barrier
                if (F == 0 && H == 0)
                    break;
                if (r >= a.max_rounds)
                {
                    if (leader)
                        ctl->truncated = 1;
                    break;
                }
                if (leader)
                {
                    // counters of round r+1 (read, or taken, before the last barrier)
                    st_agent(&ctl->rows[s2], 0U);
                    st_agent(&ctl->chunks[SYNC ? s1 : s2], 0U);
                    if (!SYNC)
                        st_agent(&ctl->qmin[s2], 0xFFFFFFFFU);
                    for (int p = 0; p < 2; ++p)
                        for (int g = 0; g < kGroups; ++g)
                            st_agent(&ctl->grab[p][s1][g].n, 0U);
                    ctl->rounds += 1;
                    ctl->frontier += F;
                }
                const std::uint32_t *qin = a.q[r & 1];
                const std::uint32_t next = r + 1;
                PushBuf pb{s_push[wid], 0U, a.q[(r + 1) & 1], &ctl->rows[s1], a.queue_cap, &ctl->queue_overflow};
                std::uint32_t qmin = 0xFFFFFFFFU; // this lane's smallest distance queued for round r+1

                std::uint32_t Hw = H;                    // chunks walked this round
                const HeavyChunk *hin = a.chunks[r & 1]; // ASYNC: queued in round r-1
                HeavyChunk *hout = a.chunks[SYNC ? 0 : ((r + 1) & 1)];
                std::uint32_t *hcount = &ctl->chunks[SYNC ? s0 : s1];
                if (SYNC)
                {
                    // phase A: the round's values, and the chunks of long rows
                    for_each_tile(&ctl->grab[0][s0][grp].n, F, gwave, waves, [&](std::uint32_t lo) {
                        const std::uint32_t i = lo + static_cast<std::uint32_t>(lane);
                        std::uint32_t nch = 0, value = 0, row = 0;
                        std::uint64_t begin = 0, count = 0;
                        if (i < F)
                        {
                            row = qin[i];
                            value = a.dist[row];
                            a.qval[i] = value;
                            begin = a.off[row];
                            count = a.off[row + 1] - begin;
                            if (count > a.light_max)
                                nch = static_cast<std::uint32_t>((count + a.chunk - 1) / a.chunk);
                        }
                        append_chunks(nch, hcount, hout, a.chunk_cap, &ctl->queue_overflow, [&](std::uint32_t c) {
                            const std::uint64_t clo = static_cast<std::uint64_t>(c) * a.chunk;
                            HeavyChunk h;
                            h.value = value;
                            h.count = static_cast<std::uint32_t>(count - clo < a.chunk ? count - clo : a.chunk);
                            h.begin = begin + clo;
                            h.row = row;
                            h.record = 0;
                            return h;
                        });
                    });
                    grid_barrier(ctl, ++barriers);
                    if (threadIdx.x == 0)
                        s_round[1] = ld_agent(hcount);
                    __syncthreads();
                    Hw = s_round[1];
                    hin = hout;
                }
                if (Hw > a.chunk_cap)
                    Hw = static_cast<std::uint32_t>(a.chunk_cap);

                // chunks: one wavefront each
                for (std::uint32_t c = gwave; c < Hw; c += waves)
                {
                    const HeavyChunk h = hin[c];
                    std::uint32_t value = h.value;
                    if (!SYNC)
                    {
                        std::uint32_t v = 0;
                        if (lane == 0)
                        {
                            const std::uint32_t live = a.dist[h.row];
                            v = live < h.value ? live : h.value;
                            atomicMax(&a.rec[h.row], v);
                        }
                        value = __shfl(v, 0, kWave);
                    }
                    chunk_e += h.count;
                    relax_edges<WM>(a.col, a.w, a.dist, a.stamp, next, pb, ctl, h.count, qmin,
                                    [&](std::uint32_t e, std::uint64_t &pos, std::uint32_t &val) {
                                        pos = h.begin + e;
                                        val = value;
                                    });
                }

                // queue rows: 64 rows per wavefront tile, long rows as chunks
                for_each_tile(&ctl->grab[1][s0][grp].n, F, gwave, waves, [&](std::uint32_t lo) {
                    const std::uint32_t i = lo + static_cast<std::uint32_t>(lane);
                    std::uint32_t deg = 0, value = 0, nch = 0, row = 0;
                    std::uint64_t begin = 0, count = 0;
                    bool wait = false; // ASYNC: the row waits, queued again for round r+1
                    if (i < F)
                    {
                        row = qin[i];
                        if (SYNC)
                        {
                            begin = a.off[row];
                            count = a.off[row + 1] - begin;
                            if (count <= a.light_max)
                            {
                                value = a.qval[i];
                                deg = static_cast<std::uint32_t>(count);
                            }
                        }
                        else
                        {
                            value = a.dist[row];
                            if (value != a.rec[row])
                            {
                                begin = a.off[row];
                                count = a.off[row + 1] - begin;
                                if (count > a.wait_min && value > hbound)
                                {
                                    // its record stays below, so the row is
                                    // not skipped when its turn comes
                                    wait = a.stamp[row] < next && atomicMax(&a.stamp[row], next) < next;
                                    qmin = value < qmin ? value : qmin;
                                }
                                else if (count > a.light_max)
                                {
                                    nch = static_cast<std::uint32_t>((count + a.chunk - 1) / a.chunk);
                                    a.rec[row] = 0U;
                                }
                                else
                                {
                                    deg = static_cast<std::uint32_t>(count);
                                    a.rec[row] = value;
                                }
                            }
                        }
                    }
                    if (!SYNC)
                    {
                        pb.append(wait, row);
                        append_chunks(nch, hcount, hout, a.chunk_cap, &ctl->queue_overflow, [&](std::uint32_t c) {
                            const std::uint64_t clo = static_cast<std::uint64_t>(c) * a.chunk;
                            HeavyChunk h;
                            h.value = value;
                            h.count = static_cast<std::uint32_t>(count - clo < a.chunk ? count - clo : a.chunk);
                            h.begin = begin + clo;
                            h.row = row;
                            h.record = 1;
                            return h;
                        });
                    }
                    const std::uint32_t incl = wave_incl_scan(deg);
                    const std::uint32_t total = __shfl(incl, kWave - 1, kWave);
                    if (total == 0)
                        return;
                    s_ex[wid][lane] = incl - deg;
                    s_begin[wid][lane] = begin;
                    s_val[wid][lane] = value;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    light_e += total;
                    const std::uint32_t *ex = s_ex[wid];
                    const std::uint64_t *bg = s_begin[wid];
                    const std::uint32_t *vl = s_val[wid];
                    relax_edges<WM>(a.col, a.w, a.dist, a.stamp, next, pb, ctl, total, qmin,
                                    [&](std::uint32_t e, std::uint64_t &pos, std::uint32_t &val) {
                                        const int j = owner_of(ex, e);
                                        pos = bg[j] + (e - ex[j]);
                                        val = vl[j];
                                    });
                    // the slice is rewritten by this wavefront's next tile
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                });
                pb.flush();
                if (!SYNC)
                {
                    // the next queue's smallest distance: one atomic per workgroup
                    for (int o = kWave / 2; o > 0; o >>= 1)
                    {
                        const std::uint32_t other = __shfl_xor(qmin, o, kWave);
                        qmin = other < qmin ? other : qmin;
                    }
                    if (lane == 0)
                        s_min[wid] = qmin;
                    __syncthreads();
                    if (threadIdx.x == 0)
                    {
                        std::uint32_t m = s_min[0];
                        for (int k = 1; k < kPWaves; ++k)
                            m = s_min[k] < m ? s_min[k] : m;
                        if (m != 0xFFFFFFFFU)
                            atomicMin(&ctl->qmin[s1], m);
                    }
                }
                grid_barrier(ctl, ++barriers);
            }
            if (lane == 0)
            {
                if (light_e)
                    atomicAdd(&ctl->light_edges, light_e);
                if (chunk_e)
                    atomicAdd(&ctl->chunk_edges, chunk_e);
            }
        }

        template <class OffT>
        __global__ void persist_seed(PersistArgs<OffT> a, std::uint32_t src)
        {
            if (threadIdx.x == 0 && blockIdx.x == 0)
            {
                a.dist[src] = 0U;
                a.q[0][0] = src;
                a.ctl->rows[0] = 1U;
                a.ctl->qmin[0] = 0U;
                a.ctl->qmin[1] = 0xFFFFFFFFU;
                a.ctl->qmin[2] = 0xFFFFFFFFU;
            }
        }

        PersistWs &persist_ws(DeviceGraph &g, std::uint64_t heavy_cap)
        {
            auto *p = static_cast<PersistWs *>(g.pws.get());
            if (p == nullptr || p->V != g.V || p->heavy_cap < heavy_cap)
            {
                g.pws.reset();
                auto w = std::make_shared<PersistWs>();
                const std::uint64_t Vp = g.V + kPad;
                w->stamp.alloc(Vp * 4, "hipMalloc(persistent stamps)");
                w->qb.alloc(Vp * 4, "hipMalloc(persistent queue)");
                w->heavy_b.alloc(heavy_cap * sizeof(HeavyChunk), "hipMalloc(persistent chunks)");
                w->ctl.alloc(sizeof(PersistCtl), "hipMalloc(persistent control)");
                w->V = g.V;
                w->heavy_cap = heavy_cap;
                g.pws = w;
                p = w.get();
            }
            return *p;
        }

        template <class OffT, int WM, bool SYNC>
        void run_persistent(DeviceGraph &g, std::uint32_t src, const hyt_run_options &o, std::uint32_t light_max,
                            std::uint32_t chunk, hyt_run_stats &st)
        {
            Workspace &ws = g.ws;
            hipStream_t s = g.stream;
            PersistWs &p = persist_ws(g, ws.heavy_cap);
            const std::uint64_t cap = std::min<std::uint64_t>(ws.heavy_cap, p.heavy_cap);
            PersistArgs<OffT> a{};
            a.off = g.off.as<OffT>();
            a.col = g.col.as<std::uint32_t>();
            a.w = g.w.as<std::uint32_t>();
            a.dist = ws.dist.as<std::uint32_t>();
            a.stamp = p.stamp.as<std::uint32_t>();
            a.rec = ws.prev.as<std::uint32_t>();
            a.qval = ws.qd.as<std::uint32_t>();
            a.q[0] = ws.q.as<std::uint32_t>();
            a.q[1] = p.qb.as<std::uint32_t>();
            a.chunks[0] = ws.heavy.as<HeavyChunk>();
            a.chunks[1] = p.heavy_b.as<HeavyChunk>();
            a.queue_cap = g.V + kPad;
            a.chunk_cap = cap;
            a.ctl = p.ctl.as<PersistCtl>();
            a.light_max = light_max;
            a.chunk = chunk;
            // waiting rows only when the caller sets heavy_delta (measured
            // slower here by default, unlike the captured rounds)
            a.heavy_delta = SYNC || o.heavy_delta == 0 ? 0U : heavy_delta_of(o.heavy_delta);
            a.wait_min = std::min(light_max, wait_degree_of(o.wait_degree, light_max));
            a.max_rounds = o.max_rounds ? o.max_rounds : g.V;
            if (a.max_rounds > 0x7FFFFFF0ULL)
                a.max_rounds = 0x7FFFFFF0ULL; // round numbers (stamps) and barrier generations are u32

            launch_fill<std::uint32_t>(a.dist, g.V, Weights<WM>::kInf, s);
            launch_fill<std::uint32_t>(a.stamp, g.V, 0U, s);
            launch_fill<std::uint32_t>(a.rec, g.V, 0xFFFFFFFFU, s);
            static_assert(sizeof(PersistCtl) % 4 == 0, "PersistCtl is whole words");
            launch_fill<std::uint32_t>(reinterpret_cast<std::uint32_t *>(a.ctl), sizeof(PersistCtl) / 4, 0U, s);
            persist_seed<OffT><<<1, 64, 0, s>>>(a, src);
            hip_check(hipGetLastError(), "launch persist_seed");

            // the resident grid: one workgroup per CU (the cooperative launch
            // fails rather than run a grid that is not co-resident)
            const void *fn = reinterpret_cast<const void *>(&sssp_persistent<OffT, WM, SYNC>);
            int per_cu = 0, cus = 0;
            hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kPBlock, 0),
                      "occupancy of the persistent kernel");
            hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g.device),
                      "multiprocessor count");
            require(per_cu >= 1 && cus >= 1, "the persistent kernel cannot be resident");
            const unsigned grid = static_cast<unsigned>(cus * std::min(per_cu, kPBlocksPerCu));
            void *args[] = {&a};
            if (o.time_kernels)
                hip_check(hipEventRecord(g.ev0, s), "hipEventRecord");
            hip_check(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(kPBlock), args, 0, s),
                      "launch persistent SSSP");
            if (o.time_kernels)
                hip_check(hipEventRecord(g.ev1, s), "hipEventRecord");
            PersistCtl h{};
            hip_check(hipMemcpyAsync(&h, a.ctl, sizeof(PersistCtl), hipMemcpyDeviceToHost, s), "read persistent control");
            hip_check(hipStreamSynchronize(s), "persistent SSSP");
            if (o.time_kernels)
            {
                float ms = 0;
                hip_check(hipEventElapsedTime(&ms, g.ev0, g.ev1), "hipEventElapsedTime");
                st.expand_ms += ms;
                st.expand_launches += 1;
            }
            if (h.queue_overflow)
                fail(HYT_ERR_INTERNAL, "persistent SSSP queue overflow");
            st.rounds = h.rounds;
            st.frontier_total = h.frontier;
            st.updates_total = h.frontier - 1;
            st.edges_total = h.light_edges + h.chunk_edges;
            st.heavy_edges = h.chunk_edges;
            st.converged = h.truncated ? 0 : 1;
            if (h.overflow)
            {
                st.overflow = 1;
                fail(HYT_ERR_OVERFLOW, "u32 distance sum exceeded 2^32-2");
            }
        }

        template <class OffT, int WM>
        void dispatch_mode(DeviceGraph &g, std::uint32_t src, const hyt_run_options &o, std::uint32_t light_max,
                           std::uint32_t chunk, hyt_run_stats &st)
        {
            if (o.mode == HYT_SSSP_ASYNC)
                run_persistent<OffT, WM, false>(g, src, o, light_max, chunk, st);
            else
                run_persistent<OffT, WM, true>(g, src, o, light_max, chunk, st);
        }

        template <class OffT>
        void dispatch_weights(DeviceGraph &g, std::uint32_t src, const hyt_run_options &o, std::uint32_t light_max,
                              std::uint32_t chunk, hyt_run_stats &st)
        {
            switch (g.weight_mode())
            {
            case kWU32: dispatch_mode<OffT, kWU32>(g, src, o, light_max, chunk, st); break;
            case kWF32: dispatch_mode<OffT, kWF32>(g, src, o, light_max, chunk, st); break;
            case kWUnitU32: dispatch_mode<OffT, kWUnitU32>(g, src, o, light_max, chunk, st); break;
            default: dispatch_mode<OffT, kWUnitF32>(g, src, o, light_max, chunk, st); break;
            }
        }
    } // namespace

    void sssp_persistent_rounds(DeviceGraph &g, std::uint32_t src, const hyt_run_options &o, std::uint32_t light_max,
                                std::uint32_t chunk, hyt_run_stats &st)
    {
        require(o.tolerance == 0.0F && o.delta == 0 && o.direction != HYT_DIR_PULL,
                "persistent SSSP runs push rounds without tolerance or buckets");
        if (g.off64)
            dispatch_weights<std::uint64_t>(g, src, o, light_max, chunk, st);
        else
            dispatch_weights<std::uint32_t>(g, src, o, light_max, chunk, st);
    }
} // namespace hyt
