// Lane sessions over the hybrid engine (used by the continuous-batching
// loop): a session's queries are lanes of the engine's [V][K] state, and one
// step runs one K-lane round of the engine (plan, move or map partitions,
// expand, mark compaction in vertex order). Admission seeds the lane and
// marks the source row; the next round first folds the current queue back
// into the marks and compacts them in vertex order, because the planner
// needs the rows of each partition to be one range of the queue.
#pragma once

#include "dgraph.hpp"

struct hyt_hybrid;

namespace hyt
{
    struct HybridEngine;

    HybridEngine &hybrid_engine(hyt_hybrid *h);
    // Offsets-only device graph of the engine (V, E, max_degree, stream).
    DeviceGraph &hybrid_device_graph(HybridEngine &E);
    // HBM the engine holds while it serves, outside its lane state (edge
    // area, offsets, partition tables, compaction staging, queue): opening a
    // session releases the single-query workspace.
    std::uint64_t hybrid_engine_bytes(HybridEngine &E);
    // HBM of the lane state for K lanes of `kind` (planning).
    std::uint64_t hybrid_lane_state_bytes(HybridEngine &E, int kind, std::uint32_t K);

    struct HybridLaneView
    {
        void *val;                // [V][K]: u32 distances, or u64 PPR residual words (mass.hpp)
        unsigned long long *rank; // [V][K] PPR pushed mass words (null for SSSP)
        unsigned long long *mark; // per row: lanes whose row awaits the next round
    };
    // Allocates the lane state for K lanes of `kind` and frees every lane.
    HybridLaneView hybrid_lanes_open(HybridEngine &E, int kind, std::uint32_t K);

    struct HybridLaneRound
    {
        Counters c;          // next_count = next queue rows, frontier_edges = its (row, lane) pairs
        std::uint64_t lanes; // lanes present in the next queue
        std::uint32_t F;     // queue rows the round expanded (after folding admissions in)
        std::uint64_t pairs; // their (row, lane) pairs
        bool items;          // the round ran the item form
    };
    // One round over queue rows [0, F) (with `pairs` (row, lane) pairs);
    // merge = admissions marked rows since the last round.
    HybridLaneRound hybrid_lanes_round(HybridEngine &E, int kind, std::uint32_t K, std::uint32_t F,
                                       std::uint64_t pairs, bool merge, float damping, float eps,
                                       std::uint32_t form, hyt_hybrid_stats &st);
} // namespace hyt
