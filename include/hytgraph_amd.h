/*
 * hytgraph_amd.h - C ABI of the MI355X (gfx950) graph engine.
 *
 * Plain C types only: pointers, sizes and POD option/stat structs. Every
 * entry point returns an int status (HYT_OK on success); on failure
 * hyt_last_error() returns a thread-local message naming the operation.
 *
 * The C++ drop-in layer (include/algorithms/{sssp,pagerank}.hpp) is built on
 * these calls and maps status codes to the exception types the original
 * project throws (std::invalid_argument / std::runtime_error).
 *
 * Threading: thread-compatible. A hyt_graph_t owns a HIP stream and a
 * workspace; concurrent calls on the same graph handle race, calls on
 * different handles do not.
 */
#ifndef HYTGRAPH_AMD_H
#define HYTGRAPH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum hyt_status {
    HYT_OK = 0,
    HYT_ERR_INVALID_ARGUMENT = 1, /* maps to std::invalid_argument */
    HYT_ERR_NO_DEVICE = 2,        /* no HIP device visible */
    HYT_ERR_HIP = 3,              /* a HIP runtime call failed */
    HYT_ERR_OUT_OF_MEMORY = 4,    /* allocation failed or budget exceeded */
    HYT_ERR_OVERFLOW = 5,         /* a u32 distance sum exceeded 2^32-2 */
    HYT_ERR_UNSUPPORTED = 6,
    HYT_ERR_INTERNAL = 7,
    HYT_ERR_IO = 8                /* a file could not be opened, read or written */
};

/* Thread-local message of the last failing call on this thread. */
const char *hyt_last_error(void);

/* Library version string. */
const char *hyt_version(void);

/* Number of visible HIP devices (0 when none; never an error). */
int hyt_device_count(int *count);

/* --------------------------------------------------- host graph generator */
/*
 * Synthetic graph generator (no counterpart in the original project, whose
 * datasets/ directory is empty). Counter-based: edge i of a graph is a pure
 * function of (seed, i), so the output is identical for any thread count.
 * Self-loops and duplicate (src,dst) pairs are removed; rows are sorted by
 * destination; weights are drawn per final CSR position.
 */
enum hyt_gen_kind { HYT_GEN_RMAT = 0, HYT_GEN_ER = 1 };
enum hyt_host_memory { HYT_HOST_PAGEABLE = 0, HYT_HOST_PINNED = 1 };

typedef struct hyt_gen_params {
    int32_t kind;          /* hyt_gen_kind */
    uint32_t scale;        /* V = 2^scale, 1..32 */
    uint32_t edge_factor;  /* generated edges = edge_factor * V (before cleanup) */
    double a, b, c;        /* RMAT quadrant probabilities (d = 1-a-b-c) */
    uint64_t seed;
    uint32_t weight_min;   /* weights uniform in [weight_min, weight_max]; */
    uint32_t weight_max;   /* both 0 = unweighted graph */
    int32_t threads;       /* 0 = min(hardware threads, 16) */
    int32_t memory;        /* hyt_host_memory: pinned uses hipHostMalloc */
} hyt_gen_params;

typedef struct hyt_host_graph *hyt_host_graph_t;

int hyt_generate(const hyt_gen_params *params, hyt_host_graph_t *out);

/*
 * The same graph as hyt_generate, bit for bit, built on the current HIP
 * device (no counterpart in the original): edges are drawn on the device in
 * source-range buckets that fit device_budget_bytes of HBM (0 = 40 GiB, about
 * 17 bytes per generated edge of the largest bucket), each bucket sorted and
 * deduplicated there and copied into a host graph of kind params->memory.
 * params->threads is not used. RMAT scale 30 (17 B generated edges) takes
 * seconds where the host generator takes tens of minutes. Fails with
 * HYT_ERR_OUT_OF_MEMORY when one source bin (1/16384 of the vertex range)
 * holds more edges than a bucket.
 */
int hyt_generate_device(const hyt_gen_params *params, uint64_t device_budget_bytes, hyt_host_graph_t *out);

/* Borrowed views of the host CSR: row_offsets u64[V+1], columns u32[E],
 * weights u32[E] (NULL when unweighted). Valid until hyt_host_graph_free. */
int hyt_host_graph_view(hyt_host_graph_t g, uint64_t *num_vertices,
                        uint64_t *num_edges, const uint64_t **row_offsets,
                        const uint32_t **columns, const uint32_t **weights);
/* The incoming CSR of a host graph (the transpose the original builds for
 * its pull kernels, build_incoming_csr, src/algorithms/sssp.cpp:21-80 and
 * cuda/pagerank.cu): in_offsets u64[V+1], in_sources u32[E], the sources of
 * each destination in ascending order (edge order within a source), in the
 * graph's memory kind. Built on first call with up to `threads` host threads
 * (0: up to 16) and kept with the graph. The hybrid engine's PageRank pull
 * rounds build the same arrays on the device when the graph (E < 2^31) and
 * the sort fit in free HBM, else this way. */
int hyt_host_graph_transpose(hyt_host_graph_t g, uint32_t threads, const uint64_t **in_offsets,
                             const uint32_t **in_sources);
int hyt_host_graph_free(hyt_host_graph_t g);

/*
 * Graph files and caller graphs as host graphs (memory: hyt_host_memory;
 * pinned graphs can feed the hybrid engine).
 *
 * hyt_load_edge_list: the original's text edge list
 * (src/graph/graph_loader.cpp:39-179, GraphLoader::load_edge_list): one
 * "src dst" (or "src dst weight") per line, blank and '#' lines skipped, ids in
 * [0, num_vertices), extra tokens rejected; edges grouped by source in input
 * order, parallel edges kept. Weights are stored as float bit patterns.
 * Same messages as the original; an unreadable file is HYT_ERR_IO.
 *
 * hyt_host_graph_save / _load: binary CSR file (no counterpart in the
 * original, which reads text only): a 128-byte header with V, E, flags and a
 * checksum per array, then offsets u64[V+1], columns u32[E], weights u32[E].
 * Read and written with positional I/O on several threads; verify = 1
 * recomputes the checksums and checks the CSR structure. weights_f32 records
 * whether the weights are float bit patterns.
 *
 * hyt_host_graph_from_csr: copies caller CSR arrays (validated as
 * CSRGraph::validate does, src/graph/csr_graph.cpp:105-169).
 */
int hyt_load_edge_list(const char *path, uint64_t num_vertices, uint32_t weighted, int32_t memory,
                       hyt_host_graph_t *out);
int hyt_host_graph_save(hyt_host_graph_t g, const char *path, uint32_t weights_f32);
int hyt_host_graph_load(const char *path, int32_t memory, uint32_t verify, hyt_host_graph_t *out,
                        uint32_t *weights_f32);
int hyt_host_graph_from_csr(uint64_t num_vertices, uint64_t num_edges, const uint64_t *row_offsets,
                            const uint32_t *columns, const uint32_t *weights, int32_t memory,
                            hyt_host_graph_t *out);

/* Degree-weighted source sampling: pick edge index uniformly with a
 * counter-based stream keyed by seed and return its source vertex. Vertices
 * with no out-edges are never picked. */
int hyt_sample_sources(const uint64_t *row_offsets, uint64_t num_vertices,
                       uint64_t seed, uint32_t count, uint32_t *out_sources);

/* ------------------------------------------------- device-resident graph */
enum hyt_graph_flags {
    HYT_GRAPH_WEIGHTS_U32 = 0,   /* weights are unsigned integers */
    HYT_GRAPH_WEIGHTS_F32 = 1,   /* weights are IEEE float bit patterns (>= +0) */
    HYT_GRAPH_OFF64 = 2          /* OR-ed in: u64 device row offsets whatever E is
                                    (they are always used from E >= 2^32 - 1) */
};

typedef struct hyt_graph *hyt_graph_t;

/*
 * Upload a CSR graph into HBM on the current device. row_offsets is u64[V+1]
 * (the original CSRGraph offset type); the device copy uses u32 offsets when
 * E < 2^32 and u64 otherwise. weights may be NULL (unweighted: every edge
 * weighs 1). The CSR invariants are validated on the host first
 * (offsets[0]=0, monotone, offsets[V]=E, columns < V).
 */
int hyt_graph_upload(uint64_t num_vertices, uint64_t num_edges,
                     const uint64_t *row_offsets, const uint32_t *columns,
                     const uint32_t *weights, uint32_t flags,
                     hyt_graph_t *out);
int hyt_graph_free(hyt_graph_t g);
/* Releases the graph's traversal workspaces and its transpose (they are
 * rebuilt by the next call that needs them); the CSR stays. */
int hyt_graph_trim(hyt_graph_t g);
int hyt_graph_info(hyt_graph_t g, uint64_t *num_vertices, uint64_t *num_edges,
                   uint64_t *device_bytes);

/* Copies the device CSR to the host: row_offsets u64[V+1], columns u32[E],
 * weights u32[E] (NULL to skip; an unweighted graph has none). */
int hyt_graph_download(hyt_graph_t g, uint64_t *row_offsets, uint32_t *columns, uint32_t *weights);

/* ------------------------------------------------- vertex order (hubs) */
/*
 * HubSorter::sort of the original (src/scheduling/hub_sort.cpp:28-183) on the
 * device graph: score[v] = outdeg(v) * indeg(v) / (max outdeg * max indeg),
 * with the original's long double division (x87 extended precision, then
 * double) reproduced bit for bit; the first hub_count = V * hub_fraction
 * (rounded to the nearest integer when within 1e-12 of it, else up)
 * vertices by (score descending, vertex ascending), then the others in
 * vertex order. vertex_order: u32[V], scores: f64[V] or NULL, both host or
 * both device memory (out_on_device). Needs V < 2^31 and max outdeg * max
 * indeg < 2^64. hub_fraction outside [0, 1] or not finite:
 * HYT_ERR_INVALID_ARGUMENT with the original's message.
 */
int hyt_hub_sort(hyt_graph_t g, double hub_fraction, uint32_t *vertex_order, double *scores,
                 uint64_t *hub_count, uint32_t out_on_device);
/*
 * CSRGraph::reorder_vertices of the original (src/graph/csr_graph.cpp:
 * 171-283) on the device graph, in place: vertex_order[new] = old (u32[V],
 * host or device memory per order_on_device); new row `new` is old row
 * `old` with each column c renamed to its new id, weights carried, edge
 * order kept. An order that is not a permutation fails with the original's
 * message for the first offending position (out-of-range id or duplicate)
 * and leaves the graph unchanged. Drops the transpose and workspaces. Needs
 * V < 2^31.
 */
int hyt_graph_reorder(hyt_graph_t g, const uint32_t *vertex_order, uint32_t order_on_device);

/* -------------------------------------------------------- contributions */
/*
 * pagerank_contributions of the original (src/algorithms/pagerank.cpp:
 * 165-243) on the device: contributions[v] = |next(v) - current[v]| for one
 * synchronous PageRank step from current (float sums in the original's
 * order, so the values equal the original's bit for bit). current: f32[V],
 * contributions: f64[V], both host or both device memory (on_device).
 * Builds the graph's transpose on first use (E < 2^31).
 */
int hyt_pagerank_contributions(hyt_graph_t g, const float *current_ranks, float damping,
                               double *contributions, uint32_t on_device);
/*
 * sssp_contributions of the original (src/algorithms/sssp.cpp:236-290) on
 * the current device: contributions[v] = current - candidate when both are
 * finite and the gain exceeds tolerance, 1 when only the candidate is
 * finite, else 0. Arrays of n elements, host or device memory (on_device).
 */
int hyt_sssp_contributions(const float *current, const float *candidate, uint64_t n, float tolerance,
                           double *contributions, uint32_t on_device);

/* ------------------------------------------------------------------ SSSP */
enum hyt_sssp_mode {
    /* Synchronous rounds: a round relaxes the out-edges of the vertices
     * improved in the previous round using the distances they held at the
     * end of that round. Reproduces the original sssp_cpu/sssp_gpu
     * iteration count exactly. */
    HYT_SSSP_SYNC = 0,
    /* Rounds read live distances (relaxations of this round are visible);
     * same final distances, fewer rounds and relaxations. */
    HYT_SSSP_ASYNC = 1
};

enum hyt_direction {
    HYT_DIR_AUTO = 0, /* per round: pull when the frontier's out-edges are a large share of E */
    HYT_DIR_PUSH = 1, /* always push along out-edges */
    HYT_DIR_PULL = 2  /* always pull along in-edges (builds the transpose on first use) */
};

typedef struct hyt_run_options {
    uint32_t mode;           /* hyt_sssp_mode (SSSP only) */
    uint32_t out_on_device;  /* 1: output pointer is device memory */
    uint64_t max_rounds;     /* 0: |V| (the original default) */
    uint32_t time_kernels;   /* 1: time the expand kernels with HIP events */
    uint32_t light_max_degree; /* 0: default; vertices above are chunked */
    uint32_t heavy_chunk;      /* 0: default edges per wave chunk */
    uint32_t direction;        /* hyt_direction (single-source SSSP and PageRank) */
    uint32_t pull_percent;     /* AUTO: pull when frontier out-edges >= pct% of E; 0 = default */
    uint32_t delta;            /* ASYNC SSSP: delta-stepping bucket width in weight units
                                  (0 = plain rounds); SYNC mode ignores it */
    uint32_t lane_form;        /* batched (K >= 2): hyt_lane_form */
    float tolerance;           /* single-source SYNC SSSP on float weights: the
                                  original's rule, v takes candidate c only when
                                  c + tolerance < best, in incoming-edge order
                                  (sssp.cpp:199-216, cuda/sssp.cu:131-153); > 0 runs
                                  ordered pull rounds over the transpose */
    uint32_t host_rounds;      /* single-source SSSP push rounds: 0 = replays of
                                  captured rounds with device-side round control
                                  (one read-back per 4 rounds); 1 = one host
                                  read-back and launch sequence per round */
    uint32_t task_order;       /* hybrid single-query runs: 0 = filter runs in
                                  partition order; 1 = by contribution (the
                                  original's ContributionScheduler order over
                                  partition contributions computed on the
                                  device: SSSP -min frontier distance, PageRank
                                  the mass the partition pushes) */
    uint32_t heavy_delta;      /* ASYNC single-source SSSP: rows above light_max
                                  (the hubs) wait while their distance exceeds
                                  the smallest distance of the round's queue by
                                  more than heavy_delta weight units, so they
                                  go out (nearly) once, with their final
                                  distance; light rows never wait. 0 = the
                                  default, 2 (with persistent: no waiting);
                                  UINT32_MAX = no waiting. Same distances
                                  either way */
    uint32_t persistent;       /* single-source SSSP push rounds without
                                  tolerance or buckets: 1 = one persistent
                                  cooperative launch runs every round (the
                                  next queue built by the relaxations, grid
                                  barriers between rounds, no compaction pass
                                  and no host read-back); SYNC keeps the
                                  original's iteration count. 0 = the rounds
                                  host_rounds selects */
    uint32_t wait_degree;      /* ASYNC single-source SSSP: light rows with more
                                  edges than this wait under the heavy_delta
                                  bound as well (0 = light_max: only the rows
                                  above light_max wait). Same distances */
} hyt_run_options;

/* Work split of the K-lane rounds (batched queries and lane sessions). */
enum hyt_lane_form {
    HYT_LANES_AUTO = 0,  /* per round: items when the queue's (row, query) pairs
                            fill under half of the group form's lanes */
    HYT_LANES_GROUP = 1, /* next_pow2(K) lanes per edge, one per query */
    HYT_LANES_ITEMS = 2  /* one lane per (edge, active query) pair */
};

typedef struct hyt_run_stats {
    uint64_t rounds;          /* iterations in the original's sense */
    int32_t converged;        /* 1 when the last round changed nothing */
    int32_t overflow;         /* 1 when a u32 sum overflowed (status OVERFLOW) */
    uint64_t frontier_total;  /* sum over rounds of frontier rows (F) */
    uint64_t edges_total;     /* sum over rounds of expanded edges (R) */
    uint64_t updates_total;   /* sum over rounds of next-frontier appends (U) */
    double expand_ms;         /* sum of expand kernel times (time_kernels=1) */
    uint64_t expand_launches; /* number of timed expand launches */
    double wall_ms;           /* host wall time of the call */
    double residual;          /* PageRank: normalised L1 of leftover residual */
    uint64_t pull_rounds;     /* rounds run as pull over the transpose */
    uint64_t pull_active_edges; /* pull rounds: in-edges whose source was active */
    uint64_t pull_edges;      /* pull rounds: in-edges scanned (part of edges_total) */
    uint64_t pull_frontier;   /* pull rounds: frontier rows (part of frontier_total) */
    uint64_t heavy_edges;     /* push rounds: edges walked as chunks of rows above
                                 light_max (part of edges_total) */
} hyt_run_stats;

/*
 * Single- or multi-source SSSP. sources: K vertex ids (host memory).
 * dist: u32[V][K] (row-major, K contiguous per vertex); UINT32_MAX marks an
 * unreachable vertex. On an F32-weight graph the u32 values are float bit
 * patterns (+inf = 0x7F800000), i.e. the float distances of the original.
 * K = 1 runs the single-source kernels; 2 <= K <= 64 runs the [V][K] batched
 * kernels, which use async semantics (rounds = rounds of the batch).
 */
int hyt_sssp(hyt_graph_t g, const uint32_t *sources, uint32_t num_sources,
             uint32_t *dist, const hyt_run_options *opts, hyt_run_stats *stats);

/*
 * No counterpart in the original: hyt_sssp for one source for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). It calls no hipMalloc/hipFree: the run's
 * scratch (distances, record, queue, chunk queue, counters) is a static
 * 128 MiB device array per GPU, held until the process ends, so graphs up to
 * about V = 2^22 (E <= 2^28) run; larger ones fail with
 * HYT_ERR_INVALID_ARGUMENT. dist must be device memory (opts->out_on_device
 * = 1); rounds are push-only (no pull direction, tolerance 0). Calls on one
 * GPU share the array: calls that overlap on one GPU race, so serialise them.
 * Distances and stats equal hyt_sssp's for the same options.
 */
int hyt_sssp_noalloc(hyt_graph_t g, uint32_t source, uint32_t *dist,
                     const hyt_run_options *opts, hyt_run_stats *stats);

/* ------------------------------------------------- PageRank / PPR (FP32) */
/*
 * Delta-push PageRank. Residuals start at (1-damping) per vertex; a round
 * pushes every vertex whose residual exceeds epsilon. Dangling mass is
 * dropped and the result is normalised to sum 1, which is the fixed point of
 * the original's uniform-dangling Jacobi iteration. ranks: float[V].
 */
int hyt_pagerank(hyt_graph_t g, float damping, float epsilon, float *ranks,
                 const hyt_run_options *opts, hyt_run_stats *stats);

/*
 * Personalized PageRank for K seeds, [V][K] FP32 state; residual (1-damping)
 * at each seed; result columns normalised to sum 1. out: float[V][K].
 */
int hyt_ppr(hyt_graph_t g, const uint32_t *seeds, uint32_t num_seeds,
            float damping, float epsilon, float *out,
            const hyt_run_options *opts, hyt_run_stats *stats);

/* Measurement: `reps` launches of a read-only streaming kernel over a
 * device buffer of `bytes` (allocated and freed by the call); gbs = the read
 * rate. The reference for rocprofv3 byte counters: each launch reads
 * exactly `bytes` from HBM. */
int hyt_stream_probe(uint64_t bytes, uint32_t reps, double *gbs);

/* ------------------------------------------------ round-by-round runs */
/*
 * A single-source SSSP run (hyt_sssp with one source) or a PageRank / PPR
 * run (hyt_pagerank, or hyt_ppr with one seed) advanced `rounds` rounds per
 * hyt_stepper_step and read between steps: the execution behind the SEP
 * drivers' execute_step() (include/sep/sep_execution_driver.hpp; the
 * original's drivers are deferred placeholders, its GPU functions only run
 * to convergence). Rounds and results equal the one-call forms: a SYNC SSSP
 * stepped one round at a time passes through the states of the original's
 * max_iterations = 1, 2, ... runs. A stepper runs on its graph's workspace:
 * one stepper per graph at a time, no other call on the graph while it lives.
 * vertex: the SSSP source, or the PPR seed (UINT32_MAX: global PageRank).
 * opts: mode, direction, split, host_rounds (no tolerance or buckets).
 */
typedef struct hyt_stepper *hyt_stepper_t;
enum hyt_step_kind { HYT_STEP_SSSP = 0, HYT_STEP_PAGERANK = 1 };
int hyt_stepper_create(hyt_graph_t g, uint32_t kind, uint32_t vertex, float damping, float epsilon,
                       const hyt_run_options *opts, hyt_stepper_t *out);
/* stats: this call's rounds; finished = 1 once a round found nothing left to
 * do (later calls run nothing). */
int hyt_stepper_step(hyt_stepper_t s, uint64_t rounds, hyt_run_stats *stats, int32_t *finished);
/* out: u32[V] distances (float bit patterns on F32 graphs) or f32[V]
 * normalised ranks, host or device per opts->out_on_device; total: the
 * run's stats so far (PageRank: residual = unpushed mass / total). */
int hyt_stepper_read(hyt_stepper_t s, void *out, hyt_run_stats *total);
int hyt_stepper_free(hyt_stepper_t s);

/* ---------------------------------------------- hybrid transfer planning */
/*
 * The HyTM planning steps of the original project as plain functions over
 * arrays (src/graph/partition.cpp:105-174, src/transfer/hytm_cost_model.cpp:
 * 233-715, src/scheduling/task_combiner.cpp:67-248). The device transfer
 * engine runs the same code on counts it gathers on the GPU.
 */
enum hyt_engine { HYT_ENGINE_FILTER = 0, HYT_ENGINE_COMPACTION = 1, HYT_ENGINE_ZERO_COPY = 2 };

typedef struct hyt_cost_params {
    double alpha, beta, gamma;       /* selector thresholds, zero-copy RTT mix */
    uint64_t request_bytes;          /* m  (bytes per memory request) */
    uint64_t requests_per_tlp;       /* MR (outstanding requests per TLP) */
    double rtt;                      /* saturated round trip, seconds or any unit */
    double cpu_bytes_per_second;     /* host compaction throughput */
    uint64_t entry_bytes;            /* d1: bytes per transferred edge entry */
    uint64_t index_bytes;            /* d2: bytes per compacted vertex index entry */
    uint64_t alignment_bytes;        /* zero-copy alignment boundary */
} hyt_cost_params;

typedef struct hyt_partition {
    uint32_t vertex_begin, vertex_end;   /* [begin, end) */
    uint64_t edge_begin, edge_end;       /* [begin, end) */
} hyt_partition;

typedef struct hyt_partition_metrics {
    uint64_t active_vertices, active_edges, total_edges;
    uint64_t filter_bytes, compaction_bytes;
    uint64_t zc_requests, zc_alignment, zc_total_requests;
} hyt_partition_metrics;

typedef struct hyt_partition_costs {
    double filter, compaction, zero_copy, zc_rtt;
    int32_t engine;                      /* hyt_engine */
    int32_t pad;
} hyt_partition_costs;

/* The original's defaults (alpha .8, beta .4, gamma .625, m 128, MR 256,
 * rtt 1, cpu 1 B/s, d1 4, d2 8, align 128). */
int hyt_cost_params_default(hyt_cost_params *out);

/* Greedy contiguous partitioning: close the range before vertex v when its
 * column bytes (4 B per edge) would exceed target_bytes; a lone oversized
 * row stays whole. *count receives the number written (<= capacity). */
int hyt_partition_graph(const uint64_t *row_offsets, uint64_t num_vertices, uint64_t target_bytes,
                        hyt_partition *out, uint64_t capacity, uint64_t *count);

/* Host metric collection for a per-vertex activity byte array. */
int hyt_partition_metrics_host(const uint64_t *row_offsets, uint64_t num_vertices, const uint8_t *active,
                               const hyt_partition *parts, uint64_t n, const hyt_cost_params *params,
                               hyt_partition_metrics *out);

/* Cost equations (1)-(3) and the strict selector, per partition. */
int hyt_cost_evaluate(const hyt_cost_params *params, const hyt_partition_metrics *metrics, uint64_t n,
                      hyt_partition_costs *out);
int hyt_select_engine(double filter_cost, double compaction_cost, double zero_copy_cost, double alpha,
                      double beta, int32_t *engine);

/* Task combination: filter runs of <= k consecutive partitions, one task of
 * all compaction partitions, one of all zero-copy partitions, ordered by first
 * partition. Task t has engine task_engine[t] and members
 * members[task_offset[t] .. task_offset[t+1]). Arrays hold n (+1) entries. */
int hyt_combine_tasks(const int32_t *engines, uint64_t n, uint64_t k, int32_t *task_engine,
                      uint64_t *task_offset, uint64_t *members, uint64_t *num_tasks);

/* ------------------------------------------------- hybrid transfer engine */
/*
 * Out-of-HBM traversal (the HyTGraph hybrid transfer management): the CSR
 * stays in pinned, device-mapped host memory; vertex state and row offsets
 * live in HBM; each round the engine gathers per-partition activity on the
 * GPU, runs the cost model and task combiner, and moves the active
 * partitions' edges by the chosen engine:
 *   filter      whole partition copied (hipMemcpyAsync on a copy stream) into
 *               an HBM slot; slots double as an LRU cache across rounds;
 *   compaction  host threads pack the active rows' edges into pinned staging,
 *               one copy, then the kernel walks the packed rows;
 *   zero-copy   the kernel reads the mapped host arrays directly.
 * No counterpart in the original runs these movers (its engines only plan).
 */
typedef struct hyt_hybrid *hyt_hybrid_t;

typedef struct hyt_hybrid_config {
    uint64_t device_budget_bytes; /* HBM for edge data: cache slots + packed staging */
    uint64_t partition_edges;     /* partition target (edges; 2^20 -> 4 MiB of columns) */
    uint64_t pack_edges;          /* packed-staging capacity in edges (0: default) */
    uint32_t filter_combine_k;    /* task combiner k (0: 4) */
    int32_t force_engine;         /* -1: cost model; else hyt_engine for every partition */
    uint32_t host_threads;        /* compaction packers (0: min(hw, 16)) */
    uint32_t calibrate;           /* 1: measure host link, packing, zero-copy; set rtt/cpu */
    uint32_t cache;               /* 1: filtered partitions stay resident (LRU) */
    uint32_t flags;               /* HYT_GRAPH_OFF64: u64 device row offsets whatever E is */
    hyt_cost_params cost;         /* used as given when calibrate == 0 */
} hyt_hybrid_config;

typedef struct hyt_hybrid_stats {
    hyt_run_stats run;
    uint64_t partitions;          /* logical partitions of the graph */
    uint64_t parts_cached;        /* partition visits served from the HBM cache */
    uint64_t parts_filter, parts_compaction, parts_zero_copy; /* visits per engine */
    uint64_t tasks;               /* executable tasks over the run */
    uint64_t bytes_filter;        /* host->device bytes copied for filter */
    uint64_t bytes_compaction;    /* host->device bytes copied for compaction */
    uint64_t bytes_zero_copy;     /* estimated: 128 B x zero-copy requests */
    double copy_ms;               /* copy-stream busy time (HIP events) */
    double pack_ms;               /* host packing time */
    double plan_ms;               /* activity readback + cost model + combine */
    double link_gbs_peak;         /* measured pinned H2D bandwidth */
    double cpu_pack_gbs;          /* measured host packing throughput */
    double zero_copy_gbs;         /* measured zero-copy kernel read bandwidth */
    uint64_t device_edge_bytes;   /* HBM used for edge data (<= budget) */
    uint64_t device_state_bytes;  /* HBM used for offsets and vertex state */
    uint64_t runs_reordered;      /* filter runs moved by the contribution order */
    double transpose_ms;          /* PageRank pull rounds: host build of the incoming
                                     CSR, in the call that built it (else 0) */
} hyt_hybrid_stats;

int hyt_hybrid_config_default(hyt_hybrid_config *out);

/* The host graph must outlive the engine. Pinned graphs (HYT_HOST_PINNED)
 * get all three movers; pageable graphs (HYT_HOST_PAGEABLE: graphs larger
 * than what can be pinned) get filter and compaction only - the cost model
 * takes zero-copy as infinitely expensive, filter copies are staged through
 * pinned bounce buffers, and force_engine = HYT_ENGINE_ZERO_COPY is
 * rejected. weights_f32: weights are float bit patterns. */
int hyt_hybrid_create(hyt_host_graph_t host, uint32_t weights_f32, const hyt_hybrid_config *config,
                      hyt_hybrid_t *out);
int hyt_hybrid_free(hyt_hybrid_t h);
int hyt_hybrid_info(hyt_hybrid_t h, uint64_t *partitions, hyt_cost_params *params_in_use,
                    hyt_hybrid_stats *calibration);

/* Single-source SSSP (async rounds) and delta-push PageRank over the hybrid
 * engine; outputs as hyt_sssp / hyt_pagerank. PageRank honours
 * opts->direction and pull_percent as hyt_pagerank does: a pull round reads
 * the host graph's incoming CSR (hyt_host_graph_transpose, built on the first
 * pull round) - every in-partition, resident or copied whole - and adds the
 * same words as a push round, so the ranks are the same bit for bit. */
int hyt_hybrid_sssp(hyt_hybrid_t h, uint32_t source, uint32_t *dist, const hyt_run_options *opts,
                    hyt_hybrid_stats *stats);
int hyt_hybrid_pagerank(hyt_hybrid_t h, float damping, float epsilon, float *ranks,
                        const hyt_run_options *opts, hyt_hybrid_stats *stats);

/*
 * K queries per run over the hybrid engine (1 <= K <= 64): multi-source SSSP
 * and personalized PageRank with [V][K] state in HBM, outputs as hyt_sssp /
 * hyt_ppr with K sources. The planner sees each frontier row once whatever
 * its query mask, so every partition that moves (or is read zero-copy) is
 * walked once for all K queries. The [V][K] state is outside the edge budget,
 * like the single-query state. No counterpart in the original project, whose
 * transfer engines plan one query (src/transfer/filter_engine.cpp:146-220,
 * compaction_engine.cpp:253-493, zero_copy_engine.cpp:119-393).
 */
int hyt_hybrid_sssp_k(hyt_hybrid_t h, const uint32_t *sources, uint32_t num_sources, uint32_t *dist,
                      const hyt_run_options *opts, hyt_hybrid_stats *stats);
int hyt_hybrid_ppr_k(hyt_hybrid_t h, const uint32_t *seeds, uint32_t num_seeds, float damping,
                     float epsilon, float *out, const hyt_run_options *opts, hyt_hybrid_stats *stats);

/* ------------------------------------------- continuous batching (serving) */
/*
 * No counterpart in the original project (which answers one query per call):
 * these serve many SSSP / personalized-PageRank queries on one GPU by running
 * them as lanes of shared [V][K] rounds and admitting new queries between
 * rounds, for request loops that batch queries as they arrive.
 *
 * A lane session holds K <= 64 query slots of one kind over one HBM-resident
 * graph (which must outlive it), with its own state and stream. Admit a
 * source into a free lane, step (each step runs rounds for every occupied
 * lane of every given session; sessions overlap on the device), and read a
 * lane once it has finished. Results equal the single-query forms: SSSP
 * distances bit-exact, PPR within the delta-push tolerance, normalised to 1.
 * A finished lane keeps its result until read and released (or re-admitted
 * after hyt_lanes_release).
 */
enum hyt_query_kind { HYT_QUERY_SSSP = 0, HYT_QUERY_PPR = 1 };

typedef struct hyt_lanes *hyt_lanes_t;

/* Device bytes of a session of `lanes` lanes (admission planning);
 * max_degree = largest out-degree of the graph. */
int hyt_lanes_state_bytes(uint64_t num_vertices, uint64_t num_edges, uint64_t max_degree, uint32_t kind,
                          uint32_t lanes, uint64_t *bytes);
/* damping / epsilon are used by PPR sessions only. */
int hyt_lanes_create(hyt_graph_t g, uint32_t kind, uint32_t lanes, float damping, float epsilon,
                     hyt_lanes_t *out);
int hyt_lanes_free(hyt_lanes_t s);
/* The lane must be free; the query starts with the next round. At most 64
 * admissions between two rounds. */
int hyt_lanes_admit(hyt_lanes_t s, uint32_t lane, uint32_t source);
/* Runs up to `rounds` rounds on every session (stops early when all are
 * idle); finished[i] = lanes of session i that finished during the call,
 * running[i] (optional) = lanes still running. */
int hyt_lanes_step(hyt_lanes_t *sessions, uint32_t count, uint32_t rounds, uint64_t *finished,
                   uint64_t *running);
/* Result of a finished lane: u32[V] distances (SSSP) or float[V] ranks (PPR)
 * into out (host or device, optional); reached = entries that are not
 * unreachable (SSSP) / not zero (PPR), checksum = u64 sum of their bit
 * patterns; rounds = rounds the query ran. All outputs optional. */
int hyt_lanes_read(hyt_lanes_t s, uint32_t lane, void *out, uint32_t out_on_device, uint64_t *reached,
                   uint64_t *checksum, uint32_t *rounds);
/* Frees a finished lane. */
int hyt_lanes_release(hyt_lanes_t s, uint32_t lane);
/* hyt_lane_form of the session's rounds (default HYT_LANES_AUTO). */
int hyt_lanes_set_form(hyt_lanes_t s, uint32_t form);
int hyt_lanes_info(hyt_lanes_t s, uint32_t *lanes, uint64_t *occupied, uint64_t *running, uint32_t *frontier,
                   uint64_t *state_bytes);

typedef struct hyt_request {
    uint32_t graph;     /* index into the graph array */
    uint32_t kind;      /* hyt_query_kind */
    uint32_t source;
    uint32_t reserved;
    double arrival_s;   /* seconds after the start of the trace (non-decreasing) */
} hyt_request;

typedef struct hyt_completion {
    double admit_s, done_s; /* seconds after the start of the trace */
    uint32_t session, lane, rounds, reserved;
    uint64_t reached, checksum; /* as hyt_lanes_read */
} hyt_completion;

typedef struct hyt_serve_config {
    uint64_t device_budget_bytes; /* resident graphs + every session's state */
    uint32_t max_lanes;           /* per session, <= 64 */
    uint32_t lane_form;           /* hyt_lane_form of every session */
    float damping, epsilon;       /* PPR */
    uint32_t reserved[4];
} hyt_serve_config;

typedef struct hyt_serve_stats {
    double wall_s, busy_s;        /* trace wall time; time with a round in flight */
    uint64_t steps, rounds, edges, admitted; /* steps = session rounds completed by the loop */
    uint64_t graph_bytes, state_bytes;
    uint64_t frontier_rows, lane_rows; /* frontier rows expanded / (row, query) pairs among them */
    uint64_t item_rounds;         /* rounds run with one lane per (edge, query) pair (sparse masks) */
    uint32_t sessions;
    uint32_t lanes[16];           /* lanes of the first 16 sessions (one per graph and kind, in order of first request) */
    uint32_t reserved;
    /* sessions over hybrid engines: bytes copied over the host link, bytes
     * read zero-copy (estimated), partition visits from the HBM cache and
     * visits that moved or mapped a partition */
    uint64_t host_bytes, host_zero_copy_bytes, parts_cached, parts_moved;
} hyt_serve_stats;

int hyt_serve_config_default(hyt_serve_config *out); /* 64 GiB, 64 lanes, 0.85, 1e-6 */

/*
 * Replays a request trace against the wall clock: arrivals queue per (graph,
 * kind) session, sessions get lanes from the device budget, and each session
 * runs its own rounds: between two of them its finished queries are read and
 * queued ones are admitted into the free lanes.
 * results: optional per-request host buffers of V words (NULL entries or a
 * NULL array skip the copy). completions: one per request.
 */
int hyt_serve_trace(const hyt_graph_t *graphs, uint32_t num_graphs, const hyt_serve_config *config,
                    const hyt_request *requests, uint64_t num_requests, void *const *results,
                    hyt_completion *completions, hyt_serve_stats *stats);

/*
 * The same loop over graphs that stay in host memory: a served graph is
 * either resident in HBM, or (resident NULL) read through hybrid engines over
 * its pinned host graph, one engine per query kind the trace uses (their edge
 * areas and vertex state count against the device budget, like resident
 * graphs). A session over an engine keeps its lanes in the engine's state and
 * admits queries between the engine's rounds; each round moves, maps or finds
 * cached every active partition once for all of its lanes. While it serves,
 * an engine must not be used otherwise.
 */
typedef struct hyt_serve_graph {
    hyt_graph_t resident;
    hyt_hybrid_t sssp_engine;
    hyt_hybrid_t ppr_engine;
} hyt_serve_graph;

int hyt_serve_trace_mixed(const hyt_serve_graph *graphs, uint32_t num_graphs, const hyt_serve_config *config,
                          const hyt_request *requests, uint64_t num_requests, void *const *results,
                          hyt_completion *completions, hyt_serve_stats *stats);

/* A lane session over a hybrid engine (the engine's lane state holds its
 * lanes; one session per engine at a time); every other hyt_lanes_* call
 * works on it as on an HBM session, hyt_lanes_step runs its rounds after
 * launching those of the HBM sessions. */
int hyt_lanes_create_hybrid(hyt_hybrid_t h, uint32_t kind, uint32_t lanes, float damping, float epsilon,
                            hyt_lanes_t *out);

/* ------------------------------------------------------------- autotuner */
/*
 * No counterpart in the original project: measures candidate settings on
 * the device (each candidate timed on the caller's sources, best of `reps`)
 * and keeps the fastest.
 *   degree-bin thresholds: light_max_degree in {64..2048} then heavy_chunk
 *     in {256..8192} for single-source SSSP (coordinate descent from the
 *     defaults);
 *   lanes: K of batched SSSP with the highest queries/s (K cycles through
 *     the sources when there are fewer than K);
 *   partition size (hyt_tune_hybrid): partition_edges of the hybrid engine
 *     with the shortest SSSP wall time under the given configuration.
 * Results never change answers, only speed.
 */
typedef struct hyt_tune_options {
    uint32_t tune_split;   /* 1: tune light_max_degree / heavy_chunk */
    uint32_t tune_lanes;   /* 1: tune K (2..max_lanes) */
    uint32_t max_lanes;    /* 0 = 64 */
    uint32_t reps;         /* timed repetitions per candidate (0 = 2) */
} hyt_tune_options;

typedef struct hyt_tune_result {
    uint32_t light_max_degree, heavy_chunk; /* fastest single-source split */
    uint32_t lanes;                         /* K with the highest batched queries/s */
    uint32_t candidates;                    /* candidates measured */
    double sssp_ms;                         /* best mean single-source time per query */
    double default_sssp_ms;                 /* the same with the default split */
    double lanes_qps;                       /* batched queries/s at the chosen K */
    uint64_t partition_edges;               /* hyt_tune_hybrid only */
    double hybrid_ms;                       /* hyt_tune_hybrid: best SSSP wall time per query */
} hyt_tune_result;

int hyt_tune(hyt_graph_t g, const uint32_t *sources, uint32_t num_sources, const hyt_tune_options *opts,
             hyt_tune_result *out);
/* candidates: partition sizes in edges (n >= 1); config: every other setting. */
int hyt_tune_hybrid(hyt_host_graph_t host, uint32_t weights_f32, const hyt_hybrid_config *config,
                    const uint64_t *candidates, uint32_t num_candidates, const uint32_t *sources,
                    uint32_t num_sources, hyt_tune_result *out);

#ifdef __cplusplus
}
#endif

#endif /* HYTGRAPH_AMD_H */
