/*
 * oracle.c - CPU restatement of the original project's hot-path algorithms.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker of the HIP path (never as the thing measured or shipped).
 *
 * Functions and what pins each one (the tests name the pin):
 *   orc_sssp_jacobi_f32   restates sssp_cpu (src/algorithms/sssp.cpp:21-80,
 *                         109-234). Pinned by the original's known answers
 *                         (tests/algorithm_tests.cpp:100-157), bit-exact
 *                         against the compiled original (oracle/_ref), and
 *                         against Dijkstra / scipy on random graphs.
 *   orc_sssp_dijkstra     textbook Dijkstra, u64 sums. Pinned against
 *                         scipy.sparse.csgraph.dijkstra.
 *   orc_pagerank_jacobi_f32 restates pagerank_cpu (src/algorithms/
 *                         pagerank.cpp:19-60, 89-163). Pinned bit-exact
 *                         against oracle/_ref and by the closed forms of
 *                         tests/algorithm_tests.cpp:22-75.
 *   orc_pagerank_f64      FP64 power iteration of the same fixed point (PPR:
 *                         teleport and dangling mass to the seed). Pinned by
 *                         a dense linear solve on small graphs.
 *   orc_delta_push_f64    FP64 restatement of the synchronous delta-push
 *                         procedure the GPU runs. Pinned by convergence to
 *                         orc_pagerank_f64 as eps -> 0 and by its mass
 *                         balance invariant.
 *   orc_hub_sort          restates HubSorter::sort (src/scheduling/
 *                         hub_sort.cpp:28-183). Pinned by the original's
 *                         known answers (tests/unit_tests.cpp:2468-2602,
 *                         2710-2750) and bit-exact against oracle/_ref.
 *   orc_reorder           restates CSRGraph::reorder_vertices (src/graph/
 *                         csr_graph.cpp:171-283). Pinned by the original's
 *                         known answers (tests/unit_tests.cpp:2604-2708),
 *                         against oracle/_ref, and by the invariant that
 *                         SSSP distances permute with the vertex ids.
 *   orc_pagerank_contributions restates pagerank_contributions (src/
 *                         algorithms/pagerank.cpp:165-243). Bit-exact against
 *                         oracle/_ref, and equal to |step - x| of
 *                         orc_pagerank_jacobi_f32's first iteration from x =
 *                         1/n (the same float step).
 *   orc_sssp_contributions restates sssp_contributions (src/algorithms/
 *                         sssp.cpp:236-290). Bit-exact against oracle/_ref
 *                         and by hand-worked cases (inf, NaN, tolerance).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_ENOMEM 4
#define ORC_EOVERFLOW 5

/* Incoming CSR by counting sort over sources in ascending order, so each
 * destination's in-edges are ordered by (source, out-edge position) - the
 * order of build_incoming_csr (sssp.cpp:21-80, pagerank.cpp:19-60). */
static int build_incoming(uint64_t V, const uint64_t *off, const uint32_t *col, uint64_t **in_off,
                          uint32_t **in_src, uint64_t **in_edge)
{
    const uint64_t E = off[V];
    uint64_t *io = calloc(V + 1, sizeof(uint64_t));
    uint32_t *is = malloc((E ? E : 1) * sizeof(uint32_t));
    uint64_t *ie = malloc((E ? E : 1) * sizeof(uint64_t));
    uint64_t *cur = malloc((V + 1) * sizeof(uint64_t));
    if (!io || !is || !ie || !cur) {
        free(io), free(is), free(ie), free(cur);
        return ORC_ENOMEM;
    }
    for (uint64_t e = 0; e < E; ++e)
        io[(uint64_t)col[e] + 1] += 1;
    for (uint64_t v = 1; v <= V; ++v)
        io[v] += io[v - 1];
    memcpy(cur, io, (V + 1) * sizeof(uint64_t));
    for (uint64_t u = 0; u < V; ++u)
        for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
            uint64_t p = cur[col[e]]++;
            is[p] = (uint32_t)u;
            ie[p] = e;
        }
    free(cur);
    *in_off = io;
    *in_src = is;
    *in_edge = ie;
    return ORC_OK;
}

/* sssp_cpu (sssp.cpp:109-234): synchronous Bellman-Ford over incoming edges,
 * float distances; a source contributes only if it improved in the previous
 * round (or is the start vertex) and its distance is finite; the strict test
 * candidate + tol < best runs in incoming-edge order; w = NULL means 1.0f;
 * max_iter 0 means V. */
int orc_sssp_jacobi_f32(uint64_t V, const uint64_t *off, const uint32_t *col, const float *w, uint32_t source,
                        float tol, uint64_t max_iter, float *out, uint64_t *iterations, int *converged)
{
    if (V == 0 || source >= V || !(tol >= 0.0f) || !isfinite(tol))
        return ORC_EINVAL;
    const uint64_t E = off[V];
    if (w)
        for (uint64_t e = 0; e < E; ++e)
            if (!isfinite(w[e]) || w[e] < 0.0f)
                return ORC_EINVAL;
    uint64_t *io, *ie;
    uint32_t *is;
    if (build_incoming(V, off, col, &io, &is, &ie))
        return ORC_ENOMEM;
    float *cur = malloc(V * sizeof(float)), *nxt = malloc(V * sizeof(float));
    unsigned char *act = calloc(V, 1), *nact = calloc(V, 1);
    if (!cur || !nxt || !act || !nact) {
        free(io), free(is), free(ie), free(cur), free(nxt), free(act), free(nact);
        return ORC_ENOMEM;
    }
    for (uint64_t v = 0; v < V; ++v)
        cur[v] = INFINITY;
    cur[source] = 0.0f;
    act[source] = 1;
    if (max_iter == 0)
        max_iter = V;
    *converged = 0;
    *iterations = 0;
    for (uint64_t it = 0; it < max_iter; ++it) {
        int changed = 0;
        memset(nact, 0, V);
        for (uint64_t v = 0; v < V; ++v) {
            float best = cur[v];
            for (uint64_t p = io[v]; p < io[v + 1]; ++p) {
                uint32_t s = is[p];
                if (!act[s])
                    continue;
                float sd = cur[s];
                if (!isfinite(sd))
                    continue;
                float cand = sd + (w ? w[ie[p]] : 1.0f);
                if (cand + tol < best)
                    best = cand;
            }
            nxt[v] = best;
            if (best + tol < cur[v]) {
                nact[v] = 1;
                changed = 1;
            }
        }
        float *tf = cur;
        cur = nxt;
        nxt = tf;
        unsigned char *ta = act;
        act = nact;
        nact = ta;
        *iterations = it + 1;
        if (!changed) {
            *converged = 1;
            break;
        }
    }
    memcpy(out, cur, V * sizeof(float));
    free(io), free(is), free(ie), free(cur), free(nxt), free(act), free(nact);
    return ORC_OK;
}

/* ---- Dijkstra with a binary heap of (dist, vertex), lazy deletion ---- */
typedef struct {
    uint64_t d;
    uint32_t v;
} hnode;

static void heap_push(hnode *h, uint64_t *n, hnode x)
{
    uint64_t i = (*n)++;
    while (i > 0) {
        uint64_t p = (i - 1) / 2;
        if (h[p].d <= x.d)
            break;
        h[i] = h[p];
        i = p;
    }
    h[i] = x;
}

static hnode heap_pop(hnode *h, uint64_t *n)
{
    hnode top = h[0], x = h[--(*n)];
    uint64_t i = 0;
    for (;;) {
        uint64_t l = 2 * i + 1, r = l + 1, m = i;
        hnode best = x;
        if (l < *n && h[l].d < best.d)
            m = l, best = h[l];
        if (r < *n && h[r].d < best.d)
            m = r;
        if (m == i)
            break;
        h[i] = h[m];
        i = m;
    }
    if (*n > 0)
        h[i] = x;
    return top;
}

/* Exact shortest-path distances with integer weights (w = NULL: all 1).
 * out: u32, UINT32_MAX = unreachable. Fails with ORC_EOVERFLOW when a
 * reachable distance is >= 2^32-1. */
int orc_sssp_dijkstra(uint64_t V, const uint64_t *off, const uint32_t *col, const uint32_t *w, uint32_t source,
                      uint32_t *out)
{
    if (V == 0 || source >= V)
        return ORC_EINVAL;
    const uint64_t E = off[V];
    uint64_t *d = malloc(V * sizeof(uint64_t));
    hnode *h = malloc((E + 1) * sizeof(hnode));
    unsigned char *done = calloc(V, 1);
    if (!d || !h || !done) {
        free(d), free(h), free(done);
        return ORC_ENOMEM;
    }
    for (uint64_t v = 0; v < V; ++v)
        d[v] = UINT64_MAX;
    uint64_t n = 0;
    d[source] = 0;
    heap_push(h, &n, (hnode){0, source});
    while (n > 0) {
        hnode t = heap_pop(h, &n);
        if (done[t.v])
            continue;
        done[t.v] = 1;
        for (uint64_t e = off[t.v]; e < off[t.v + 1]; ++e) {
            uint64_t nd = t.d + (w ? w[e] : 1u);
            uint32_t x = col[e];
            if (nd < d[x]) {
                d[x] = nd;
                heap_push(h, &n, (hnode){nd, x});
            }
        }
    }
    int rc = ORC_OK;
    for (uint64_t v = 0; v < V; ++v) {
        if (d[v] == UINT64_MAX)
            out[v] = UINT32_MAX;
        else if (d[v] >= UINT32_MAX)
            rc = ORC_EOVERFLOW, out[v] = UINT32_MAX;
        else
            out[v] = (uint32_t)d[v];
    }
    free(d), free(h), free(done);
    return rc;
}

/* pagerank_cpu (pagerank.cpp:89-163): x0 = 1/n; per iteration the float
 * dangling mass (ascending vertex order), then for each destination the float
 * sum of cur[src]/(float)outdeg(src) in incoming order, next = base +
 * dangling_share + d*sum, float L1 residual; stop when residual <= tol. */
int orc_pagerank_jacobi_f32(uint64_t V, const uint64_t *off, const uint32_t *col, float damping, float tol,
                            uint64_t max_iter, float *out, uint64_t *iterations, float *residual_out,
                            int *converged)
{
    if (V == 0 || !(damping >= 0.0f && damping < 1.0f) || !(tol > 0.0f) || !isfinite(tol) || max_iter == 0)
        return ORC_EINVAL;
    uint64_t *io, *ie;
    uint32_t *is;
    if (build_incoming(V, off, col, &io, &is, &ie))
        return ORC_ENOMEM;
    float *cur = malloc(V * sizeof(float)), *nxt = malloc(V * sizeof(float));
    if (!cur || !nxt) {
        free(io), free(is), free(ie), free(cur), free(nxt);
        return ORC_ENOMEM;
    }
    const float n = (float)V;
    const float base = (1.0f - damping) / n;
    for (uint64_t v = 0; v < V; ++v)
        cur[v] = 1.0f / n;
    *converged = 0;
    *iterations = 0;
    *residual_out = 0.0f;
    for (uint64_t it = 0; it < max_iter; ++it) {
        float dangling = 0.0f;
        for (uint64_t u = 0; u < V; ++u)
            if (off[u + 1] == off[u])
                dangling += cur[u];
        const float share = damping * dangling / n;
        float residual = 0.0f;
        for (uint64_t v = 0; v < V; ++v) {
            float sum = 0.0f;
            for (uint64_t p = io[v]; p < io[v + 1]; ++p) {
                uint32_t s = is[p];
                uint64_t deg = off[s + 1] - off[s];
                if (deg != 0)
                    sum += cur[s] / (float)deg;
            }
            nxt[v] = base + share + damping * sum;
            residual += fabsf(nxt[v] - cur[v]);
        }
        float *t = cur;
        cur = nxt;
        nxt = t;
        *iterations = it + 1;
        *residual_out = residual;
        if (residual <= tol) {
            *converged = 1;
            break;
        }
    }
    memcpy(out, cur, V * sizeof(float));
    free(io), free(is), free(ie), free(cur), free(nxt);
    return ORC_OK;
}

/* FP64 power iteration of x = (1-d) t + d (P~^T x + M t), where t is uniform
 * (seed = -1, PageRank) or e_seed (PPR) and M is the dangling mass; pushes
 * along out-edges (order does not matter in FP64 at the tolerances used).
 * Runs until the L1 change is <= tol or max_iter. out sums to 1. */
int orc_pagerank_f64(uint64_t V, const uint64_t *off, const uint32_t *col, double damping, int64_t seed,
                     double tol, uint64_t max_iter, double *out, uint64_t *iterations)
{
    if (V == 0 || !(damping >= 0.0 && damping < 1.0) || (seed >= 0 && (uint64_t)seed >= V))
        return ORC_EINVAL;
    double *cur = malloc(V * sizeof(double)), *nxt = malloc(V * sizeof(double));
    if (!cur || !nxt) {
        free(cur), free(nxt);
        return ORC_ENOMEM;
    }
    for (uint64_t v = 0; v < V; ++v)
        cur[v] = seed < 0 ? 1.0 / (double)V : (v == (uint64_t)seed ? 1.0 : 0.0);
    *iterations = 0;
    for (uint64_t it = 0; it < max_iter; ++it) {
        double dangling = 0.0;
        memset(nxt, 0, V * sizeof(double));
        for (uint64_t u = 0; u < V; ++u) {
            uint64_t deg = off[u + 1] - off[u];
            if (deg == 0) {
                dangling += cur[u];
                continue;
            }
            double share = damping * cur[u] / (double)deg;
            for (uint64_t e = off[u]; e < off[u + 1]; ++e)
                nxt[col[e]] += share;
        }
        double tele = (1.0 - damping) + damping * dangling;
        if (seed < 0) {
            double add = tele / (double)V;
            for (uint64_t v = 0; v < V; ++v)
                nxt[v] += add;
        } else {
            nxt[seed] += tele;
        }
        double diff = 0.0;
        for (uint64_t v = 0; v < V; ++v)
            diff += fabs(nxt[v] - cur[v]);
        double *t = cur;
        cur = nxt;
        nxt = t;
        *iterations = it + 1;
        if (diff <= tol)
            break;
    }
    double s = 0.0;
    for (uint64_t v = 0; v < V; ++v)
        s += cur[v];
    for (uint64_t v = 0; v < V; ++v)
        out[v] = cur[v] / s;
    free(cur), free(nxt);
    return ORC_OK;
}

/* FP64 synchronous delta-push, the procedure of the HIP kernels: residual
 * (1-d) at every vertex (seed < 0) or at the seed; per round, every vertex
 * with residual > eps (snapshotted at the start of the round) moves it into
 * rank and sends d*r/outdeg to each out-neighbour; dangling residual is
 * dropped. Result (rank + residual) normalised to sum 1. */
int orc_delta_push_f64(uint64_t V, const uint64_t *off, const uint32_t *col, double damping, double eps,
                       int64_t seed, uint64_t max_rounds, double *out, uint64_t *rounds,
                       uint64_t *edges_pushed)
{
    if (V == 0 || !(damping >= 0.0 && damping < 1.0) || !(eps > 0.0) || (seed >= 0 && (uint64_t)seed >= V))
        return ORC_EINVAL;
    double *res = calloc(V, sizeof(double)), *rank = calloc(V, sizeof(double)),
           *snap = malloc(V * sizeof(double));
    uint32_t *q = malloc(V * sizeof(uint32_t));
    if (!res || !rank || !snap || !q) {
        free(res), free(rank), free(snap), free(q);
        return ORC_ENOMEM;
    }
    uint64_t F = 0;
    if (seed < 0) {
        for (uint64_t v = 0; v < V; ++v)
            res[v] = 1.0 - damping;
    } else {
        res[seed] = 1.0 - damping;
    }
    for (uint64_t v = 0; v < V; ++v)
        if (res[v] > eps)
            q[F++] = (uint32_t)v;
    *rounds = 0;
    *edges_pushed = 0;
    if (max_rounds == 0)
        max_rounds = 100000;
    while (F > 0 && *rounds < max_rounds) {
        for (uint64_t i = 0; i < F; ++i) {
            uint32_t u = q[i];
            snap[i] = res[u];
            rank[u] += res[u];
            res[u] = 0.0;
        }
        for (uint64_t i = 0; i < F; ++i) {
            uint32_t u = q[i];
            uint64_t deg = off[u + 1] - off[u];
            if (deg == 0)
                continue;
            double share = damping * snap[i] / (double)deg;
            for (uint64_t e = off[u]; e < off[u + 1]; ++e)
                res[col[e]] += share;
            *edges_pushed += deg;
        }
        F = 0;
        for (uint64_t v = 0; v < V; ++v)
            if (res[v] > eps)
                q[F++] = (uint32_t)v;
        *rounds += 1;
    }
    double s = 0.0;
    for (uint64_t v = 0; v < V; ++v)
        s += rank[v] + res[v];
    for (uint64_t v = 0; v < V; ++v)
        out[v] = s > 0.0 ? (rank[v] + res[v]) / s : 0.0;
    free(res), free(rank), free(snap), free(q);
    return ORC_OK;
}

/* HubSorter::sort (src/scheduling/hub_sort.cpp:28-183): in-degrees by
 * counting; score = outdeg * indeg / (max outdeg * max indeg) evaluated in
 * long double (x87 extended on x86-64) and rounded to double, 0 when the
 * denominator is 0; hub_count = V * fraction rounded to the nearest integer
 * when within 1e-12 of it, else up, at most V; order = the first hub_count
 * vertices by (score desc, id asc), then the rest in id order. */
static const double *hub_score_key;

static int hub_before(const void *pa, const void *pb)
{
    const uint32_t a = *(const uint32_t *)pa, b = *(const uint32_t *)pb;
    const double sa = hub_score_key[a], sb = hub_score_key[b];
    if (sa != sb)
        return sa > sb ? -1 : 1;
    return a < b ? -1 : (a > b ? 1 : 0);
}

int orc_hub_sort(uint64_t V, const uint64_t *off, const uint32_t *col, double fraction, uint32_t *order,
                 double *scores, uint64_t *hub_count)
{
    if (!isfinite(fraction) || fraction < 0.0 || fraction > 1.0)
        return ORC_EINVAL;
    *hub_count = 0;
    if (V == 0)
        return ORC_OK;
    const uint64_t E = off[V];
    uint64_t *indeg = calloc(V, sizeof(uint64_t));
    uint32_t *rank = malloc(V * sizeof(uint32_t));
    unsigned char *hub = calloc(V, 1);
    if (!indeg || !rank || !hub) {
        free(indeg), free(rank), free(hub);
        return ORC_ENOMEM;
    }
    for (uint64_t e = 0; e < E; ++e)
        indeg[col[e]] += 1;
    uint64_t max_out = 0, max_in = 0;
    for (uint64_t v = 0; v < V; ++v) {
        const uint64_t d = off[v + 1] - off[v];
        if (d > max_out)
            max_out = d;
        if (indeg[v] > max_in)
            max_in = indeg[v];
    }
    const long double den = (long double)max_out * (long double)max_in;
    for (uint64_t v = 0; v < V; ++v)
        scores[v] = den > 0.0L ? (double)((long double)(off[v + 1] - off[v]) * (long double)indeg[v] / den) : 0.0;
    const long double want = (long double)V * (long double)fraction;
    const long double nearest = roundl(want);
    const long double count = fabsl(want - nearest) < 1e-12L ? nearest : ceill(want);
    uint64_t hubs = (uint64_t)count;
    if (hubs > V)
        hubs = V;
    for (uint64_t v = 0; v < V; ++v)
        rank[v] = (uint32_t)v;
    hub_score_key = scores;
    qsort(rank, V, sizeof(uint32_t), hub_before);
    uint64_t at = 0;
    for (uint64_t r = 0; r < hubs; ++r) {
        hub[rank[r]] = 1;
        order[at++] = rank[r];
    }
    for (uint64_t v = 0; v < V; ++v)
        if (!hub[v])
            order[at++] = (uint32_t)v;
    *hub_count = hubs;
    free(indeg), free(rank), free(hub);
    return ORC_OK;
}

/* CSRGraph::reorder_vertices (src/graph/csr_graph.cpp:171-283): order[new] =
 * old must be a permutation (the first position holding an id >= V or an id
 * seen before decides the error: *error_kind 1 = out of range, 2 =
 * duplicate); new row `new` is old row order[new] with columns renamed by
 * the inverse permutation, weights (u32 bit patterns, may be NULL) carried
 * in edge order. */
int orc_reorder(uint64_t V, const uint64_t *off, const uint32_t *col, const uint32_t *w, const uint32_t *order,
                uint64_t *off2, uint32_t *col2, uint32_t *w2, int *error_kind)
{
    *error_kind = 0;
    unsigned char *seen = calloc(V ? V : 1, 1);
    uint32_t *inv = malloc((V ? V : 1) * sizeof(uint32_t));
    if (!seen || !inv) {
        free(seen), free(inv);
        return ORC_ENOMEM;
    }
    for (uint64_t i = 0; i < V; ++i) {
        const uint32_t o = order[i];
        if (o >= V) {
            *error_kind = 1;
            break;
        }
        if (seen[o]) {
            *error_kind = 2;
            break;
        }
        seen[o] = 1;
        inv[o] = (uint32_t)i;
    }
    if (*error_kind) {
        free(seen), free(inv);
        return ORC_EINVAL;
    }
    uint64_t at = 0;
    off2[0] = 0;
    for (uint64_t n = 0; n < V; ++n) {
        const uint32_t u = order[n];
        for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
            col2[at] = inv[col[e]];
            if (w)
                w2[at] = w[e];
            ++at;
        }
        off2[n + 1] = at;
    }
    free(seen), free(inv);
    return ORC_OK;
}

/* pagerank_contributions (src/algorithms/pagerank.cpp:165-243): one
 * synchronous step from x with the float arithmetic of pagerank_cpu (dangling
 * mass in ascending vertex order, in-edge sums in incoming order); out[v] =
 * |next(v) - x[v]| as double. */
int orc_pagerank_contributions(uint64_t V, const uint64_t *off, const uint32_t *col, float damping, const float *x,
                               double *out)
{
    if (V == 0 || !(damping >= 0.0f && damping < 1.0f))
        return ORC_EINVAL;
    uint64_t *io, *ie;
    uint32_t *is;
    if (build_incoming(V, off, col, &io, &is, &ie))
        return ORC_ENOMEM;
    const float n = (float)V;
    const float base = (1.0f - damping) / n;
    float dangling = 0.0f;
    for (uint64_t u = 0; u < V; ++u)
        if (off[u + 1] == off[u])
            dangling += x[u];
    const float share = damping * dangling / n;
    for (uint64_t v = 0; v < V; ++v) {
        float sum = 0.0f;
        for (uint64_t p = io[v]; p < io[v + 1]; ++p) {
            const uint32_t s = is[p];
            const uint64_t deg = off[s + 1] - off[s];
            if (deg != 0)
                sum += x[s] / (float)deg;
        }
        const float next = base + share + damping * sum;
        out[v] = (double)fabsf(next - x[v]);
    }
    free(io), free(is), free(ie);
    return ORC_OK;
}

/* sssp_contributions (src/algorithms/sssp.cpp:236-290): current - candidate
 * (float) when both are finite and the gain exceeds tol; 1 when only the
 * candidate is finite; else 0. */
int orc_sssp_contributions(uint64_t n, const float *cur, const float *cand, float tol, double *out)
{
    if (!(tol >= 0.0f) || !isfinite(tol))
        return ORC_EINVAL;
    for (uint64_t i = 0; i < n; ++i) {
        double value = 0.0;
        if (isfinite(cur[i]) && isfinite(cand[i])) {
            const float gain = cur[i] - cand[i];
            if (gain > tol)
                value = (double)gain;
        } else if (!isfinite(cur[i]) && isfinite(cand[i])) {
            value = 1.0;
        }
        out[i] = value;
    }
    return ORC_OK;
}
