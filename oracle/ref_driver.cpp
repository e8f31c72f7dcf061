// ref_driver.cpp - extern "C" entry points over the ORIGINAL project's CPU
// functions, compiled from the original sources where they lie (see
// oracle/Makefile, target `ref`). TEST INFRASTRUCTURE ONLY: used to pin the
// oracle restatement and as the cpu_baseline of bench.py ("reference" kind).
//
// Wraps: sssp_cpu (src/algorithms/sssp.cpp:109), pagerank_cpu
// (src/algorithms/pagerank.cpp:89), via CSRGraph (src/graph/csr_graph.cpp:11).
#include "algorithms/pagerank.hpp"
#include "algorithms/sssp.hpp"
#include "graph/csr_graph.hpp"

#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <vector>

using hytgraph::graph::CSRGraph;

namespace
{
    CSRGraph make_graph(std::uint64_t V, const std::uint64_t *off, const std::uint32_t *col, const float *w)
    {
        const std::uint64_t E = off[V];
        std::vector<CSRGraph::offset_type> o(off, off + V + 1);
        std::vector<CSRGraph::vertex_id> c(col, col + E);
        std::vector<CSRGraph::weight_type> ww;
        if (w)
            ww.assign(w, w + E);
        return CSRGraph(V, std::move(o), std::move(c), std::move(ww));
    }

    int code_of(const std::exception_ptr &ep)
    {
        try
        {
            std::rethrow_exception(ep);
        }
        catch (const std::invalid_argument &)
        {
            return 1;
        }
        catch (...)
        {
            return 7;
        }
    }
} // namespace

extern "C" {

// Returns 0, 1 (std::invalid_argument) or 7 (any other exception).
int ref_sssp_cpu(std::uint64_t V, const std::uint64_t *off, const std::uint32_t *col, const float *w,
                 std::uint32_t source, float tol, std::uint64_t max_iter, float *out, std::uint64_t *iterations,
                 int *converged)
{
    try
    {
        const CSRGraph g = make_graph(V, off, col, w);
        hytgraph::algorithms::SSSPOptions o;
        o.source = source;
        o.tolerance = tol;
        o.max_iterations = max_iter;
        const auto r = hytgraph::algorithms::sssp_cpu(g, o);
        std::memcpy(out, r.distances.data(), r.distances.size() * sizeof(float));
        *iterations = r.iterations;
        *converged = r.converged ? 1 : 0;
        return 0;
    }
    catch (...)
    {
        return code_of(std::current_exception());
    }
}

int ref_pagerank_cpu(std::uint64_t V, const std::uint64_t *off, const std::uint32_t *col, float damping, float tol,
                     std::uint64_t max_iter, float *out, std::uint64_t *iterations, float *residual, int *converged)
{
    try
    {
        const CSRGraph g = make_graph(V, off, col, nullptr);
        hytgraph::algorithms::PageRankOptions o;
        o.damping_factor = damping;
        o.tolerance = tol;
        o.max_iterations = max_iter;
        const auto r = hytgraph::algorithms::pagerank_cpu(g, o);
        std::memcpy(out, r.ranks.data(), r.ranks.size() * sizeof(float));
        *iterations = r.iterations;
        *residual = r.residual;
        *converged = r.converged ? 1 : 0;
        return 0;
    }
    catch (...)
    {
        return code_of(std::current_exception());
    }
}

} // extern "C"
