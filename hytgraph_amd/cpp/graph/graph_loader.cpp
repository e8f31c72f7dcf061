// GraphLoader over the C-ABI graph-file entry points (graph_io.cpp): the
// loaders fill a library host graph, which is copied into CSRGraph's vectors.
#include "graph/graph_loader.hpp"

#include "hytgraph_amd.h"

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace hytgraph::graph
{
    namespace
    {
        // The C-ABI message is "<entry point>: <message>"; the C++ interface
        // reports the message itself.
        [[noreturn]] void raise(int status, const char *entry)
        {
            std::string msg = hyt_last_error();
            const std::string prefix = std::string(entry) + ": ";
            if (msg.compare(0, prefix.size(), prefix) == 0)
                msg.erase(0, prefix.size());
            if (status == HYT_ERR_INVALID_ARGUMENT)
                throw std::invalid_argument(msg);
            throw std::runtime_error(msg);
        }

        struct HostGraphHandle
        {
            hyt_host_graph_t h = nullptr;
            ~HostGraphHandle()
            {
                if (h)
                    (void)hyt_host_graph_free(h);
            }
        };

        CSRGraph to_csr(hyt_host_graph_t h, bool weights_are_f32)
        {
            std::uint64_t V = 0, E = 0;
            const std::uint64_t *off = nullptr;
            const std::uint32_t *col = nullptr, *w = nullptr;
            if (const int st = hyt_host_graph_view(h, &V, &E, &off, &col, &w); st != HYT_OK)
                raise(st, "hyt_host_graph_view");
            std::vector<CSRGraph::offset_type> o(off, off + V + 1);
            std::vector<CSRGraph::vertex_id> c(col, col + E);
            std::vector<CSRGraph::weight_type> wt;
            if (w != nullptr)
            {
                wt.resize(E);
                if (weights_are_f32)
                    std::memcpy(wt.data(), w, E * 4);
                else
                    for (std::uint64_t e = 0; e < E; ++e)
                        wt[e] = static_cast<float>(w[e]);
            }
            return CSRGraph(V, std::move(o), std::move(c), std::move(wt));
        }
    } // namespace

    CSRGraph GraphLoader::load_edge_list(const std::string &path, CSRGraph::offset_type num_vertices, bool weighted)
    {
        HostGraphHandle g;
        if (const int st = hyt_load_edge_list(path.c_str(), num_vertices, weighted ? 1U : 0U, HYT_HOST_PAGEABLE, &g.h);
            st != HYT_OK)
            raise(st, "hyt_load_edge_list");
        return to_csr(g.h, true);
    }

    void GraphLoader::save_binary(const CSRGraph &graph, const std::string &path)
    {
        const auto &w = graph.edge_weights();
        std::vector<std::uint32_t> bits(w.size());
        if (!w.empty())
            std::memcpy(bits.data(), w.data(), w.size() * 4);
        HostGraphHandle g;
        if (const int st = hyt_host_graph_from_csr(graph.num_vertices(), graph.num_edges(), graph.row_offsets().data(),
                                                   graph.column_indices().data(), w.empty() ? nullptr : bits.data(),
                                                   HYT_HOST_PAGEABLE, &g.h);
            st != HYT_OK)
            raise(st, "hyt_host_graph_from_csr");
        if (const int st = hyt_host_graph_save(g.h, path.c_str(), 1U); st != HYT_OK)
            raise(st, "hyt_host_graph_save");
    }

    CSRGraph GraphLoader::load_binary(const std::string &path)
    {
        HostGraphHandle g;
        std::uint32_t f32 = 0;
        if (const int st = hyt_host_graph_load(path.c_str(), HYT_HOST_PAGEABLE, 1U, &g.h, &f32); st != HYT_OK)
            raise(st, "hyt_host_graph_load");
        return to_csr(g.h, f32 != 0);
    }
} // namespace hytgraph::graph
