// GraphLoader - the interface of the original project's
// include/graph/graph_loader.hpp (GraphLoader::load_edge_list: same rules,
// messages and exception types), plus a binary CSR file format for large
// graphs (save_binary / load_binary), which the original does not have.
#pragma once

#include "graph/csr_graph.hpp"

#include <string>

namespace hytgraph::graph
{
    class GraphLoader
    {
    public:
        // Text edge list: each non-empty, non-'#' line is "source destination"
        // (or "source destination weight" when weighted), zero-based ids in
        // [0, num_vertices). Edges are grouped by source in input order;
        // parallel edges are kept. std::invalid_argument for malformed input,
        // std::runtime_error when the file cannot be read.
        [[nodiscard]] static CSRGraph load_edge_list(const std::string &path, CSRGraph::offset_type num_vertices,
                                                     bool weighted = false);

        // Binary CSR file (checksummed; hyt_host_graph_save / _load). Weights
        // are stored as float bit patterns. load_binary verifies checksums and
        // structure; std::invalid_argument for a corrupt file,
        // std::runtime_error for I/O errors.
        static void save_binary(const CSRGraph &graph, const std::string &path);
        [[nodiscard]] static CSRGraph load_binary(const std::string &path);
    };
} // namespace hytgraph::graph
