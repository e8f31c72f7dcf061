// Graph files and user graphs as library host graphs:
//   * the original project's text edge list (src/graph/graph_loader.cpp:39-179):
//     "src dst" or "src dst weight" per line, '#' comments and blank lines
//     skipped, ids in [0, num_vertices), trailing tokens rejected, edges grouped
//     by source in input order (stable), parallel edges kept;
//   * a binary CSR file with a checksum per array, written and read with
//     positional I/O by several threads straight into the host graph's arrays
//     (pinned and device-mapped when the hybrid engine will read them), so a
//     scale-27..30 graph loads at disk speed instead of being regenerated;
//   * a copy of caller-owned CSR arrays into a host graph.
//
// Binary layout (little endian): a 128-byte header, then row offsets
// u64[V+1], columns u32[E] and, when weighted, weights u32[E], back to back.
#include "common.hpp"
#include "host_graph.hpp"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace hyt
{
    namespace
    {
        constexpr char kMagic[8] = {'H', 'Y', 'T', 'C', 'S', 'R', '0', '1'};
        constexpr std::uint32_t kVersion = 1;
        constexpr std::uint32_t kFlagWeighted = 1, kFlagWeightsF32 = 2;

        struct FileHeader
        {
            char magic[8];
            std::uint32_t version, flags;
            std::uint64_t V, E;
            std::uint64_t off_sum, col_sum, w_sum; // array checksums
            std::uint64_t header_sum;              // checksum of the bytes above
            std::uint8_t pad[64];
        };
        static_assert(sizeof(FileHeader) == 128, "header is 128 bytes");

        int io_threads(std::uint64_t bytes)
        {
            const unsigned hw = std::max(1U, std::thread::hardware_concurrency());
            const std::uint64_t by_size = bytes / (64ULL << 20) + 1; // >= 64 MiB per thread
            return static_cast<int>(std::min<std::uint64_t>({hw, 16ULL, by_size}));
        }

        // Runs fn(lo, hi) over [0, n) split into `parts` contiguous ranges on threads.
        template <class Fn>
        void parallel_ranges(std::uint64_t n, int parts, Fn fn)
        {
            if (parts <= 1 || n == 0)
            {
                fn(std::uint64_t{0}, n);
                return;
            }
            std::vector<std::thread> pool;
            std::vector<std::exception_ptr> errors(static_cast<std::size_t>(parts));
            for (int t = 0; t < parts; ++t)
                pool.emplace_back([&, t] {
                    try
                    {
                        fn(n * t / parts, n * (t + 1) / parts);
                    }
                    catch (...)
                    {
                        errors[static_cast<std::size_t>(t)] = std::current_exception();
                    }
                });
            for (auto &th : pool)
                th.join();
            for (auto &e : errors)
                if (e)
                    std::rethrow_exception(e);
        }

        // Checksum of a byte array: 4 MiB blocks hashed independently (four
        // interleaved SplitMix64 chains over the block's 64-bit words, the
        // tail zero-padded), then the block hashes folded in order.
        constexpr std::uint64_t kBlock = 4ULL << 20;

        std::uint64_t block_hash(const unsigned char *p, std::uint64_t n, std::uint64_t index)
        {
            std::uint64_t acc[4] = {index, index ^ 0x1ULL, index ^ 0x2ULL, index ^ 0x3ULL};
            const std::uint64_t words = n / 8;
            std::uint64_t i = 0;
            for (; i + 4 <= words; i += 4)
                for (int l = 0; l < 4; ++l)
                {
                    std::uint64_t w;
                    std::memcpy(&w, p + (i + l) * 8, 8);
                    acc[l] = mix64(acc[l] ^ w);
                }
            for (; i < words; ++i)
            {
                std::uint64_t w;
                std::memcpy(&w, p + i * 8, 8);
                acc[i % 4] = mix64(acc[i % 4] ^ w);
            }
            if (n % 8)
            {
                std::uint64_t w = 0;
                std::memcpy(&w, p + words * 8, n % 8);
                acc[0] = mix64(acc[0] ^ w);
            }
            return mix64(acc[0] ^ mix64(acc[1] ^ mix64(acc[2] ^ mix64(acc[3] ^ n))));
        }

        std::uint64_t checksum(const void *data, std::uint64_t bytes)
        {
            const unsigned char *p = static_cast<const unsigned char *>(data);
            const std::uint64_t nb = (bytes + kBlock - 1) / kBlock;
            std::vector<std::uint64_t> h(nb);
            parallel_ranges(nb, io_threads(bytes), [&](std::uint64_t lo, std::uint64_t hi) {
                for (std::uint64_t b = lo; b < hi; ++b)
                    h[b] = block_hash(p + b * kBlock, std::min(kBlock, bytes - b * kBlock), b);
            });
            std::uint64_t total = mix64(bytes);
            for (std::uint64_t b = 0; b < nb; ++b)
                total = mix64(total ^ h[b]);
            return total;
        }

        std::uint64_t header_checksum(const FileHeader &h)
        {
            return checksum(&h, offsetof(FileHeader, header_sum));
        }

        struct Fd
        {
            int fd = -1;
            ~Fd()
            {
                if (fd >= 0)
                    ::close(fd);
            }
        };

        [[noreturn]] void io_fail(const std::string &what, const std::string &path)
        {
            fail(HYT_ERR_IO, what + ": " + path + ": " + std::strerror(errno));
        }

        // Positional transfer of [0, bytes) at file offset `at`, several threads.
        template <bool Write>
        void transfer(int fd, unsigned char *buf, std::uint64_t bytes, std::uint64_t at, const std::string &path)
        {
            parallel_ranges(bytes, io_threads(bytes), [&](std::uint64_t lo, std::uint64_t hi) {
                while (lo < hi)
                {
                    const std::size_t n = static_cast<std::size_t>(std::min<std::uint64_t>(hi - lo, 1ULL << 30));
                    const ssize_t r = Write ? ::pwrite(fd, buf + lo, n, static_cast<off_t>(at + lo))
                                            : ::pread(fd, buf + lo, n, static_cast<off_t>(at + lo));
                    if (r < 0 && errno == EINTR)
                        continue;
                    if (r < 0)
                        io_fail(Write ? "write failed" : "read failed", path);
                    if (r == 0)
                        fail(HYT_ERR_IO, "unexpected end of file: " + path);
                    lo += static_cast<std::uint64_t>(r);
                }
            });
        }

        // Structural checks of a CSR (the original's CSRGraph::validate rules,
        // src/graph/csr_graph.cpp:105-169), parallel over vertices.
        void validate_csr(const HostGraph &g)
        {
            require(g.off[0] == 0, "row offsets must start at zero");
            require(g.off[g.V] == g.E, "last row offset must equal the edge count");
            const int T = io_threads((g.V + g.E) * 4);
            parallel_ranges(g.V, T, [&](std::uint64_t lo, std::uint64_t hi) {
                for (std::uint64_t v = lo; v < hi; ++v)
                    require(g.off[v] <= g.off[v + 1], "row offsets must be non-decreasing");
            });
            parallel_ranges(g.E, T, [&](std::uint64_t lo, std::uint64_t hi) {
                for (std::uint64_t e = lo; e < hi; ++e)
                    require(g.col[e] < g.V, "column index out of range");
            });
        }

        // ------------------------------------------------------- text edge list
        bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; }

        // Unsigned integer token as an istream extracts one: optional sign,
        // digits (at least one), stops at the first non-digit; a '-' negates
        // modulo 2^64; overflow fails.
        bool parse_u64(const char *&p, const char *end, std::uint64_t &out)
        {
            while (p < end && is_space(*p))
                ++p;
            bool neg = false;
            if (p < end && (*p == '+' || *p == '-'))
                neg = *p++ == '-';
            const char *d0 = p;
            std::uint64_t v = 0;
            while (p < end && *p >= '0' && *p <= '9')
            {
                const std::uint64_t digit = static_cast<std::uint64_t>(*p - '0');
                if (v > (~0ULL - digit) / 10)
                    return false;
                v = v * 10 + digit;
                ++p;
            }
            if (p == d0)
                return false;
            out = neg ? (0 - v) : v;
            return true;
        }

        // Decimal float token: [+-] digits [. digits] [(e|E) [+-] digits], at
        // least one mantissa digit; converted with strtof (correctly rounded).
        // Forms an istream does not read as a number (inf, nan, hex) fail.
        bool parse_f32(const char *&p, const char *end, float &out)
        {
            while (p < end && is_space(*p))
                ++p;
            const char *t0 = p;
            const char *q = p;
            if (q < end && (*q == '+' || *q == '-'))
                ++q;
            int digits = 0;
            while (q < end && *q >= '0' && *q <= '9')
                ++q, ++digits;
            if (q < end && *q == '.')
            {
                ++q;
                while (q < end && *q >= '0' && *q <= '9')
                    ++q, ++digits;
            }
            if (digits == 0)
                return false;
            if (q < end && (*q == 'e' || *q == 'E'))
            {
                const char *x = q + 1;
                if (x < end && (*x == '+' || *x == '-'))
                    ++x;
                const char *x0 = x;
                while (x < end && *x >= '0' && *x <= '9')
                    ++x;
                if (x == x0)
                    return false; // an exponent mark without digits
                q = x;
            }
            const std::string tok(t0, q);
            errno = 0;
            const float v = std::strtof(tok.c_str(), nullptr);
            if (errno == ERANGE && !std::isfinite(v))
                return false;
            out = v;
            p = q;
            return true;
        }

        struct TextEdge
        {
            std::uint32_t src, dst, w;
        };

        std::unique_ptr<HostGraph> load_edge_list(const std::string &path, std::uint64_t V, bool weighted, int memory)
        {
            const std::string who = "GraphLoader::load_edge_list: ";
            Fd f;
            f.fd = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
            if (f.fd < 0)
                fail(HYT_ERR_IO, who + "unable to open file: " + path);
            struct stat sb;
            if (::fstat(f.fd, &sb) != 0)
                fail(HYT_ERR_IO, who + "error while reading file: " + path);
            std::string text(static_cast<std::size_t>(sb.st_size), '\0');
            if (!text.empty())
                transfer<false>(f.fd, reinterpret_cast<unsigned char *>(text.data()), text.size(), 0, path);
            require(V <= 0xFFFFFFFFULL, who + "num_vertices must fit 32-bit vertex ids");

            std::vector<TextEdge> edges;
            std::uint64_t line_number = 0;
            const char *p = text.data(), *end = p + text.size();
            while (p < end)
            {
                const char *nl = static_cast<const char *>(std::memchr(p, '\n', static_cast<std::size_t>(end - p)));
                const char *le = nl ? nl : end;
                ++line_number;
                const char *a = p, *b = le;
                p = nl ? nl + 1 : end;
                while (a < b && (*a == ' ' || *a == '\t' || *a == '\r'))
                    ++a;
                while (b > a && (b[-1] == ' ' || b[-1] == '\t' || b[-1] == '\r'))
                    --b;
                if (a == b || *a == '#')
                    continue;
                const std::string at = std::to_string(line_number);
                std::uint64_t s = 0, d = 0;
                if (!parse_u64(a, b, s) || !parse_u64(a, b, d))
                    fail(HYT_ERR_INVALID_ARGUMENT, who + "malformed edge at line " + at);
                if (s >= V || d >= V)
                    fail(HYT_ERR_INVALID_ARGUMENT, who + "vertex ID out of range at line " + at);
                TextEdge e{static_cast<std::uint32_t>(s), static_cast<std::uint32_t>(d), 0U};
                if (weighted)
                {
                    float w = 0.0F;
                    if (!parse_f32(a, b, w))
                        fail(HYT_ERR_INVALID_ARGUMENT, who + "missing edge weight at line " + at);
                    std::memcpy(&e.w, &w, 4);
                }
                while (a < b && is_space(*a))
                    ++a;
                if (a < b)
                    fail(HYT_ERR_INVALID_ARGUMENT, who + "unexpected data at line " + at);
                edges.push_back(e);
            }

            // group by source, input order kept within a source (counting sort)
            auto g = std::make_unique<HostGraph>(V, edges.size(), weighted, memory);
            std::fill(g->off, g->off + V + 1, 0ULL);
            for (const TextEdge &e : edges)
                ++g->off[e.src + 1];
            for (std::uint64_t v = 0; v < V; ++v)
                g->off[v + 1] += g->off[v];
            std::vector<std::uint64_t> next(g->off, g->off + V);
            for (const TextEdge &e : edges)
            {
                const std::uint64_t at = next[e.src]++;
                g->col[at] = e.dst;
                if (weighted)
                    g->w[at] = e.w;
            }
            return g;
        }
    } // namespace
} // namespace hyt

extern "C" {

int hyt_host_graph_save(hyt_host_graph_t graph, const char *path, uint32_t weights_f32)
{
    return hyt::guarded("hyt_host_graph_save", [&] {
        using namespace hyt;
        require(path != nullptr, "null path");
        const HostGraph &g = *host_graph_of(graph);
        FileHeader h{};
        std::memcpy(h.magic, kMagic, 8);
        h.version = kVersion;
        h.flags = (g.weighted ? kFlagWeighted : 0U) | (g.weighted && weights_f32 ? kFlagWeightsF32 : 0U);
        h.V = g.V;
        h.E = g.E;
        const std::uint64_t ob = (g.V + 1) * 8, cb = g.E * 4, wb = g.weighted ? g.E * 4 : 0;
        h.off_sum = checksum(g.off, ob);
        h.col_sum = checksum(g.col, cb);
        h.w_sum = g.weighted ? checksum(g.w, wb) : 0;
        h.header_sum = header_checksum(h);
        Fd f;
        f.fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
        if (f.fd < 0)
            io_fail("cannot create", path);
        if (::ftruncate(f.fd, static_cast<off_t>(sizeof(FileHeader) + ob + cb + wb)) != 0)
            io_fail("cannot size", path);
        transfer<true>(f.fd, reinterpret_cast<unsigned char *>(&h), sizeof(h), 0, path);
        transfer<true>(f.fd, reinterpret_cast<unsigned char *>(g.off), ob, sizeof(FileHeader), path);
        transfer<true>(f.fd, reinterpret_cast<unsigned char *>(g.col), cb, sizeof(FileHeader) + ob, path);
        if (wb)
            transfer<true>(f.fd, reinterpret_cast<unsigned char *>(g.w), wb, sizeof(FileHeader) + ob + cb, path);
        if (::fsync(f.fd) != 0)
            io_fail("fsync failed", path);
    });
}

int hyt_host_graph_load(const char *path, int32_t memory, uint32_t verify, hyt_host_graph_t *out,
                        uint32_t *weights_f32)
{
    return hyt::guarded("hyt_host_graph_load", [&] {
        using namespace hyt;
        require(path != nullptr && out != nullptr, "null pointer");
        require(memory == HYT_HOST_PAGEABLE || memory == HYT_HOST_PINNED, "memory must be 0 or 1");
        *out = nullptr;
        Fd f;
        f.fd = ::open(path, O_RDONLY | O_CLOEXEC);
        if (f.fd < 0)
            io_fail("cannot open", path);
        struct stat sb;
        if (::fstat(f.fd, &sb) != 0)
            io_fail("cannot stat", path);
        const std::uint64_t size = static_cast<std::uint64_t>(sb.st_size);
        require(size >= sizeof(FileHeader), std::string("not a binary CSR file (too short): ") + path);
        FileHeader h{};
        transfer<false>(f.fd, reinterpret_cast<unsigned char *>(&h), sizeof(h), 0, path);
        require(std::memcmp(h.magic, kMagic, 8) == 0, std::string("not a binary CSR file (magic): ") + path);
        require(h.version == kVersion, "unsupported binary CSR version " + std::to_string(h.version));
        require(h.header_sum == header_checksum(h), std::string("binary CSR header checksum mismatch: ") + path);
        require(h.V >= 1 && h.V <= 0xFFFFFFFFULL, "vertex count must be in [1, 2^32-1]");
        const bool weighted = (h.flags & kFlagWeighted) != 0;
        const std::uint64_t ob = (h.V + 1) * 8, cb = h.E * 4, wb = weighted ? h.E * 4 : 0;
        require(size == sizeof(FileHeader) + ob + cb + wb, std::string("binary CSR file size mismatch: ") + path);
        auto g = std::make_unique<HostGraph>(h.V, h.E, weighted, memory);
        transfer<false>(f.fd, reinterpret_cast<unsigned char *>(g->off), ob, sizeof(FileHeader), path);
        transfer<false>(f.fd, reinterpret_cast<unsigned char *>(g->col), cb, sizeof(FileHeader) + ob, path);
        if (wb)
            transfer<false>(f.fd, reinterpret_cast<unsigned char *>(g->w), wb, sizeof(FileHeader) + ob + cb, path);
        if (verify)
        {
            require(checksum(g->off, ob) == h.off_sum, std::string("row offsets checksum mismatch: ") + path);
            require(checksum(g->col, cb) == h.col_sum, std::string("columns checksum mismatch: ") + path);
            if (wb)
                require(checksum(g->w, wb) == h.w_sum, std::string("weights checksum mismatch: ") + path);
            validate_csr(*g);
        }
        if (weights_f32)
            *weights_f32 = (h.flags & kFlagWeightsF32) ? 1U : 0U;
        *out = new hyt_host_graph{std::move(g)};
    });
}

int hyt_load_edge_list(const char *path, uint64_t num_vertices, uint32_t weighted, int32_t memory,
                       hyt_host_graph_t *out)
{
    return hyt::guarded("hyt_load_edge_list", [&] {
        using namespace hyt;
        require(path != nullptr && out != nullptr, "null pointer");
        require(memory == HYT_HOST_PAGEABLE || memory == HYT_HOST_PINNED, "memory must be 0 or 1");
        *out = nullptr;
        *out = new hyt_host_graph{load_edge_list(path, num_vertices, weighted != 0, memory)};
    });
}

int hyt_host_graph_from_csr(uint64_t num_vertices, uint64_t num_edges, const uint64_t *row_offsets,
                            const uint32_t *columns, const uint32_t *weights, int32_t memory,
                            hyt_host_graph_t *out)
{
    return hyt::guarded("hyt_host_graph_from_csr", [&] {
        using namespace hyt;
        require(row_offsets != nullptr && out != nullptr && (columns != nullptr || num_edges == 0), "null pointer");
        require(memory == HYT_HOST_PAGEABLE || memory == HYT_HOST_PINNED, "memory must be 0 or 1");
        require(num_vertices >= 1 && num_vertices <= 0xFFFFFFFFULL, "vertex count must be in [1, 2^32-1]");
        *out = nullptr;
        auto g = std::make_unique<HostGraph>(num_vertices, num_edges, weights != nullptr, memory);
        const int T = io_threads((num_vertices + num_edges) * 8);
        auto copy = [&](void *dst, const void *src, std::uint64_t bytes) {
            parallel_ranges(bytes, T, [&](std::uint64_t lo, std::uint64_t hi) {
                std::memcpy(static_cast<unsigned char *>(dst) + lo, static_cast<const unsigned char *>(src) + lo,
                            hi - lo);
            });
        };
        copy(g->off, row_offsets, (num_vertices + 1) * 8);
        copy(g->col, columns, num_edges * 4);
        if (weights)
            copy(g->w, weights, num_edges * 4);
        validate_csr(*g);
        *out = new hyt_host_graph{std::move(g)};
    });
}

} // extern "C"
