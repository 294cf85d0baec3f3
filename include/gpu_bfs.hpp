/*
 * gpu_bfs.hpp -- C++17 host API over libmsbfs (msbfs.h), header only.
 *
 * The reference's host side is the Rust `gpu_bfs` crate; this header restates
 * its public surface with the same names, argument meanings and error
 * behaviour so that C++ callers (and tests written after the reference's
 * own) read like the Rust ones:
 *
 *   core::CsrGraph / CscGraph / ReachabilityBitmask  crates/core/src/graph.rs
 *   BfsEngine::run_ms_bfs                            crates/core/src/lib.rs:21-39
 *   GpuContext::init_blocking                        crates/gpu_bfs/src/context.rs:174-388
 *   GpuBfsMethod, GpuBfsEngine                       crates/gpu_bfs/src/engine.rs:33-480
 *   CsrGraph::from_mtx_reader, MtxError              crates/core/src/mtx.rs, error.rs:45-81
 *
 * Rust's `Result<T>` becomes a return value or a thrown gpu_bfs::Error whose
 * what() is the reference's message (bail! texts of engine.rs:233-249,
 * MtxError's Display).  Every traversal runs on the GPU through the C ABI;
 * nothing here computes distances on the host.
 */
#ifndef GPU_BFS_HPP
#define GPU_BFS_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "msbfs.h"

namespace gpu_bfs {

/* ---- errors ---- */

class Error : public std::runtime_error {
public:
    Error(int status, const std::string &msg) : std::runtime_error(msg), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

/* A source id >= |V| (engine.rs:233-249). */
class SourceOutOfBounds : public Error {
public:
    using Error::Error;
};

/* MtxError (crates/core/src/error.rs:45-81): kind and fields of the rejected line. */
class MtxError : public Error {
public:
    MtxError(const std::string &msg, const msbfs_mtx_error &d) : Error(MSBFS_ERR_PARSE, msg), detail(d) {}
    msbfs_mtx_error detail;
};

namespace detail {
inline void check(int rc) {
    if (rc == MSBFS_OK) return;
    const char *m = msbfs_last_error();
    std::string msg = (m && *m) ? m : msbfs_status_string(rc);
    if (rc == MSBFS_ERR_SOURCE_OOB) throw SourceOutOfBounds(rc, msg);
    if (rc == MSBFS_ERR_PARSE) {
        msbfs_mtx_error d{};
        msbfs_mtx_last_error(&d);
        throw MtxError(msg, d);
    }
    throw Error(rc, msg);
}
template <typename T> struct dist_code;
template <> struct dist_code<uint32_t> { static constexpr int value = MSBFS_DIST_U32; };
template <> struct dist_code<uint16_t> { static constexpr int value = MSBFS_DIST_U16; };
template <> struct dist_code<uint8_t> { static constexpr int value = MSBFS_DIST_U8; };
} // namespace detail

/* ---- graphs (crates/core/src/graph.rs) ---- */

struct CscGraph {
    uint64_t num_vertices = 0, num_edges = 0;
    std::vector<uint32_t> col_offsets, row_indices;
    uint64_t in_degree(uint64_t v) const { return col_offsets[v + 1] - col_offsets[v]; }
    std::pair<const uint32_t *, const uint32_t *> predecessors(uint64_t v) const {
        return {row_indices.data() + col_offsets[v], row_indices.data() + col_offsets[v + 1]};
    }
};

struct CsrGraph {
    uint64_t num_vertices = 0, num_edges = 0;
    std::vector<uint32_t> row_offsets{0}, col_indices;

    /* graph.rs:51-80: drops self loops and out-of-range endpoints, sorts, dedups. */
    static CsrGraph from_edge_list(uint64_t num_vertices,
                                   const std::vector<std::pair<uint32_t, uint32_t>> &edges) {
        std::vector<uint32_t> us(edges.size()), vs(edges.size());
        for (size_t i = 0; i < edges.size(); ++i) { us[i] = edges[i].first; vs[i] = edges[i].second; }
        msbfs_host_graph *h = nullptr;
        detail::check(msbfs_csr_from_edge_list(num_vertices, edges.size(), us.data(), vs.data(), &h));
        return take(h);
    }
    /* CsrGraph::from_mtx_reader (graph.rs:83-88) over a file path. */
    static CsrGraph from_mtx_file(const std::string &path) {
        msbfs_host_graph *h = nullptr;
        detail::check(msbfs_mtx_read(path.c_str(), &h));
        return take(h);
    }
    /* MtxReader::parse over in-memory text. */
    static CsrGraph from_mtx_str(const std::string &text) {
        msbfs_host_graph *h = nullptr;
        detail::check(msbfs_mtx_parse(text.data(), text.size(), &h));
        return take(h);
    }
    /* generators.rs:28-80 (row-major grid, 4-neighbour). */
    static CsrGraph grid2d(uint64_t width, uint64_t height) {
        msbfs_host_graph *h = nullptr;
        detail::check(msbfs_gen_grid2d(width, height, &h));
        return take(h);
    }
    static CsrGraph rmat(uint32_t scale, uint32_t edge_factor, uint64_t seed, bool directed = false,
                         double a = 0.57, double b = 0.19, double c = 0.19, double d = 0.05) {
        msbfs_host_graph *h = nullptr;
        detail::check(msbfs_gen_rmat(scale, edge_factor, a, b, c, d, seed, directed ? 1 : 0, &h));
        return take(h);
    }

    CscGraph to_csc() const {
        CscGraph c;
        c.num_vertices = num_vertices;
        c.num_edges = num_edges;
        c.col_offsets.assign(num_vertices + 1, 0);
        c.row_indices.assign(num_edges, 0);
        detail::check(msbfs_csr_to_csc(num_vertices, num_edges, row_offsets.data(), col_indices.data(),
                                       c.col_offsets.data(), c.row_indices.data()));
        return c;
    }
    bool is_symmetric() const {
        int r = msbfs_csr_is_symmetric(num_vertices, num_edges, row_offsets.data(), col_indices.data());
        if (r < 0) detail::check(-r);
        return r == 1;
    }
    void validate() const {
        if (row_offsets.size() != num_vertices + 1 || col_indices.size() != num_edges)
            throw Error(MSBFS_ERR_INVALID_GRAPH,
                        "Invalid row_offsets length: expected " + std::to_string(num_vertices + 1) +
                            " (|V| + 1), found " + std::to_string(row_offsets.size()));
        detail::check(msbfs_csr_validate(num_vertices, num_edges, row_offsets.data(), col_indices.data()));
    }
    bool is_empty() const { return num_vertices == 0; }
    uint64_t degree(uint64_t u) const { return row_offsets[u + 1] - row_offsets[u]; }
    std::pair<const uint32_t *, const uint32_t *> neighbors(uint64_t u) const {
        return {col_indices.data() + row_offsets[u], col_indices.data() + row_offsets[u + 1]};
    }
    uint64_t max_degree() const {
        uint64_t m = 0;
        for (uint64_t u = 0; u < num_vertices; ++u) m = degree(u) > m ? degree(u) : m;
        return m;
    }
    double avg_degree() const { return num_vertices ? double(num_edges) / double(num_vertices) : 0.0; }

private:
    static CsrGraph take(msbfs_host_graph *h) {
        CsrGraph g;
        const uint32_t *ro = nullptr, *ci = nullptr;
        int rc = msbfs_host_graph_get(h, &g.num_vertices, &g.num_edges, &ro, &ci);
        if (rc == MSBFS_OK) {
            g.row_offsets.assign(ro, ro + g.num_vertices + 1);
            g.col_indices.assign(ci, ci + g.num_edges);
        }
        msbfs_host_graph_free(h);
        detail::check(rc);
        return g;
    }
};

/* ReachabilityBitmask (graph.rs:446-646): vertex-major words, bit s%32 of
 * words[v*words_per_vertex + s/32]. */
struct ReachabilityBitmask {
    uint64_t num_vertices = 0, num_sources = 0, words_per_vertex = 0;
    std::vector<uint32_t> words;

    static ReachabilityBitmask zeroed(uint64_t num_vertices, uint64_t num_sources) {
        ReachabilityBitmask b;
        b.num_vertices = num_vertices;
        b.num_sources = num_sources;
        b.words_per_vertex = num_sources ? (num_sources + 31) / 32 : 0;
        b.words.assign(num_vertices * b.words_per_vertex, 0u);
        return b;
    }
    bool is_reachable(uint64_t s_idx, uint64_t v) const {
        return (words[v * words_per_vertex + s_idx / 32] >> (s_idx % 32)) & 1u;
    }
    uint64_t reachable_source_count(uint64_t v) const {
        uint64_t n = 0;
        for (uint64_t w = 0; w < words_per_vertex; ++w) n += __builtin_popcount(words[v * words_per_vertex + w]);
        return n;
    }
    uint64_t total_reachable_pairs() const {
        uint64_t n = 0;
        for (uint32_t w : words) n += __builtin_popcount(w);
        return n;
    }
    bool operator==(const ReachabilityBitmask &o) const {
        return num_vertices == o.num_vertices && num_sources == o.num_sources && words == o.words;
    }
};

/* ---- device ---- */

enum class GpuBfsMethod : int {
    NaiveSequential = MSBFS_METHOD_NAIVE_SEQUENTIAL,
    BatchedDenseSpMM = MSBFS_METHOD_BATCHED_DENSE_SPMM,
    BitParallel32 = MSBFS_METHOD_BIT_PARALLEL_32,
    BitParallel64 = MSBFS_METHOD_BIT_PARALLEL_64,
    BitParallel64Pull = MSBFS_METHOD_BIT_PARALLEL_64_PULL,
    BooleanReachability = MSBFS_METHOD_BOOLEAN_REACHABILITY,
    PushPullAdaptive = MSBFS_METHOD_PUSH_PULL_ADAPTIVE,
    SparseQueue = MSBFS_METHOD_SPARSE_QUEUE,
    SparseBitParallel64 = MSBFS_METHOD_SPARSE_BIT_PARALLEL_64,
    MultiWordBitParallel = MSBFS_METHOD_MULTI_WORD_BIT_PARALLEL,
    IndirectMacroBatch = MSBFS_METHOD_INDIRECT_MACRO_BATCH,
    IndirectSparseQueuePush = MSBFS_METHOD_INDIRECT_SPARSE_QUEUE_PUSH,
    MultiWordBitParallelPush = MSBFS_METHOD_MULTI_WORD_BIT_PARALLEL_PUSH,
    WorkgroupPersistentBfs = MSBFS_METHOD_WORKGROUP_PERSISTENT_BFS,
    Auto = MSBFS_METHOD_AUTO,
};

/* GpuContext (context.rs): one HIP device. */
class GpuContext {
public:
    static std::shared_ptr<GpuContext> init_blocking(int device = 0) {
        return std::shared_ptr<GpuContext>(new GpuContext(device));
    }
    ~GpuContext() { msbfs_context_destroy(h_); }
    GpuContext(const GpuContext &) = delete;
    GpuContext &operator=(const GpuContext &) = delete;

    msbfs_context *handle() const { return h_; }
    msbfs_device_info info() const {
        msbfs_device_info d{};
        detail::check(msbfs_context_info(h_, &d));
        return d;
    }
    void synchronize() { detail::check(msbfs_context_synchronize(h_)); }

private:
    explicit GpuContext(int device) { detail::check(msbfs_context_create(device, &h_)); }
    msbfs_context *h_ = nullptr;
};

/* A CSR resident in device memory (and its transpose).  Passing one to the
 * engine skips the per-call upload that a host CsrGraph costs. */
class DeviceGraph {
public:
    DeviceGraph(std::shared_ptr<GpuContext> ctx, const CsrGraph &g)
        : ctx_(std::move(ctx)), num_vertices(g.num_vertices), num_edges(g.num_edges) {
        detail::check(msbfs_graph_create(ctx_->handle(), g.num_vertices, g.num_edges, g.row_offsets.data(),
                                         g.col_indices.data(), &h_));
    }
    ~DeviceGraph() { msbfs_graph_destroy(h_); }
    DeviceGraph(const DeviceGraph &) = delete;
    DeviceGraph &operator=(const DeviceGraph &) = delete;

    msbfs_graph *handle() const { return h_; }
    msbfs_graph_info info() const {
        msbfs_graph_info i{};
        detail::check(msbfs_graph_get_info(h_, &i));
        return i;
    }

private:
    std::shared_ptr<GpuContext> ctx_;
    msbfs_graph *h_ = nullptr;

public:
    const uint64_t num_vertices, num_edges;
};

/* GpuBfsEngine (engine.rs:57-480). */
class GpuBfsEngine {
public:
    GpuBfsEngine(std::shared_ptr<GpuContext> context, GpuBfsMethod method)
        : ctx_(std::move(context)), method_(method) {
        detail::check(msbfs_engine_create(ctx_->handle(), int(method), &h_));
    }
    ~GpuBfsEngine() { msbfs_engine_destroy(h_); }
    GpuBfsEngine(const GpuBfsEngine &) = delete;
    GpuBfsEngine &operator=(const GpuBfsEngine &) = delete;

    /* engine.rs:68-155 */
    static std::unique_ptr<GpuBfsEngine> create(std::shared_ptr<GpuContext> c, GpuBfsMethod m) {
        return std::unique_ptr<GpuBfsEngine>(new GpuBfsEngine(std::move(c), m));
    }
    static std::unique_ptr<GpuBfsEngine> default_with_method(GpuBfsMethod m) {
        return create(GpuContext::init_blocking(0), m);
    }
    static std::unique_ptr<GpuBfsEngine> auto_(std::shared_ptr<GpuContext> c) {
        return create(std::move(c), GpuBfsMethod::Auto);
    }
    static std::unique_ptr<GpuBfsEngine> workgroup_persistent_bfs(std::shared_ptr<GpuContext> c) {
        return create(std::move(c), GpuBfsMethod::WorkgroupPersistentBfs);
    }
    static std::unique_ptr<GpuBfsEngine> push_pull_adaptive(std::shared_ptr<GpuContext> c) {
        return create(std::move(c), GpuBfsMethod::PushPullAdaptive);
    }
    static std::unique_ptr<GpuBfsEngine> multi_word_bit_parallel(std::shared_ptr<GpuContext> c) {
        return create(std::move(c), GpuBfsMethod::MultiWordBitParallel);
    }
    static std::unique_ptr<GpuBfsEngine> indirect_sparse_queue_push(std::shared_ptr<GpuContext> c) {
        return create(std::move(c), GpuBfsMethod::IndirectSparseQueuePush);
    }

    /* engine.rs:157-198: the reference's decision table. */
    static GpuBfsMethod select_auto_method(const CsrGraph &g, uint64_t num_sources) {
        return GpuBfsMethod(msbfs_select_auto_method(g.num_vertices, g.num_edges,
                                                     uint32_t(g.max_degree()), num_sources));
    }
    /* engine.rs:201-206 */
    GpuBfsMethod resolve_method(const CsrGraph &g, const std::vector<uint32_t> &sources) const {
        return method_ == GpuBfsMethod::Auto ? select_auto_method(g, sources.size()) : method_;
    }
    GpuBfsMethod method() const { return method_; }
    const std::shared_ptr<GpuContext> &context() const { return ctx_; }
    msbfs_engine *handle() const { return h_; }

    /* PushPullConfig { alpha, beta } (dispatch/push_pull.rs:22-36). */
    void set_push_pull_config(uint32_t alpha, uint32_t beta) {
        detail::check(msbfs_engine_set_push_pull(h_, alpha, beta));
    }
    void set_tuning(const msbfs_tuning &t) { detail::check(msbfs_engine_set_tuning(h_, &t)); }
    msbfs_tuning tuning() const {
        msbfs_tuning t{};
        detail::check(msbfs_engine_get_tuning(h_, &t));
        return t;
    }
    msbfs_run_stats last_stats() const {
        msbfs_run_stats s{};
        detail::check(msbfs_engine_last_stats(h_, &s));
        return s;
    }

    /* BfsEngine::run_ms_bfs (lib.rs:21-39): dist[s*|V| + v], u32::MAX unreached. */
    std::vector<uint32_t> run_ms_bfs(const CsrGraph &g, const std::vector<uint32_t> &sources) {
        std::vector<uint32_t> out;
        run_ms_bfs_into(g, sources, out);
        return out;
    }
    std::vector<uint32_t> run_ms_bfs(const DeviceGraph &g, const std::vector<uint32_t> &sources) {
        return run_as<uint32_t>(g, sources);
    }
    /* Compact outputs: 0xFFFF / 0xFF unreached; a distance the type cannot hold throws
     * (MSBFS_ERR_DIAMETER_U16 / _U8). */
    template <typename T>
    std::vector<T> run_ms_bfs_as(const DeviceGraph &g, const std::vector<uint32_t> &sources) {
        return run_as<T>(g, sources);
    }

    /* engine.rs:219-297: S == 0 clears `out`; otherwise resizes it to S*|V|. */
    void run_ms_bfs_into(const CsrGraph &g, const std::vector<uint32_t> &sources, std::vector<uint32_t> &out) {
        if (sources.empty()) {
            out.clear();
            return;
        }
        if (g.num_vertices == 0)
            throw SourceOutOfBounds(MSBFS_ERR_SOURCE_OOB, "Source vertex " + std::to_string(sources[0]) +
                                                              " at index 0 exceeds graph vertex count 0");
        DeviceGraph dg(ctx_, g);
        out.resize(sources.size() * g.num_vertices);
        detail::check(msbfs_run_ms_bfs(h_, dg.handle(), sources.data(), sources.size(), MSBFS_DIST_U32,
                                       out.data(), out.size()));
    }

    /* engine.rs:300-404: runs `method` whatever the engine's own is. */
    std::vector<uint32_t> run_ms_bfs_with_method(GpuBfsMethod method, const CsrGraph &g,
                                                 const std::vector<uint32_t> &sources) {
        GpuBfsEngine other(ctx_, method);
        other.set_tuning(tuning());
        return other.run_ms_bfs(g, sources);
    }

    /* engine.rs:420-479 */
    ReachabilityBitmask run_ms_reachability(const CsrGraph &g, const std::vector<uint32_t> &sources) {
        if (sources.empty() || g.num_vertices == 0) return ReachabilityBitmask::zeroed(g.num_vertices, sources.size());
        DeviceGraph dg(ctx_, g);
        return run_ms_reachability(dg, sources);
    }
    ReachabilityBitmask run_ms_reachability(const DeviceGraph &g, const std::vector<uint32_t> &sources) {
        ReachabilityBitmask b = ReachabilityBitmask::zeroed(g.num_vertices, sources.size());
        if (sources.empty() || g.num_vertices == 0) return b;
        detail::check(msbfs_run_ms_reachability(h_, g.handle(), sources.data(), sources.size(), b.words.data(),
                                                b.words.size()));
        return b;
    }

private:
    template <typename T> std::vector<T> run_as(const DeviceGraph &g, const std::vector<uint32_t> &sources) {
        std::vector<T> out;
        if (sources.empty()) return out;
        if (g.num_vertices == 0)
            throw SourceOutOfBounds(MSBFS_ERR_SOURCE_OOB, "Source vertex " + std::to_string(sources[0]) +
                                                              " at index 0 exceeds graph vertex count 0");
        out.resize(sources.size() * g.num_vertices);
        detail::check(msbfs_run_ms_bfs(h_, g.handle(), sources.data(), sources.size(), detail::dist_code<T>::value,
                                       out.data(), out.size()));
        return out;
    }

    std::shared_ptr<GpuContext> ctx_;
    GpuBfsMethod method_;
    msbfs_engine *h_ = nullptr;
};

} // namespace gpu_bfs

#endif
