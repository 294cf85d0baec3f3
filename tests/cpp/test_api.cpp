// C++ API tests (include/gpu_bfs.hpp), written after the reference's Rust
// tests for the same calls:
//   crates/gpu_bfs/tests/adversarial_challenge.rs:195-244  grid Manhattan oracle
//   crates/gpu_bfs/tests/adversarial_bounds.rs:175-345     single vertex, chain, star,
//                                                          out-of-bounds source text
//   crates/gpu_bfs/tests/advanced_optimizations.rs:169-250 disconnected sentinel,
//                                                          duplicate sources
//   crates/core/src/graph.rs tests (from_edge_list, to_csc), mtx.rs tests
// Expected values come from closed forms (|dx|+|dy| on a grid, |i-j| on a
// path, 0/1/2 on a star, component membership) and from agreement between
// methods and output types, never from a host BFS.
//
//   test_api --cpu   host-side checks only (no GPU needed)
//   test_api         host-side and GPU checks
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gpu_bfs.hpp"

using namespace gpu_bfs;

static int failures = 0;
#define CHECK(cond)                                                                  \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                              \
        }                                                                            \
    } while (0)

template <typename F> static std::string thrown_message(F &&f, int *status = nullptr) {
    try {
        f();
    } catch (const Error &e) {
        if (status) *status = e.status();
        return e.what();
    }
    return "";
}

static CsrGraph path_graph(uint32_t n) {
    std::vector<std::pair<uint32_t, uint32_t>> e;
    for (uint32_t i = 0; i + 1 < n; ++i) {
        e.push_back({i, i + 1});
        e.push_back({i + 1, i});
    }
    return CsrGraph::from_edge_list(n, e);
}

// ------------------------------------------------------------------ host

static void host_checks() {
    // from_edge_list: self loops, out-of-range endpoints and duplicates dropped, sorted
    CsrGraph g = CsrGraph::from_edge_list(4, {{2, 1}, {0, 3}, {0, 1}, {1, 1}, {0, 3}, {5, 0}, {3, 9}});
    CHECK(g.num_vertices == 4 && g.num_edges == 3);
    CHECK((g.row_offsets == std::vector<uint32_t>{0, 2, 2, 3, 3}));
    CHECK((g.col_indices == std::vector<uint32_t>{1, 3, 1}));
    CHECK(g.degree(0) == 2 && g.max_degree() == 2 && !g.is_symmetric());
    g.validate();
    CscGraph c = g.to_csc();
    CHECK((c.col_offsets == std::vector<uint32_t>{0, 0, 2, 2, 3}));
    CHECK((c.row_indices == std::vector<uint32_t>{0, 2, 0}));
    CHECK(c.in_degree(1) == 2);

    CsrGraph grid = CsrGraph::grid2d(5, 4);
    CHECK(grid.num_vertices == 20 && grid.num_edges == 2 * (4 * 4 + 5 * 3));
    CHECK(grid.is_symmetric());

    // Matrix Market: symmetric file mirrored, diagonal dropped
    CsrGraph m = CsrGraph::from_mtx_str("%%MatrixMarket matrix coordinate pattern symmetric\n% c\n"
                                        "3 3 3\n2 1\n3 2\n3 3\n");
    CHECK(m.num_vertices == 3 && m.num_edges == 4 && m.is_symmetric());
    // MtxError::IndexOutOfBounds with its line and fields (error.rs:45-81)
    try {
        CsrGraph::from_mtx_str("%%MatrixMarket matrix coordinate pattern general\n3 3 1\n4 1\n");
        CHECK(false);
    } catch (const MtxError &e) {
        CHECK(e.detail.kind == MSBFS_MTX_INDEX_OUT_OF_BOUNDS && e.detail.line_number == 3);
        CHECK(e.detail.row == 4 && e.detail.col == 1 && e.detail.max_rows == 3);
        CHECK(std::string(e.what()) ==
              "Coordinate out of bounds at line 3: entry (4, 1) exceeds matrix dimensions (3, 3)");
    }
    // an invalid CSR is rejected by validate()
    CsrGraph bad = g;
    bad.col_indices[1] = 7;
    int st = 0;
    std::string msg = thrown_message([&] { bad.validate(); }, &st);
    CHECK(st == MSBFS_ERR_INVALID_GRAPH && !msg.empty());

    // bitmask layout: bit s%32 of words[v*wpv + s/32]
    ReachabilityBitmask b = ReachabilityBitmask::zeroed(3, 33);
    CHECK(b.words_per_vertex == 2 && b.words.size() == 6);
    b.words[1 * 2 + 1] = 1u;  // vertex 1, source 32
    CHECK(b.is_reachable(32, 1) && !b.is_reachable(0, 1) && b.total_reachable_pairs() == 1);
}

// ------------------------------------------------------------------ GPU

static void gpu_checks() {
    auto ctx = GpuContext::init_blocking(0);
    std::printf("device: %s (%s)\n", ctx->info().name, ctx->info().arch);
    std::vector<GpuBfsMethod> methods = {GpuBfsMethod::Auto, GpuBfsMethod::WorkgroupPersistentBfs,
                                         GpuBfsMethod::PushPullAdaptive, GpuBfsMethod::BitParallel64Pull,
                                         GpuBfsMethod::IndirectSparseQueuePush};

    // grid: dist((x0,y0),(x,y)) = |x-x0| + |y-y0|
    const uint32_t W = 37, H = 23;
    CsrGraph grid = CsrGraph::grid2d(W, H);
    std::vector<uint32_t> src;
    for (uint32_t i = 0; i < 70; ++i) src.push_back((i * 2654435761u) % (W * H));
    for (GpuBfsMethod meth : methods) {
        GpuBfsEngine eng(ctx, meth);
        std::vector<uint32_t> d = eng.run_ms_bfs(grid, src);
        CHECK(d.size() == src.size() * W * H);
        uint64_t wrong = 0;
        for (size_t s = 0; s < src.size(); ++s)
            for (uint32_t v = 0; v < W * H; ++v) {
                int dx = int(v % W) - int(src[s] % W), dy = int(v / W) - int(src[s] / W);
                wrong += d[s * W * H + v] != uint32_t(std::abs(dx) + std::abs(dy));
            }
        CHECK(wrong == 0);
        // reachability of a connected graph: every pair
        ReachabilityBitmask r = eng.run_ms_reachability(grid, src);
        CHECK(r.total_reachable_pairs() == uint64_t(src.size()) * W * H);
    }

    // path of 3000 vertices: |i-j|, beyond the u8 range, within u16
    CsrGraph path = path_graph(3000);
    std::vector<uint32_t> ps = {0, 2999, 1500, 7};
    {
        GpuBfsEngine eng(ctx, GpuBfsMethod::Auto);
        std::vector<uint32_t> d = eng.run_ms_bfs(path, ps);
        uint64_t wrong = 0;
        for (size_t s = 0; s < ps.size(); ++s)
            for (uint32_t v = 0; v < 3000; ++v)
                wrong += d[s * 3000 + v] != uint32_t(v > ps[s] ? v - ps[s] : ps[s] - v);
        CHECK(wrong == 0);
        DeviceGraph dp(ctx, path);
        std::vector<uint16_t> d16 = eng.run_ms_bfs_as<uint16_t>(dp, ps);
        wrong = 0;
        for (size_t i = 0; i < d.size(); ++i) wrong += d16[i] != d[i];
        CHECK(wrong == 0);
        int st = 0;
        thrown_message([&] { eng.run_ms_bfs_as<uint8_t>(dp, ps); }, &st);
        CHECK(st == MSBFS_ERR_DIAMETER_U8);
    }

    // star with 5000 leaves: 0 at the source, 1 hub<->leaf, 2 leaf<->leaf
    {
        std::vector<std::pair<uint32_t, uint32_t>> e;
        for (uint32_t i = 1; i <= 5000; ++i) {
            e.push_back({0, i});
            e.push_back({i, 0});
        }
        CsrGraph star = CsrGraph::from_edge_list(5001, e);
        std::vector<uint32_t> ss = {0, 1, 4999, 2500};
        for (GpuBfsMethod meth : methods) {
            GpuBfsEngine eng(ctx, meth);
            std::vector<uint32_t> d = eng.run_ms_bfs(star, ss);
            uint64_t wrong = 0;
            for (size_t s = 0; s < ss.size(); ++s)
                for (uint32_t v = 0; v <= 5000; ++v) {
                    uint32_t want = v == ss[s] ? 0 : (v == 0 || ss[s] == 0) ? 1 : 2;
                    wrong += d[s * 5001 + v] != want;
                }
            CHECK(wrong == 0);
        }
    }

    // two disconnected paths plus an isolated vertex: u32::MAX / 0xFF across components,
    // reachability bitmask equal to the finite entries
    {
        std::vector<std::pair<uint32_t, uint32_t>> e;
        for (uint32_t i = 0; i + 1 < 50; ++i) {
            e.push_back({i, i + 1});
            e.push_back({i + 1, i});
            e.push_back({50 + i, 51 + i});
            e.push_back({51 + i, 50 + i});
        }
        CsrGraph two = CsrGraph::from_edge_list(101, e);  // vertex 100 isolated
        std::vector<uint32_t> ts = {0, 60, 100, 49, 0};   // duplicate source 0
        DeviceGraph dt(ctx, two);
        for (GpuBfsMethod meth : methods) {
            GpuBfsEngine eng(ctx, meth);
            std::vector<uint32_t> d = eng.run_ms_bfs(dt, ts);
            std::vector<uint8_t> d8 = eng.run_ms_bfs_as<uint8_t>(dt, ts);
            ReachabilityBitmask r = eng.run_ms_reachability(dt, ts);
            uint64_t wrong = 0;
            for (size_t s = 0; s < ts.size(); ++s)
                for (uint32_t v = 0; v < 101; ++v) {
                    uint32_t a = ts[s], comp_a = a < 50 ? 0 : a < 100 ? 1 : 2;
                    uint32_t comp_v = v < 50 ? 0 : v < 100 ? 1 : 2;
                    uint32_t want = comp_a != comp_v ? 0xFFFFFFFFu : (v > a ? v - a : a - v);
                    uint32_t got = d[s * 101 + v];
                    wrong += got != want;
                    wrong += d8[s * 101 + v] != (want == 0xFFFFFFFFu ? 0xFFu : want);
                    wrong += r.is_reachable(s, v) != (want != 0xFFFFFFFFu);
                }
            CHECK(wrong == 0);
            // duplicate sources give identical rows
            CHECK(std::memcmp(&d[0], &d[4 * 101], 101 * sizeof(uint32_t)) == 0);
        }
    }

    // single vertex, S = 0, and the reference's out-of-bounds texts
    {
        GpuBfsEngine eng(ctx, GpuBfsMethod::Auto);
        CsrGraph one = CsrGraph::from_edge_list(1, {});
        std::vector<uint32_t> d = eng.run_ms_bfs(one, {0, 0});
        CHECK((d == std::vector<uint32_t>{0, 0}));
        std::vector<uint32_t> out = {1, 2, 3};
        eng.run_ms_bfs_into(grid, {}, out);
        CHECK(out.empty());
        int st = 0;
        std::string msg = thrown_message([&] { eng.run_ms_bfs(grid, {3, W * H}); }, &st);
        CHECK(st == MSBFS_ERR_SOURCE_OOB);
        CHECK(msg == "Source vertex " + std::to_string(W * H) + " at index 1 exceeds graph vertex count " +
                         std::to_string(W * H));
        CsrGraph empty = CsrGraph::from_edge_list(0, {});
        msg = thrown_message([&] { eng.run_ms_bfs(empty, {5}); }, &st);
        CHECK(msg == "Source vertex 5 at index 0 exceeds graph vertex count 0");
        CHECK(eng.run_ms_reachability(empty, {5}).words.empty());
        // with_method runs the named method, whatever the engine's own
        std::vector<uint32_t> a = eng.run_ms_bfs_with_method(GpuBfsMethod::PushPullAdaptive, grid, src);
        std::vector<uint32_t> b = eng.run_ms_bfs(grid, src);
        CHECK(a == b);
        CHECK(eng.resolve_method(grid, src) == GpuBfsEngine::select_auto_method(grid, src.size()));
    }
}

int main(int argc, char **argv) {
    bool cpu_only = argc > 1 && std::strcmp(argv[1], "--cpu") == 0;
    try {
        host_checks();
        std::printf("host checks done\n");
        if (!cpu_only) {
            gpu_checks();
            std::printf("gpu checks done\n");
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "unexpected exception: %s\n", e.what());
        return 2;
    }
    std::printf("%s: %d failed checks\n", failures ? "FAIL" : "OK", failures);
    return failures ? 1 : 0;
}
