// The library's host-side code (webgpu-bfs_amd/csrc/host_graph.cpp and
// mtx.cpp: CSR builders, transpose, validation, lattice test, generators,
// source sampling, Matrix Market parser) and the oracle (oracle/oracle.c),
// compiled from source with AddressSanitizer and UndefinedBehaviorSanitizer
// by g++/gcc (`make -C tests/cpp test_sanitize`) and driven over the inputs
// their bounds depend on:
//   - empty, single-vertex, one-row and one-column graphs, partial last rows;
//   - Matrix Market text handed over in heap buffers of exactly `len` bytes
//     with no terminating NUL, every prefix of a valid file, byte mutations,
//     numbers past 2^64, and bodies large enough to be parsed in several
//     threaded chunks (crates/core/src/mtx.rs:124-287 is the behaviour);
//   - every oracle BFS form (fifo.rs, rayon_ms.rs, bit_parallel.rs) and the
//     u16/u8 output codecs at ragged lengths.
// Expected values come from closed forms (grid edge counts and Manhattan
// distances, complete graphs) and from agreement between the library's and
// the oracle's independent implementations of the same step.  A sanitizer
// report ends the program with a non-zero status (halt_on_error).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "msbfs.h"

extern "C" {
// oracle/oracle.c
size_t orc_csr_from_edge_list(size_t V, size_t n, const uint32_t *us, const uint32_t *vs,
                              uint32_t *row_offsets, uint32_t *cols);
void orc_csr_to_csc(size_t V, size_t E, const uint32_t *ro, const uint32_t *ci,
                    uint32_t *col_offsets, uint32_t *row_indices);
int orc_ms_bfs_fifo(size_t V, const uint32_t *ro, const uint32_t *ci, const uint32_t *sources,
                    size_t S, uint32_t *dist);
int orc_ms_bfs_parallel(size_t V, const uint32_t *ro, const uint32_t *ci,
                        const uint32_t *sources, size_t S, uint32_t *dist, int nthreads);
int orc_ms_bfs_bit_parallel(size_t V, const uint32_t *ro, const uint32_t *ci,
                            const uint32_t *sources, size_t S, uint32_t *dist);
uint64_t orc_traversed_edges(size_t V, const uint32_t *ro, size_t S, const uint32_t *dist);
void orc_reachability_from_dist(size_t V, size_t S, const uint32_t *dist, uint32_t *words);
void orc_grid_manhattan(size_t width, size_t height, const uint32_t *sources, size_t S,
                        uint32_t *dist);
void orc_pack_u16(size_t n, const uint32_t *d, uint32_t *out);
void orc_unpack_u16(size_t n, const uint32_t *words, uint32_t *d);
void orc_pack_u8(size_t n, const uint32_t *d, uint32_t *out);
void orc_unpack_u8(size_t n, const uint32_t *words, uint32_t *d);
}

// webgpu-bfs_amd/csrc/common.hpp (library-internal, linked from host_graph.cpp)
namespace msbfs {
uint32_t host_levels_from(uint64_t V, const uint32_t *ro, const uint32_t *ci, uint32_t s);
}

static const uint32_t INF = 0xFFFFFFFFu;
static int failures = 0;
static int checks = 0;
#define CHECK(cond)                                                                       \
    do {                                                                                  \
        ++checks;                                                                         \
        if (!(cond)) {                                                                    \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                                   \
        }                                                                                 \
    } while (0)

// A library-owned host graph copied into vectors sized exactly V+1 and E, so
// that any read past them is a heap overflow the sanitizer sees.
struct Csr {
    uint64_t V = 0, E = 0;
    std::vector<uint32_t> ro, ci;
};

static Csr take(msbfs_host_graph *h) {
    Csr c;
    c.ro.assign(1, 0);
    CHECK(h != nullptr);
    if (!h) return c;
    const uint32_t *ro = nullptr, *ci = nullptr;
    CHECK(msbfs_host_graph_get(h, &c.V, &c.E, &ro, &ci) == MSBFS_OK);
    c.ro.assign(ro, ro + c.V + 1);
    c.ci.assign(ci, ci + c.E);
    msbfs_host_graph_free(h);
    return c;
}

static bool same(const Csr &a, const Csr &b) {
    return a.V == b.V && a.E == b.E && a.ro == b.ro && a.ci == b.ci;
}

// The oracle's CSR of the same pairs (qsort + dedup), beside the library's.
static Csr oracle_csr(uint64_t V, const std::vector<uint32_t> &us, const std::vector<uint32_t> &vs) {
    Csr c;
    c.V = V;
    c.ro.assign(V + 1, 0);
    std::vector<uint32_t> cols(us.size() + 1);
    c.E = orc_csr_from_edge_list(V, us.size(), us.data(), vs.data(), c.ro.data(), cols.data());
    c.ci.assign(cols.begin(), cols.begin() + c.E);
    return c;
}

// Transpose, symmetry, validation and the three oracle BFS forms on one graph;
// returns the fifo distances of `S` sources drawn by msbfs_sample_sources.
static std::vector<uint32_t> graph_checks(const Csr &g, uint64_t S, uint64_t seed,
                                          std::vector<uint32_t> *src_out = nullptr) {
    CHECK(msbfs_csr_validate(g.V, g.E, g.ro.data(), g.ci.data()) == MSBFS_OK);
    std::vector<uint32_t> co(g.V + 1), ri(g.E), oco(g.V + 1), ori(g.E);
    CHECK(msbfs_csr_to_csc(g.V, g.E, g.ro.data(), g.ci.data(), co.data(), ri.data()) == MSBFS_OK);
    orc_csr_to_csc(g.V, g.E, g.ro.data(), g.ci.data(), oco.data(), ori.data());
    CHECK(co == oco && ri == ori);
    const int sym = msbfs_csr_is_symmetric(g.V, g.E, g.ro.data(), g.ci.data());
    CHECK(sym == (int)(co == g.ro && ri == g.ci));
    uint32_t d[4];
    const int k = msbfs_csr_lattice_offsets(g.V, g.E, g.ro.data(), g.ci.data(), d);
    CHECK(k >= 0 && k <= 4);

    S = std::min<uint64_t>(S, g.V);
    std::vector<uint32_t> src(S);
    CHECK(msbfs_sample_sources(g.ro.data(), g.V, S, seed, 0, src.data()) == MSBFS_OK);
    for (uint64_t i = 0; i < S; ++i) CHECK(src[i] < g.V && (i == 0 || src[i - 1] < src[i]));
    std::vector<uint32_t> fifo(S * g.V), par(S * g.V), bit(S * g.V);
    CHECK(orc_ms_bfs_fifo(g.V, g.ro.data(), g.ci.data(), src.data(), S, fifo.data()) == 0);
    CHECK(orc_ms_bfs_parallel(g.V, g.ro.data(), g.ci.data(), src.data(), S, par.data(), 3) == 0);
    CHECK(orc_ms_bfs_bit_parallel(g.V, g.ro.data(), g.ci.data(), src.data(), S, bit.data()) == 0);
    CHECK(fifo == par && fifo == bit);
    for (uint64_t s = 0; s < S; ++s) CHECK(fifo[s * g.V + src[s]] == 0);
    if (S) {
        // levels of one search: its largest finite distance + 1
        uint32_t far = 0;
        for (uint64_t v = 0; v < g.V; ++v)
            if (fifo[v] != INF) far = std::max(far, fifo[v]);
        CHECK(msbfs::host_levels_from(g.V, g.ro.data(), g.ci.data(), src[0]) == far + 1);
    }
    CHECK(msbfs::host_levels_from(g.V, g.ro.data(), g.ci.data(), (uint32_t)g.V) == 0);

    // traversed edges and reachability words against direct counts
    uint64_t te = 0;
    const uint64_t wpv = (S + 31) / 32;
    std::vector<uint32_t> words(g.V * wpv + 1);
    orc_reachability_from_dist(g.V, S, fifo.data(), words.data());
    for (uint64_t s = 0; s < S; ++s)
        for (uint64_t v = 0; v < g.V; ++v) {
            const bool r = fifo[s * g.V + v] != INF;
            if (r) te += g.ro[v + 1] - g.ro[v];
            CHECK(((words[v * wpv + s / 32] >> (s % 32)) & 1u) == (uint32_t)r);
        }
    CHECK(orc_traversed_edges(g.V, g.ro.data(), S, fifo.data()) == te);
    if (src_out) *src_out = src;
    return fifo;
}

static void grid_checks() {
    const uint64_t dims[][2] = {{1, 1}, {1, 9}, {9, 1}, {2, 2}, {5, 4}, {33, 17}, {64, 3}};
    for (auto &wh : dims) {
        const uint64_t w = wh[0], h = wh[1];
        msbfs_host_graph *hg = nullptr;
        CHECK(msbfs_gen_grid2d(w, h, &hg) == MSBFS_OK);
        Csr g = take(hg);
        CHECK(g.V == w * h);
        CHECK(g.E == 2 * ((w - 1) * h + w * (h - 1)));
        CHECK(msbfs_csr_is_symmetric(g.V, g.E, g.ro.data(), g.ci.data()) == 1);
        uint32_t d[4];
        const int k = msbfs_csr_lattice_offsets(g.V, g.E, g.ro.data(), g.ci.data(), d);
        if (w > 1 && h > 1) CHECK(k == 2 && d[0] == 1 && d[1] == w);
        if ((w == 1) != (h == 1)) CHECK(k == 1 && d[0] == 1);
        if (w == 1 && h == 1) CHECK(k == 0);
        // from a corner: |dx| + |dy| reaches w - 1 + h - 1
        CHECK(msbfs::host_levels_from(g.V, g.ro.data(), g.ci.data(), 0) == w + h - 1);
        std::vector<uint32_t> src;
        std::vector<uint32_t> fifo = graph_checks(g, 70, w * 131 + h, &src);
        std::vector<uint32_t> man(fifo.size());
        orc_grid_manhattan(w, h, src.data(), src.size(), man.data());
        CHECK(fifo == man);
    }
    msbfs_host_graph *hg = nullptr;
    CHECK(msbfs_gen_grid2d(0, 5, &hg) == MSBFS_ERR_INVALID_ARG && hg == nullptr);
    CHECK(msbfs_gen_grid2d(5, 0, &hg) == MSBFS_ERR_INVALID_ARG && hg == nullptr);
    CHECK(msbfs_gen_grid2d(1ull << 20, 1ull << 13, &hg) == MSBFS_ERR_INVALID_ARG);
}

static void generator_checks() {
    // road-like: partial last row, spanning tree => one component; 999
    // vertices in 40 columns have 974 + 959 grid edges, 998 of them in the
    // tree, so at most 935 extra edges, and with all of them the whole grid
    const uint64_t road[][3] = {{1, 1, 0}, {7, 3, 0}, {1000, 37, 50}, {999, 40, 935}, {500, 500, 0}};
    for (auto &r : road) {
        msbfs_host_graph *hg = nullptr;
        CHECK(msbfs_gen_road_like(r[0], r[1], r[2], 11, &hg) == MSBFS_OK);
        Csr g = take(hg);
        CHECK(g.V == r[0]);
        CHECK(g.E == 2 * (r[0] - 1 + r[2]));
        if (r[2] == 935) CHECK(g.E == 2 * (974 + 959));
        CHECK(msbfs_csr_is_symmetric(g.V, g.E, g.ro.data(), g.ci.data()) == 1);
        uint32_t d[4];
        const int k = msbfs_csr_lattice_offsets(g.V, g.E, g.ro.data(), g.ci.data(), d);
        CHECK(k >= 0 && k <= 2);
        for (int i = 0; i < k; ++i) CHECK(d[i] == 1 || d[i] == r[1]);
        std::vector<uint32_t> fifo = graph_checks(g, 5, r[0]);
        for (uint32_t x : fifo) CHECK(x != INF);
    }
    // R-MAT, directed and undirected
    for (int directed = 0; directed < 2; ++directed)
        for (uint32_t scale : {1u, 4u, 10u}) {
            msbfs_host_graph *hg = nullptr;
            CHECK(msbfs_gen_rmat(scale, 8, 0.57, 0.19, 0.19, 0.05, 5 + scale, directed, &hg) ==
                  MSBFS_OK);
            Csr g = take(hg);
            CHECK(g.V == (1ull << scale));
            if (!directed) CHECK(msbfs_csr_is_symmetric(g.V, g.E, g.ro.data(), g.ci.data()) == 1);
            graph_checks(g, 130, scale);
        }
    // G(n,p): p = 0 has no edges, p = 1 is complete, both sampling branches
    for (double p : {0.0, 1.0, 0.02, 0.6}) {
        msbfs_host_graph *hg = nullptr;
        CHECK(msbfs_gen_erdos_renyi_np(40, p, 3, &hg) == MSBFS_OK);
        Csr g = take(hg);
        if (p == 0.0) CHECK(g.E == 0);
        if (p == 1.0) CHECK(g.E == 40 * 39);
        std::vector<uint32_t> fifo = graph_checks(g, 40, 17);
        if (p == 1.0)
            for (uint64_t i = 0; i < fifo.size(); ++i) CHECK(fifo[i] == (i / 40 == i % 40 ? 0u : 1u));
    }
    msbfs_host_graph *hg = nullptr;
    CHECK(msbfs_gen_erdos_renyi_np(1, 0.5, 3, &hg) == MSBFS_OK);
    Csr one = take(hg);
    CHECK(one.V == 1 && one.E == 0);
    graph_checks(one, 1, 1);
    CHECK(msbfs_gen_erdos_renyi_np(0, 0.5, 3, &hg) == MSBFS_ERR_INVALID_ARG);
    CHECK(msbfs_gen_erdos_renyi_np(10, 1.5, 3, &hg) == MSBFS_ERR_INVALID_ARG);
    CHECK(msbfs_gen_road_like(0, 3, 0, 1, &hg) == MSBFS_ERR_INVALID_ARG);
    CHECK(msbfs_gen_road_like(999, 40, 936, 11, &hg) == MSBFS_ERR_INVALID_ARG);
}

static void builder_checks() {
    std::mt19937_64 rng(7);
    for (uint64_t V : {1ull, 2ull, 17ull, 300ull}) {
        for (int rep = 0; rep < 4; ++rep) {
            // pairs with self loops, duplicates and endpoints up to 2V (out of range)
            const uint64_t n = rng() % (4 * V + 3);
            std::vector<uint32_t> us(n), vs(n);
            for (uint64_t i = 0; i < n; ++i) {
                us[i] = (uint32_t)(rng() % (2 * V));
                vs[i] = (i % 5 == 0) ? us[i] : (uint32_t)(rng() % (2 * V));
            }
            msbfs_host_graph *hg = nullptr;
            CHECK(msbfs_csr_from_edge_list(V, n, us.data(), vs.data(), &hg) == MSBFS_OK);
            Csr g = take(hg);
            CHECK(same(g, oracle_csr(V, us, vs)));
            graph_checks(g, 3, V + rep);

            // undirected: the oracle CSR of the mirrored pairs
            CHECK(msbfs_csr_from_undirected_pairs(V, n, us.data(), vs.data(), &hg) == MSBFS_OK);
            Csr u = take(hg);
            std::vector<uint32_t> mu(us), mv(vs);
            mu.insert(mu.end(), vs.begin(), vs.end());
            mv.insert(mv.end(), us.begin(), us.end());
            CHECK(same(u, oracle_csr(V, mu, mv)));
            CHECK(msbfs_csr_is_symmetric(u.V, u.E, u.ro.data(), u.ci.data()) == 1);
        }
    }
    // empty graph, empty pair list
    msbfs_host_graph *hg = nullptr;
    CHECK(msbfs_csr_from_edge_list(0, 0, nullptr, nullptr, &hg) == MSBFS_OK);
    Csr z = take(hg);
    CHECK(z.V == 0 && z.E == 0 && z.ro.size() == 1);
    CHECK(msbfs_csr_validate(0, 0, z.ro.data(), nullptr) == MSBFS_OK);

    // validation rejects each malformation
    const uint32_t ro_ok[] = {0, 1, 2}, ci_ok[] = {1, 0};
    CHECK(msbfs_csr_validate(2, 2, ro_ok, ci_ok) == MSBFS_OK);
    const uint32_t ro_dec[] = {0, 2, 1}, ro_end[] = {0, 1, 1}, ro_start[] = {1, 1, 2};
    const uint32_t ci_oob[] = {1, 2};
    CHECK(msbfs_csr_validate(2, 2, ro_dec, ci_ok) == MSBFS_ERR_INVALID_GRAPH);
    CHECK(msbfs_csr_validate(2, 2, ro_end, ci_ok) == MSBFS_ERR_INVALID_GRAPH);
    CHECK(msbfs_csr_validate(2, 2, ro_start, ci_ok) == MSBFS_ERR_INVALID_GRAPH);
    CHECK(msbfs_csr_validate(2, 2, ro_ok, ci_oob) == MSBFS_ERR_INVALID_GRAPH);

    // sampling: all candidates, too many, isolated vertices skipped
    const uint32_t ro_iso[] = {0, 0, 1, 1, 2, 2};  // vertices 1 and 3 have edges
    uint32_t out[5];
    CHECK(msbfs_sample_sources(ro_iso, 5, 2, 9, 1, out) == MSBFS_OK && out[0] == 1 && out[1] == 3);
    CHECK(msbfs_sample_sources(ro_iso, 5, 3, 9, 1, out) == MSBFS_ERR_INVALID_ARG);
    CHECK(msbfs_sample_sources(ro_iso, 5, 5, 9, 0, out) == MSBFS_OK);
    for (uint32_t i = 0; i < 5; ++i) CHECK(out[i] == i);
    CHECK(msbfs_sample_sources(nullptr, 0, 0, 9, 0, nullptr) == MSBFS_OK);
}

// ------------------------------------------------------------ Matrix Market

// Parses `text` from a heap buffer of exactly its length, with no NUL after it.
static int parse_exact(const std::string &text, Csr *g = nullptr, msbfs_mtx_error *err = nullptr) {
    char *buf = (char *)std::malloc(text.size() ? text.size() : 1);
    if (!text.empty()) std::memcpy(buf, text.data(), text.size());
    msbfs_host_graph *h = nullptr;
    const int rc = msbfs_mtx_parse(buf, text.size(), &h);
    std::free(buf);
    if (rc == MSBFS_OK) {
        Csr c = take(h);
        CHECK(msbfs_csr_validate(c.V, c.E, c.ro.data(), c.ci.data()) == MSBFS_OK);
        if (g) *g = std::move(c);
    } else {
        CHECK(h == nullptr);
        CHECK(rc == MSBFS_ERR_PARSE);
        msbfs_mtx_error e{};
        CHECK(msbfs_mtx_last_error(&e) == MSBFS_OK);
        CHECK(e.kind >= MSBFS_MTX_IO && e.kind <= MSBFS_MTX_INDEX_OUT_OF_BOUNDS);
        if (err) *err = e;
    }
    return rc;
}

static void mtx_checks() {
    const std::string good =
        "%%MatrixMarket matrix coordinate real general\r\n"
        "% comment\n"
        "\n"
        "4 4 6\n"
        "1 2 0.5\n"
        "2 3 1e3\n"
        "3 1\n"
        "\t4 4 7\n"
        "2 1 -1\n"
        "1 2 9";
    Csr g;
    CHECK(parse_exact(good, &g) == MSBFS_OK);
    CHECK(g.V == 4 && g.E == 4);  // self loop dropped, duplicate (1,2) merged
    const uint32_t ro_want[] = {0, 1, 3, 4, 4}, ci_want[] = {1, 0, 2, 0};
    CHECK(std::equal(g.ro.begin(), g.ro.end(), ro_want) && std::equal(g.ci.begin(), g.ci.end(), ci_want));

    Csr s;
    CHECK(parse_exact("%%MatrixMarket matrix coordinate pattern symmetric\n3 3 2\n2 1\n3 2\n", &s) ==
          MSBFS_OK);
    CHECK(s.V == 3 && s.E == 4 && msbfs_csr_is_symmetric(s.V, s.E, s.ro.data(), s.ci.data()) == 1);

    // every prefix of both files: parsed or rejected, never read past the end
    for (const std::string &t : {good, std::string("%%MatrixMarket matrix coordinate pattern symmetric\n3 3 2\n2 1\n3 2\n")})
        for (size_t n = 0; n <= t.size(); ++n) parse_exact(t.substr(0, n));

    struct Bad {
        const char *text;
        int kind;
        uint64_t line;
    } bad[] = {
        {"", MSBFS_MTX_INVALID_HEADER, 0},
        {"%%MatrixMarket matrix array real general\n2 2\n", MSBFS_MTX_UNSUPPORTED_FORMAT, 1},
        {"%%MatrixMarket matrix coordinate real general\n2 x 1\n", MSBFS_MTX_INVALID_DIMENSIONS, 2},
        {"%%MatrixMarket matrix coordinate real general\n2 2 1\n1\n", MSBFS_MTX_INVALID_COORDINATE_LINE, 3},
        {"%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1\n", MSBFS_MTX_ZERO_INDEX, 3},
        {"%%MatrixMarket matrix coordinate real general\n2 2 1\n\n1 3\n", MSBFS_MTX_INDEX_OUT_OF_BOUNDS, 4},
        {"%%MatrixMarket matrix coordinate real general\n2 2 1\n1 18446744073709551616\n",
         MSBFS_MTX_INVALID_COORDINATE_LINE, 3},
        {"%%MatrixMarket matrix coordinate real general\n4294967296 1 0\n", MSBFS_MTX_INVALID_DIMENSIONS, 2},
    };
    for (auto &b : bad) {
        msbfs_mtx_error e{};
        CHECK(parse_exact(b.text, nullptr, &e) == MSBFS_ERR_PARSE);
        CHECK(e.kind == b.kind);
        if (b.line) CHECK(e.line_number == b.line);
    }

    // byte mutations of the good file
    std::mt19937_64 rng(1);
    const char alphabet[] = "0123456789 \t\n\r%-+.eExX";
    for (int i = 0; i < 20000; ++i) {
        std::string t = good;
        const int edits = 1 + (int)(rng() % 4);
        for (int k = 0; k < edits; ++k) {
            const size_t at = rng() % t.size();
            const uint64_t r = rng();
            if (r % 3 == 0) t[at] = alphabet[(r >> 8) % (sizeof(alphabet) - 1)];
            else if (r % 3 == 1) t[at] = (char)(r >> 8);
            else t.erase(at, 1 + (r >> 8) % 8);
            if (t.empty()) break;
        }
        parse_exact(t);
    }

    // a body of several MiB is parsed in threaded chunks: same graph as the
    // oracle's CSR of the same pairs, and the first error in file order wins
    const uint64_t V = 5000, n = 400000;
    std::vector<uint32_t> us(n), vs(n);
    std::string big = "%%MatrixMarket matrix coordinate integer general\n5000 5000 400000\n";
    for (uint64_t i = 0; i < n; ++i) {
        us[i] = (uint32_t)(rng() % V);
        vs[i] = (uint32_t)(rng() % V);
        big += std::to_string(us[i] + 1) + " " + std::to_string(vs[i] + 1) + " 1\n";
    }
    Csr bg;
    CHECK(big.size() > (3u << 20));
    CHECK(parse_exact(big, &bg) == MSBFS_OK);
    CHECK(same(bg, oracle_csr(V, us, vs)));
    const size_t early = big.find('\n', big.size() / 3) + 1;  // a line start a third in
    std::string broken = big + "0 1\n";                      // a zero index on the last line
    broken.insert(early, "1 9999\n");                        // and an earlier out-of-bounds one
    msbfs_mtx_error e{};
    CHECK(parse_exact(broken, nullptr, &e) == MSBFS_ERR_PARSE);
    CHECK(e.kind == MSBFS_MTX_INDEX_OUT_OF_BOUNDS && e.col == 9999);
    CHECK(e.line_number == 1 + (uint64_t)std::count(big.begin(), big.begin() + early, '\n'));
    CHECK(parse_exact(big + "0 1\n", nullptr, &e) == MSBFS_ERR_PARSE);
    CHECK(e.kind == MSBFS_MTX_ZERO_INDEX && e.line_number == 3 + n);

    msbfs_host_graph *h = nullptr;
    e = msbfs_mtx_error{};
    CHECK(msbfs_mtx_read("/nonexistent/dir/graph.mtx", &h) == MSBFS_ERR_PARSE && h == nullptr);
    CHECK(msbfs_mtx_last_error(&e) == MSBFS_OK && e.kind == MSBFS_MTX_IO);
}

// ------------------------------------------------------------ codecs

static void codec_checks() {
    std::mt19937_64 rng(3);
    for (size_t n : {0u, 1u, 2u, 3u, 4u, 5u, 7u, 64u, 1001u}) {
        std::vector<uint32_t> d(n), back(n);
        for (size_t i = 0; i < n; ++i) {
            const uint32_t pick[] = {0, 1, 0xFE, 0xFF, 0x100, 0xFFFE, 0xFFFF, 0x10000, INF,
                                     (uint32_t)rng()};
            d[i] = pick[rng() % 10];
        }
        std::vector<uint32_t> w16((n + 1) / 2 + (n == 0)), w8((n + 3) / 4 + (n == 0));
        orc_pack_u16(n, d.data(), w16.data());
        orc_unpack_u16(n, w16.data(), back.data());
        for (size_t i = 0; i < n; ++i) CHECK(back[i] == (d[i] >= 0xFFFF ? INF : d[i]));
        orc_pack_u8(n, d.data(), w8.data());
        orc_unpack_u8(n, w8.data(), back.data());
        for (size_t i = 0; i < n; ++i) CHECK(back[i] == (d[i] >= 0xFF ? INF : d[i]));
        if (n % 2) CHECK((w16[n / 2] >> 16) == 0xFFFF);  // the pad half reads as unreachable
    }
}

int main() {
    grid_checks();
    generator_checks();
    builder_checks();
    mtx_checks();
    codec_checks();
    std::printf("%s: %d checks, %d failed\n", failures ? "FAIL" : "OK", checks, failures);
    return failures ? 1 : 0;
}
