// Host-side graph data model of libmsbfs: canonical CSR construction,
// transpose, symmetry and validation checks, seeded generators for the
// benchmark workloads, and source sampling.  Semantics follow the
// reference's crates/core (file:line cited per function); the code and the
// random streams are this library's own.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

#include "common.hpp"

namespace msbfs {

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

int host_threads() {
    unsigned n = std::thread::hardware_concurrency();
    if (n == 0) n = 1;
    if (n > 64) n = 64;
    return (int)n;
}

template <class F>
static void parallel_for(uint64_t n, int threads, F &&f) {
    if (threads <= 1 || n < 4096) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    uint64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        uint64_t b = (uint64_t)t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&, b, e] { f(b, e); });
    }
    for (auto &x : th) x.join();
}

// pairs are (u << 32 | v).  Counting sort by u, then sort + dedup each row.
void csr_from_pairs(uint64_t V, const std::vector<uint64_t> &pairs, HostCsr &out, int threads) {
    out.V = V;
    std::vector<uint64_t> cnt(V + 1, 0);
    for (uint64_t p : pairs) {
        uint32_t u = (uint32_t)(p >> 32), v = (uint32_t)p;
        if (u != v && u < V && v < V) cnt[u + 1]++;
    }
    for (uint64_t i = 0; i < V; ++i) cnt[i + 1] += cnt[i];
    std::vector<uint32_t> tmp(cnt[V]);
    {
        std::vector<uint64_t> pos(cnt.begin(), cnt.end() - 1);
        for (uint64_t p : pairs) {
            uint32_t u = (uint32_t)(p >> 32), v = (uint32_t)p;
            if (u != v && u < V && v < V) tmp[pos[u]++] = v;
        }
    }
    std::vector<uint32_t> keep(V, 0);
    parallel_for(V, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t u = b; u < e; ++u) {
            uint32_t *beg = tmp.data() + cnt[u], *end = tmp.data() + cnt[u + 1];
            std::sort(beg, end);
            keep[u] = (uint32_t)(std::unique(beg, end) - beg);
        }
    });
    out.ro.assign(V + 1, 0);
    for (uint64_t u = 0; u < V; ++u) out.ro[u + 1] = out.ro[u] + keep[u];
    out.ci.resize(out.ro[V]);
    parallel_for(V, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t u = b; u < e; ++u)  // (an empty graph's vectors hold no storage)
            if (keep[u])
                std::memcpy(out.ci.data() + out.ro[u], tmp.data() + cnt[u], keep[u] * sizeof(uint32_t));
    });
}

void csr_transpose(uint64_t V, const uint32_t *ro, const uint32_t *ci, uint32_t *co,
                   uint32_t *ri) {
    uint64_t E = ro[V];
    std::vector<uint32_t> pos(V + 1, 0);
    for (uint64_t e = 0; e < E; ++e) pos[ci[e] + 1]++;
    co[0] = 0;
    for (uint64_t i = 0; i < V; ++i) co[i + 1] = co[i] + pos[i + 1];
    for (uint64_t i = 0; i < V; ++i) pos[i] = co[i];
    for (uint64_t u = 0; u < V; ++u)
        for (uint32_t e = ro[u]; e < ro[u + 1]; ++e) ri[pos[ci[e]]++] = (uint32_t)u;
}

int csr_lattice_offsets(uint64_t V, const uint32_t *ro, const uint32_t *ci, uint32_t d[4]) {
    for (int k = 0; k < 4; ++k) d[k] = 0;
    int K = 0;
    for (uint64_t u = 0; u < V; ++u)
        for (uint32_t j = ro[u]; j < ro[u + 1]; ++j) {
            if (ci[j] <= u) continue;  // self-loops reach nothing; lower ids are mirrored
            const uint32_t dd = ci[j] - (uint32_t)u;
            int k = 0;
            while (k < K && d[k] != dd) ++k;
            if (k < K) continue;
            if (K == 4) return -1;
            d[K++] = dd;
        }
    std::sort(d, d + K);
    return K;
}

uint32_t host_levels_from(uint64_t V, const uint32_t *ro, const uint32_t *ci, uint32_t s) {
    if (V == 0 || s >= V) return 0;
    std::vector<uint32_t> dist(V, 0xFFFFFFFFu), q;
    q.reserve(V);
    dist[s] = 0;
    q.push_back(s);
    for (size_t h = 0; h < q.size(); ++h) {
        const uint32_t u = q[h];
        for (uint32_t j = ro[u]; j < ro[u + 1]; ++j)
            if (dist[ci[j]] == 0xFFFFFFFFu) {
                dist[ci[j]] = dist[u] + 1;
                q.push_back(ci[j]);
            }
    }
    return dist[q.back()] + 1;  // (the queue ends with a farthest vertex)
}

bool csr_is_symmetric(uint64_t V, uint64_t E, const uint32_t *ro, const uint32_t *ci) {
    std::vector<uint32_t> co(V + 1), ri(E);
    csr_transpose(V, ro, ci, co.data(), ri.data());
    return std::equal(co.begin(), co.end(), ro) && std::equal(ri.begin(), ri.end(), ci);
}

}  // namespace msbfs

using namespace msbfs;

msbfs_host_graph *msbfs::wrap_host_graph(HostCsr &&c) {
    auto *h = new msbfs_host_graph;
    h->csr = std::move(c);
    return h;
}

extern "C" {

const char *msbfs_last_error(void) { return g_last_error.c_str(); }

const char *msbfs_status_string(int s) {
    switch (s) {
    case MSBFS_OK: return "ok";
    case MSBFS_ERR_INVALID_ARG: return "invalid argument";
    case MSBFS_ERR_SOURCE_OOB: return "source out of bounds";
    case MSBFS_ERR_EMPTY_GRAPH: return "empty graph";
    case MSBFS_ERR_OUT_TOO_SMALL: return "output buffer too small";
    case MSBFS_ERR_HIP: return "HIP runtime error";
    case MSBFS_ERR_OOM: return "out of device memory";
    case MSBFS_ERR_DIAMETER_U16: return "distance does not fit u16";
    case MSBFS_ERR_INVALID_GRAPH: return "invalid graph";
    case MSBFS_ERR_UNSUPPORTED: return "unsupported";
    case MSBFS_ERR_NO_DEVICE: return "no device";
    case MSBFS_ERR_BUSY: return "busy";
    case MSBFS_ERR_PARSE: return "Matrix Market parse error";
    default: return "unknown status";
    }
}

int msbfs_host_graph_get(const msbfs_host_graph *h, uint64_t *V, uint64_t *E,
                         const uint32_t **ro, const uint32_t **ci) {
    if (!h) return fail(MSBFS_ERR_INVALID_ARG, "null host graph");
    if (V) *V = h->csr.V;
    if (E) *E = h->csr.E();
    if (ro) *ro = h->csr.ro.data();
    if (ci) *ci = h->csr.ci.data();
    return MSBFS_OK;
}

void msbfs_host_graph_free(msbfs_host_graph *h) { delete h; }

int msbfs_csr_from_edge_list(uint64_t V, uint64_t n, const uint32_t *us, const uint32_t *vs,
                             msbfs_host_graph **out) {
    if (!out || (n && (!us || !vs))) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    if (V > 0xFFFFFFFFull) return fail(MSBFS_ERR_INVALID_ARG, "num_vertices exceeds u32 range");
    std::vector<uint64_t> pairs(n);
    for (uint64_t i = 0; i < n; ++i) pairs[i] = ((uint64_t)us[i] << 32) | vs[i];
    HostCsr c;
    csr_from_pairs(V, pairs, c, host_threads());
    *out = wrap_host_graph(std::move(c));
    return MSBFS_OK;
}

int msbfs_csr_from_undirected_pairs(uint64_t V, uint64_t n, const uint32_t *us,
                                    const uint32_t *vs, msbfs_host_graph **out) {
    if (!out || (n && (!us || !vs))) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    if (V > 0xFFFFFFFFull) return fail(MSBFS_ERR_INVALID_ARG, "num_vertices exceeds u32 range");
    std::vector<uint64_t> pairs;
    pairs.reserve(2 * n);
    for (uint64_t i = 0; i < n; ++i) {
        pairs.push_back(((uint64_t)us[i] << 32) | vs[i]);
        pairs.push_back(((uint64_t)vs[i] << 32) | us[i]);
    }
    HostCsr c;
    csr_from_pairs(V, pairs, c, host_threads());
    *out = wrap_host_graph(std::move(c));
    return MSBFS_OK;
}

int msbfs_csr_validate(uint64_t V, uint64_t E, const uint32_t *ro, const uint32_t *ci) {
    if (!ro) return fail(MSBFS_ERR_INVALID_GRAPH, "row_offsets is null");
    if (E && !ci) return fail(MSBFS_ERR_INVALID_GRAPH, "col_indices is null");
    if (ro[0] != 0) return fail(MSBFS_ERR_INVALID_GRAPH, "row_offsets[0] = %u, expected 0", ro[0]);
    for (uint64_t i = 0; i < V; ++i)
        if (ro[i] > ro[i + 1])
            return fail(MSBFS_ERR_INVALID_GRAPH,
                        "row_offsets not monotonic at %llu: %u > %u",
                        (unsigned long long)i, ro[i], ro[i + 1]);
    if ((uint64_t)ro[V] != E)
        return fail(MSBFS_ERR_INVALID_GRAPH, "row_offsets[V] = %u but num_edges = %llu", ro[V],
                    (unsigned long long)E);
    for (uint64_t e = 0; e < E; ++e)
        if ((uint64_t)ci[e] >= V)
            return fail(MSBFS_ERR_INVALID_GRAPH,
                        "col_indices[%llu] = %u out of bounds for %llu vertices",
                        (unsigned long long)e, ci[e], (unsigned long long)V);
    return MSBFS_OK;
}

int msbfs_csr_to_csc(uint64_t V, uint64_t E, const uint32_t *ro, const uint32_t *ci,
                     uint32_t *co, uint32_t *ri) {
    int rc = msbfs_csr_validate(V, E, ro, ci);
    if (rc) return rc;
    if (!co || (E && !ri)) return fail(MSBFS_ERR_INVALID_ARG, "null output");
    csr_transpose(V, ro, ci, co, ri);
    return MSBFS_OK;
}

int msbfs_csr_is_symmetric(uint64_t V, uint64_t E, const uint32_t *ro, const uint32_t *ci) {
    int rc = msbfs_csr_validate(V, E, ro, ci);
    if (rc) return -rc;
    return csr_is_symmetric(V, E, ro, ci) ? 1 : 0;
}

int msbfs_csr_lattice_offsets(uint64_t V, uint64_t E, const uint32_t *ro, const uint32_t *ci,
                              uint32_t *offsets) {
    if (!offsets) return -fail(MSBFS_ERR_INVALID_ARG, "null offsets");
    int rc = msbfs_csr_validate(V, E, ro, ci);
    if (rc) return -rc;
    if (!csr_is_symmetric(V, E, ro, ci)) {
        for (int k = 0; k < 4; ++k) offsets[k] = 0;
        return 0;
    }
    const int k = csr_lattice_offsets(V, ro, ci, offsets);
    if (k < 0) for (int i = 0; i < 4; ++i) offsets[i] = 0;
    return k < 0 ? 0 : k;
}

// Grid2D::new (crates/core/src/generators.rs:28-80): row-major ids, each
// row list emitted in North, West, East, South order (already ascending).
int msbfs_gen_grid2d(uint64_t width, uint64_t height, msbfs_host_graph **out) {
    if (!out) return fail(MSBFS_ERR_INVALID_ARG, "null out");
    if (width == 0 || height == 0)
        return fail(MSBFS_ERR_INVALID_ARG, "invalid grid dimensions %llux%llu",
                    (unsigned long long)width, (unsigned long long)height);
    if (width * height > 0xFFFFFFFFull) return fail(MSBFS_ERR_INVALID_ARG, "grid too large");
    HostCsr c;
    c.V = width * height;
    c.ro.reserve(c.V + 1);
    c.ci.reserve(2 * ((width - 1) * height + width * (height - 1)));
    c.ro.push_back(0);
    for (uint64_t y = 0; y < height; ++y)
        for (uint64_t x = 0; x < width; ++x) {
            uint64_t v = y * width + x;
            if (y > 0) c.ci.push_back((uint32_t)(v - width));
            if (x > 0) c.ci.push_back((uint32_t)(v - 1));
            if (x + 1 < width) c.ci.push_back((uint32_t)(v + 1));
            if (y + 1 < height) c.ci.push_back((uint32_t)(v + width));
            c.ro.push_back((uint32_t)c.ci.size());
        }
    *out = wrap_host_graph(std::move(c));
    return MSBFS_OK;
}

// Road-like high-diameter graph: vertices on a `width`-column grid (last row
// partial), a uniformly shuffled list of all grid edges, Kruskal with
// union-find keeps the edges that join two components (a random spanning
// tree), and the first `extra` rejected edges of the same shuffled order are
// added back (a uniform sample of distinct non-tree grid edges).
int msbfs_gen_road_like(uint64_t V, uint64_t width, uint64_t extra, uint64_t seed,
                        msbfs_host_graph **out) {
    if (!out) return fail(MSBFS_ERR_INVALID_ARG, "null out");
    if (V == 0 || width == 0) return fail(MSBFS_ERR_INVALID_ARG, "empty road-like graph");
    if (V > 0xFFFFFFFFull) return fail(MSBFS_ERR_INVALID_ARG, "graph too large");
    std::vector<uint64_t> edges;
    for (uint64_t v = 0; v < V; ++v) {
        uint64_t x = v % width;
        if (x + 1 < width && v + 1 < V) edges.push_back((v << 32) | (v + 1));
        if (v + width < V) edges.push_back((v << 32) | (v + width));
    }
    Rng rng(seed, 0x524F4144ull /* "ROAD" */);
    for (uint64_t i = edges.size(); i > 1; --i) std::swap(edges[i - 1], edges[rng.below(i)]);
    std::vector<uint32_t> parent(V), rank(V, 0);
    std::iota(parent.begin(), parent.end(), 0u);
    auto find = [&](uint32_t a) {
        while (parent[a] != a) {
            parent[a] = parent[parent[a]];
            a = parent[a];
        }
        return a;
    };
    std::vector<uint64_t> kept, rejected;
    kept.reserve(V - 1);
    for (uint64_t e : edges) {
        uint32_t a = find((uint32_t)(e >> 32)), b = find((uint32_t)e);
        if (a == b) {
            rejected.push_back(e);
            continue;
        }
        if (rank[a] < rank[b]) std::swap(a, b);
        parent[b] = a;
        if (rank[a] == rank[b]) rank[a]++;
        kept.push_back(e);
    }
    if (extra > rejected.size())
        return fail(MSBFS_ERR_INVALID_ARG, "requested %llu extra edges, only %llu available",
                    (unsigned long long)extra, (unsigned long long)rejected.size());
    kept.insert(kept.end(), rejected.begin(), rejected.begin() + extra);
    std::vector<uint64_t> pairs;
    pairs.reserve(2 * kept.size());
    for (uint64_t e : kept) {
        pairs.push_back(e);
        pairs.push_back((e << 32) | (e >> 32));
    }
    HostCsr c;
    csr_from_pairs(V, pairs, c, host_threads());
    *out = wrap_host_graph(std::move(c));
    return MSBFS_OK;
}

// R-MAT (crates/core/src/generators.rs:277-342): for each of
// edge_factor * 2^scale samples, descend `scale` quadrant choices with
// thresholds a, a+b, a+b+c; drop self loops; symmetrise unless directed.
// Samples are drawn in fixed chunks of 2^20, chunk k from stream k, so the
// result does not depend on the host thread count.
int msbfs_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, double d,
                   uint64_t seed, int directed, msbfs_host_graph **out) {
    if (!out) return fail(MSBFS_ERR_INVALID_ARG, "null out");
    if (scale == 0) return fail(MSBFS_ERR_INVALID_ARG, "R-MAT scale must be >= 1");
    if (scale > 30) return fail(MSBFS_ERR_INVALID_ARG, "R-MAT scale %u too large", scale);
    double sum = a + b + c + d;
    if (std::fabs(sum - 1.0) > 1e-6)
        return fail(MSBFS_ERR_INVALID_ARG, "R-MAT probabilities sum to %f", sum);
    const uint64_t V = 1ull << scale;
    const uint64_t M = (uint64_t)edge_factor * V;
    const double ta = a, tb = a + b, tc = a + b + c;
    const uint64_t CH = 1ull << 20;
    const uint64_t nch = (M + CH - 1) / CH;
    std::vector<uint64_t> raw(M, ~0ull);
    std::atomic<uint64_t> next{0};
    int T = host_threads();
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&] {
            for (;;) {
                uint64_t k = next.fetch_add(1);
                if (k >= nch) break;
                Rng rng(seed, 0x524D4154ull + k);
                uint64_t e0 = k * CH, e1 = std::min(M, e0 + CH);
                for (uint64_t i = e0; i < e1; ++i) {
                    uint32_t u = 0, v = 0;
                    for (int step = (int)scale - 1; step >= 0; --step) {
                        uint32_t bit = 1u << step;
                        double r = rng.uniform();
                        if (r < ta) {
                        } else if (r < tb) {
                            v |= bit;
                        } else if (r < tc) {
                            u |= bit;
                        } else {
                            u |= bit;
                            v |= bit;
                        }
                    }
                    if (u != v) raw[i] = ((uint64_t)u << 32) | v;
                }
            }
        });
    for (auto &x : th) x.join();
    std::vector<uint64_t> pairs;
    pairs.reserve(directed ? M : 2 * M);
    for (uint64_t p : raw) {
        if (p == ~0ull) continue;
        pairs.push_back(p);
        if (!directed) pairs.push_back((p << 32) | (p >> 32));
    }
    raw.clear();
    raw.shrink_to_fit();
    HostCsr cs;
    csr_from_pairs(V, pairs, cs, T);
    *out = wrap_host_graph(std::move(cs));
    return MSBFS_OK;
}

// G(n, p) by Batagelj-Brandes geometric skipping (Batagelj & Brandes,
// "Efficient generation of large random networks", 2005), undirected; the
// reference uses the same method (crates/core/src/generators.rs:139-191).
int msbfs_gen_erdos_renyi_np(uint64_t V, double p, uint64_t seed, msbfs_host_graph **out) {
    if (!out) return fail(MSBFS_ERR_INVALID_ARG, "null out");
    if (V == 0) return fail(MSBFS_ERR_INVALID_ARG, "G(n,p) needs n >= 1");
    if (!(p >= 0.0 && p <= 1.0)) return fail(MSBFS_ERR_INVALID_ARG, "invalid probability %f", p);
    if (V > 0xFFFFFFFFull) return fail(MSBFS_ERR_INVALID_ARG, "graph too large");
    std::vector<uint64_t> pairs;
    if (V > 1 && p > 0.0) {
        Rng rng(seed, 0x45524E50ull /* "ERNP" */);
        if (p >= 0.5) {
            for (uint64_t u = 0; u + 1 < V; ++u)
                for (uint64_t v = u + 1; v < V; ++v)
                    if (rng.uniform() < p) {
                        pairs.push_back((u << 32) | v);
                        pairs.push_back((v << 32) | u);
                    }
        } else {
            const double lq = std::log(1.0 - p);
            uint64_t u = 0, v = 1;
            while (u < V - 1) {
                double r = rng.uniform();
                if (r <= 0.0) r = 0x1.0p-53;
                double kf = std::floor(std::log(r) / lq);
                uint64_t k = kf > 1e18 ? (uint64_t)1e18 : (uint64_t)kf;
                v += k;
                // Carry the overflow into the next row, whose first
                // candidate is u + 2.  (The reference writes `v += u` here,
                // generators.rs:170-175, which skips u extra pairs per row
                // and halves the expected edge count.)
                while (v >= V && u < V - 1) {
                    v -= V - 1 - u;
                    u += 1;
                    v += 1;
                }
                if (u < V - 1) {
                    pairs.push_back((u << 32) | v);
                    pairs.push_back((v << 32) | u);
                    v += 1;
                }
            }
        }
    }
    HostCsr c;
    csr_from_pairs(V, pairs, c, 1);
    *out = wrap_host_graph(std::move(c));
    return MSBFS_OK;
}

int msbfs_sample_sources(const uint32_t *ro, uint64_t V, uint64_t count, uint64_t seed,
                         int non_isolated, uint32_t *out) {
    if (!out && count) return fail(MSBFS_ERR_INVALID_ARG, "null out");
    if (non_isolated && !ro) return fail(MSBFS_ERR_INVALID_ARG, "row_offsets needed");
    std::vector<uint32_t> cand;
    cand.reserve(V);
    for (uint64_t v = 0; v < V; ++v)
        if (!non_isolated || ro[v + 1] > ro[v]) cand.push_back((uint32_t)v);
    if (count > cand.size())
        return fail(MSBFS_ERR_INVALID_ARG, "requested %llu sources from %llu candidates",
                    (unsigned long long)count, (unsigned long long)cand.size());
    Rng rng(seed, 0x53524353ull /* "SRCS" */);
    for (uint64_t i = 0; i < count; ++i) {
        uint64_t j = i + rng.below(cand.size() - i);
        std::swap(cand[i], cand[j]);
    }
    std::sort(cand.begin(), cand.begin() + count);
    std::copy(cand.begin(), cand.begin() + count, out);
    return MSBFS_OK;
}

}  // extern "C"
