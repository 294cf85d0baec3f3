// Host engine of libmsbfs: device context, resident graphs, the engine
// object with its reusable workspace, method routing and kernel launches.
//
// Mirrors the reference's crates/gpu_bfs/src/engine.rs (GpuBfsEngine): the
// same method enum, the same Auto decision table, the same argument checks
// and error behaviour; underneath, the wgpu buffers/pipelines and the
// host-driven level loops are replaced by device-resident graphs and
// persistent HIP kernels that iterate every level on the device.
#include "engine_internal.hpp"

#include <cstring>
#include <thread>

using namespace msbfs;

#define MSBFS_VERSION "0.1.0"

extern "C" const char *msbfs_version(void) { return MSBFS_VERSION; }

// ------------------------------------------------------------------ context

extern "C" int msbfs_context_create(int device, msbfs_context **out) {
    clear_error();
    if (!out) return fail(MSBFS_ERR_INVALID_ARG, "null out");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(MSBFS_ERR_NO_DEVICE, "no HIP device available (%s)",
                    e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    if (device < 0 || device >= n)
        return fail(MSBFS_ERR_INVALID_ARG, "device %d out of range (%d devices)", device, n);
    auto *c = new msbfs_context;
    c->device = device;
    if ((e = hipSetDevice(device)) != hipSuccess ||
        (e = hipGetDeviceProperties(&c->prop, device)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete c;
        return hip_fail(e, "context init");
    }
    c->cus = c->prop.multiProcessorCount;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
            hipSuccess ||
        lds <= 0)
        lds = (int)c->prop.sharedMemPerBlock;
    c->lds_max = (size_t)lds;
    int coop = 0;
    (void)hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device);
    c->coop = coop;
    int rc = noalloc_setup(c);  // this device's copy of the no-alloc scratch array
    if (rc) {
        context_release(c);
        return rc;
    }
    *out = c;
    return MSBFS_OK;
}

void msbfs::context_release(msbfs_context *c) {
    if (!c || c->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    if (c->coop_done) (void)hipEventDestroy(c->coop_done);
    for (auto &se : c->stream_done) (void)hipEventDestroy(se.second);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int msbfs::LaunchOrder::begin() {
    if (c_->coop_valid) HIP_TRY(hipStreamWaitEvent(st_, c_->coop_done, 0));
    if (coop_)
        for (auto &se : c_->stream_done)
            if (se.first != st_) HIP_TRY(hipStreamWaitEvent(st_, se.second, 0));
    return MSBFS_OK;
}

int msbfs::LaunchOrder::end() {
    if (coop_) {
        if (!c_->coop_done)
            HIP_TRY(hipEventCreateWithFlags(&c_->coop_done, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(c_->coop_done, st_));
        c_->coop_valid = true;
        return MSBFS_OK;
    }
    for (auto &se : c_->stream_done)
        if (se.first == st_) {
            HIP_TRY(hipEventRecord(se.second, st_));
            return MSBFS_OK;
        }
    hipEvent_t ev = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(ev, st_));
    c_->stream_done.emplace_back(st_, ev);
    return MSBFS_OK;
}

extern "C" int msbfs_context_destroy(msbfs_context *c) {
    context_release(c);
    return MSBFS_OK;
}

extern "C" int msbfs_context_info(const msbfs_context *c, msbfs_device_info *info) {
    if (!c || !info) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    std::memset(info, 0, sizeof(*info));
    std::snprintf(info->name, sizeof(info->name), "%s", c->prop.name);
    std::snprintf(info->arch, sizeof(info->arch), "%s", c->prop.gcnArchName);
    info->compute_units = c->cus;
    info->lds_bytes_per_block = (int32_t)c->lds_max;
    info->total_mem_bytes = c->prop.totalGlobalMem;
    info->cooperative_launch = c->coop;
    return MSBFS_OK;
}

extern "C" int msbfs_context_synchronize(msbfs_context *c) {
    if (!c) return fail(MSBFS_ERR_INVALID_ARG, "null context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipDeviceSynchronize());
    return MSBFS_OK;
}

// ------------------------------------------------------------------ graph

extern "C" int msbfs_graph_create(msbfs_context *c, uint64_t V, uint64_t E, const uint32_t *ro,
                                  const uint32_t *ci, msbfs_graph **out) {
    clear_error();
    return graph_create_impl(c, V, E, ro, ci, nullptr, out);
}

// vbound (batch union graphs, else null): per vertex, the source count of
// its graph, kept in the pull's light and heavy in-list items.
int msbfs::graph_create_impl(msbfs_context *c, uint64_t V, uint64_t E, const uint32_t *ro,
                             const uint32_t *ci, const uint32_t *vbound, msbfs_graph **out) {
    if (!c || !out) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    if (V > 0xFFFFFFFEull) return fail(MSBFS_ERR_INVALID_GRAPH, "num_vertices exceeds u32");
    if (E > 0xFFFFFFFFull) return fail(MSBFS_ERR_INVALID_GRAPH, "num_edges exceeds u32");
    int rc = msbfs_csr_validate(V, E, ro, ci);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(c->device));
    auto *g = new msbfs_graph;
    c->refs.fetch_add(1);
    g->ctx = c;
    g->V = V;
    g->E = E;
    for (uint64_t v = 0; v < V; ++v) g->max_degree = std::max(g->max_degree, ro[v + 1] - ro[v]);
    g->symmetric = csr_is_symmetric(V, E, ro, ci) ? 1 : 0;
    std::vector<uint32_t> co, ri;
    if (!g->symmetric) {
        co.resize(V + 1);
        ri.resize(E);
        csr_transpose(V, ro, ci, co.data(), ri.data());
    }
    auto upload = [&](uint32_t **dst, const uint32_t *src, uint64_t n) -> hipError_t {
        hipError_t e = hipMalloc((void **)dst, std::max<uint64_t>(n, 1) * sizeof(uint32_t));
        if (e != hipSuccess) return e;
        g->bytes += n * sizeof(uint32_t);
        if (n) return hipMemcpy(*dst, src, n * sizeof(uint32_t), hipMemcpyHostToDevice);
        return hipSuccess;
    };
    hipError_t e = upload(&g->ro, ro, V + 1);
    if (e == hipSuccess) e = upload(&g->ci, ci, E);
    if (e == hipSuccess) {
        if (g->symmetric) {
            g->co = g->ro;
            g->ri = g->ci;
        } else {
            e = upload(&g->co, co.data(), V + 1);
            if (e == hipSuccess) e = upload(&g->ri, ri.data(), E);
        }
    }
    // Fixed-width neighbour table for low-degree graphs: the deep family
    // fetches a vertex's whole neighbourhood with one vector load.  Empty
    // slots hold the vertex itself: a frontier vertex's own bits are already
    // visited, so claiming it finds nothing new and needs no test.
    if (e == hipSuccess && V > 0 && g->max_degree <= 8) {
        g->ell_width = g->max_degree <= 4 ? 4 : 8;
        std::vector<uint32_t> ell(V * g->ell_width);
        for (uint64_t v = 0; v < V; ++v) {
            const uint32_t d = ro[v + 1] - ro[v];
            for (uint32_t k = 0; k < (uint32_t)g->ell_width; ++k)
                ell[v * g->ell_width + k] = k < d ? ci[ro[v] + k] : (uint32_t)v;
        }
        e = upload(&g->ell, ell.data(), ell.size());
    }
    // Lattice adjacency: when every edge of a symmetric graph joins ids that
    // differ by one of at most four offsets (grids and grid-derived road
    // models numbered row by row), a few bits per vertex hold the whole
    // graph, small enough for LDS (road-like 114,599 vertices on a
    // 339-column grid: offsets 1 and 339, 28.6 KB).  Self-loops reach
    // nothing new and are left out.
    if (e == hipSuccess && g->symmetric && V > 1 && E > 0) {
        uint32_t d[4];
        const int K = csr_lattice_offsets(V, ro, ci, d);
        if (K > 0) {
            const int lat = K <= 2 ? 2 : 4;
            const uint32_t fpw = 32 / lat;
            // zero fields below vertex 0 (whole words, at least the largest
            // offset), so the kernel looks up a lower neighbour's field
            // without a bounds test; the total padded to whole 16-byte
            // groups: the kernel copies it to LDS in the size it reserves
            const uint64_t pad = (d[K - 1] + fpw - 1) / fpw * fpw;
            std::vector<uint32_t> words(((pad + V + fpw - 1) / fpw + 3) / 4 * 4, 0u);
            for (uint64_t u = 0; u < V; ++u)
                for (uint32_t j = ro[u]; j < ro[u + 1]; ++j) {
                    if (ci[j] <= u) continue;
                    const uint32_t dd = ci[j] - (uint32_t)u;
                    int k = 0;
                    while (d[k] != dd) ++k;
                    const uint64_t x = pad + u;
                    words[x / fpw] |= 1u << ((x % fpw) * lat + k);
                }
            e = upload(&g->lat, words.data(), words.size());
            if (e == hipSuccess) {
                g->lat_bits = lat;
                g->lat_pad = (uint32_t)pad;
                g->lat_words = (uint32_t)words.size();
                std::memcpy(g->lat_d, d, sizeof(d));
                // The routing of one-source tasks to the lattice kernel
                // needs the graph's level count, which a first traversal
                // records; for a lattice small enough for LDS, one host
                // BFS from vertex 0 gives it now (its eccentricity + 1, at
                // least half the diameter), so the first traversal is
                // routed like the later ones (road-like S=2048: 1.6 ms
                // instead of 3.4 ms in the workgroup kernel).
                if ((uint64_t)g->lat_words * 4 < (uint64_t)c->lds_max)
                    g->deep_levels.store(host_levels_from(V, ro, ci, 0), std::memory_order_relaxed);
            }
        }
    }
    // Vertices whose edge lists the wide family hands to a whole wave.
    if (e == hipSuccess) {
        std::vector<uint32_t> hout, hin;
        for (uint64_t v = 0; v < V; ++v)
            if (ro[v + 1] - ro[v] > kHeavyDegree) hout.push_back((uint32_t)v);
        // Heavy edge lists split into work items of at most `len` edges.
        auto chunks = [vbound](const std::vector<uint32_t> &hv, const uint32_t *off, uint32_t len) {
            std::vector<uint32_t> out;
            for (uint32_t v : hv)
                for (uint32_t b = off[v]; b < off[v + 1]; b += len) {
                    out.push_back(v);
                    out.push_back(b);
                    out.push_back(std::min(off[v + 1], b + len));
                    out.push_back(vbound ? vbound[v] : 0u);
                }
            return out;
        };
        g->n_heavy_out = (uint32_t)hout.size();
        e = upload(&g->heavy_out, hout.data(), hout.size());
        std::vector<uint32_t> cout = chunks(hout, ro, kHeavyChunkOut);
        g->n_hchunk_out = (uint32_t)(cout.size() / 4);
        if (e == hipSuccess) e = upload(&g->hchunk_out, cout.data(), cout.size());
        if (e == hipSuccess) {
            if (g->symmetric) {
                g->heavy_in = g->heavy_out;
                g->n_heavy_in = g->n_heavy_out;
                hin = hout;
            } else {
                for (uint64_t v = 0; v < V; ++v)
                    if (co[v + 1] - co[v] > kHeavyDegree) hin.push_back((uint32_t)v);
                g->n_heavy_in = (uint32_t)hin.size();
                e = upload(&g->heavy_in, hin.data(), hin.size());
            }
            std::vector<uint32_t> cin = chunks(hin, g->symmetric ? ro : co.data(), kHeavyChunk);
            g->n_hchunk_in = (uint32_t)(cin.size() / 4);
            if (e == hipSuccess) e = upload(&g->hchunk_in, cin.data(), cin.size());
        }
        // The pull's light vertices with their in-list bounds: vertices
        // without in-edges (38% of R-MAT scale 20) can gain nothing and are
        // not visited at all.
        if (e == hipSuccess) {
            const uint32_t *cv = g->symmetric ? ro : co.data();
            std::vector<uint32_t> lin;
            for (uint64_t v = 0; v < V; ++v) {
                const uint32_t d = cv[v + 1] - cv[v];
                if (d == 0 || d > kHeavyDegree) continue;
                lin.push_back((uint32_t)v);
                lin.push_back(cv[v]);
                lin.push_back(cv[v + 1]);
                lin.push_back(vbound ? vbound[v] : 0u);
            }
            g->n_light_in = (uint32_t)(lin.size() / 4);
            e = upload(&g->light_in, lin.data(), lin.size());
            // The same items by decreasing in-degree (stable): tiles of
            // similar in-lists, the longest first.  Used when a vertex's
            // words fill whole cache lines, so vertex order buys no
            // locality (R-MAT scale 20, 2,048 sources: 5.46 -> 5.15 ms; with
            // 64 sources, 8-byte rows, vertex order stays 4% faster).
            if (e == hipSuccess) {
                std::vector<uint4> it(g->n_light_in);
                std::memcpy(it.data(), lin.data(), lin.size() * 4);
                std::stable_sort(it.begin(), it.end(), [](const uint4 &a, const uint4 &b) {
                    return a.z - a.y > b.z - b.y;
                });
                e = upload(&g->light_in_deg, (const uint32_t *)it.data(), lin.size());
            }
        }
    }
    if (e != hipSuccess) {
        msbfs_graph_destroy(g);
        return hip_fail(e, "graph upload");
    }
    *out = g;
    return MSBFS_OK;
}

extern "C" int msbfs_graph_destroy(msbfs_graph *g) {
    graph_release(g);
    return MSBFS_OK;
}

void msbfs::graph_release(msbfs_graph *g) {
    if (!g || g->refs.fetch_sub(1) != 1) return;
    msbfs_context *ctx = g->ctx;
    (void)hipSetDevice(ctx->device);
    if (g->co && g->co != g->ro) (void)hipFree(g->co);
    if (g->ri && g->ri != g->ci) (void)hipFree(g->ri);
    if (g->ro) (void)hipFree(g->ro);
    if (g->ci) (void)hipFree(g->ci);
    if (g->ell) (void)hipFree(g->ell);
    if (g->lat) (void)hipFree(g->lat);
    if (g->heavy_in && g->heavy_in != g->heavy_out) (void)hipFree(g->heavy_in);
    if (g->heavy_out) (void)hipFree(g->heavy_out);
    if (g->hchunk_in) (void)hipFree(g->hchunk_in);
    if (g->hchunk_out) (void)hipFree(g->hchunk_out);
    if (g->light_in) (void)hipFree(g->light_in);
    if (g->light_in_deg) (void)hipFree(g->light_in_deg);
    delete g;
    context_release(ctx);
}

extern "C" int msbfs_graph_get_info(const msbfs_graph *g, msbfs_graph_info *info) {
    if (!g || !info) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    info->num_vertices = g->V;
    info->num_edges = g->E;
    info->max_degree = g->max_degree;
    info->symmetric = (uint32_t)g->symmetric;
    info->device_bytes = g->bytes;
    return MSBFS_OK;
}

// ------------------------------------------------------------------ engine

extern "C" int msbfs_engine_create(msbfs_context *c, int method, msbfs_engine **out) {
    clear_error();
    if (!c || !out) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    if (method < 0 || method > MSBFS_METHOD_AUTO)
        return fail(MSBFS_ERR_INVALID_ARG, "unknown method %d", method);
    HIP_TRY(hipSetDevice(c->device));
    auto *e = new msbfs_engine;
    c->refs.fetch_add(1);
    e->ctx = c;
    e->method = method;
    for (auto &ev : e->ev) {
        hipError_t er = hipEventCreate(&ev);
        if (er != hipSuccess) {
            msbfs_engine_destroy(e);
            return hip_fail(er, "hipEventCreate");
        }
    }
    *out = e;
    return MSBFS_OK;
}

void msbfs::engine_release(msbfs_engine *e) {
    if (!e || e->refs.fetch_sub(1) != 1) return;
    msbfs_context *ctx = e->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    graph_release(e->pending_graph);
    for (auto &ev : e->ev)
        if (ev) (void)hipEventDestroy(ev);
    delete e;
    context_release(ctx);
}

extern "C" int msbfs_engine_destroy(msbfs_engine *e) {
    engine_release(e);
    return MSBFS_OK;
}

extern "C" int msbfs_engine_method(const msbfs_engine *e) { return e ? e->method : -1; }

extern "C" int msbfs_engine_set_push_pull(msbfs_engine *e, uint32_t alpha, uint32_t beta) {
    if (!e) return fail(MSBFS_ERR_INVALID_ARG, "null engine");
    std::lock_guard<std::mutex> lk(e->mu);
    e->tune.alpha = alpha;
    e->tune.beta = beta;
    return MSBFS_OK;
}

extern "C" int msbfs_tuning_validate(const msbfs_tuning *t) {
    if (!t) return fail(MSBFS_ERR_INVALID_ARG, "null tuning");
    if (t->tile_words && (t->tile_words > 64 || (t->tile_words & (t->tile_words - 1))))
        return fail(MSBFS_ERR_INVALID_ARG, "tile_words must be a power of two <= 64");
    if (t->words_per_lane > 2)
        return fail(MSBFS_ERR_INVALID_ARG, "words_per_lane must be 0, 1 or 2");
    if (t->deep_bits && (t->deep_bits > 32 || (t->deep_bits & (t->deep_bits - 1))))
        return fail(MSBFS_ERR_INVALID_ARG, "deep_bits must be 1, 2, 4, 8, 16 or 32");
    // the deep kernels stride whole waves: a workgroup is 1..16 waves
    if (t->deep_threads && (t->deep_threads % 64 || t->deep_threads > 1024))
        return fail(MSBFS_ERR_INVALID_ARG, "deep_threads must be a multiple of 64 <= 1024");
    if (t->wide_blocks_per_cu > 16)
        return fail(MSBFS_ERR_INVALID_ARG, "wide_blocks_per_cu must be <= 16");
    if (t->direction_rule > 1)
        return fail(MSBFS_ERR_INVALID_ARG, "direction_rule must be 0 or 1");
    if (t->stepped > 2) return fail(MSBFS_ERR_INVALID_ARG, "stepped must be 0, 1 or 2");
    if (t->deep_rows > 2) return fail(MSBFS_ERR_INVALID_ARG, "deep_rows must be 0, 1 or 2");
    if (t->deep_lattice > 2) return fail(MSBFS_ERR_INVALID_ARG, "deep_lattice must be 0, 1 or 2");
    if (t->batch_union > 2) return fail(MSBFS_ERR_INVALID_ARG, "batch_union must be 0, 1 or 2");
    if (t->union_words && (t->union_words > 64 || (t->union_words & (t->union_words - 1))))
        return fail(MSBFS_ERR_INVALID_ARG, "union_words must be a power of two <= 64");
    if (t->deep_blocks_per_cu > 32)
        return fail(MSBFS_ERR_INVALID_ARG, "deep_blocks_per_cu must be <= 32");
    return MSBFS_OK;
}

extern "C" int msbfs_engine_set_tuning(msbfs_engine *e, const msbfs_tuning *t) {
    if (!e || !t) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    int rc = msbfs_tuning_validate(t);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(e->mu);
    e->tune = *t;
    return MSBFS_OK;
}

extern "C" int msbfs_engine_get_tuning(const msbfs_engine *e, msbfs_tuning *t) {
    if (!e || !t) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    *t = e->tune;
    return MSBFS_OK;
}

// GpuBfsEngine::select_auto_method (crates/gpu_bfs/src/engine.rs:157-198).
extern "C" int msbfs_select_auto_method(uint64_t V, uint64_t E, uint32_t max_degree,
                                        uint64_t S) {
    if (V == 0 || S == 0) return MSBFS_METHOD_NAIVE_SEQUENTIAL;
    const bool road_or_narrow = max_degree <= 8 && E <= 4 * V;
    if (road_or_narrow) return MSBFS_METHOD_WORKGROUP_PERSISTENT_BFS;
    if (max_degree >= 128) return MSBFS_METHOD_INDIRECT_SPARSE_QUEUE_PUSH;
    if (S >= 64) return MSBFS_METHOD_MULTI_WORD_BIT_PARALLEL;
    if (S >= 32) return MSBFS_METHOD_BIT_PARALLEL_64;
    if (S > 1) return MSBFS_METHOD_BIT_PARALLEL_32;
    return MSBFS_METHOD_SPARSE_QUEUE;
}

namespace msbfs {

static size_t deep_lds_bytes(uint64_t V, int bits) { return ((V * bits + 31) / 32) * 4; }

// Deep family: sources per task so that the task count fills the device,
// with the task's visited bits fitting in one workgroup's LDS.
int deep_bits_for(const msbfs_context *c, const msbfs_tuning &t, uint64_t V, uint64_t S) {
    if (V >= (1ull << 24)) return 0;
    auto fits = [&](int b) { return deep_lds_bytes(V, b) + kDeepStaticLds <= c->lds_max; };
    if (t.deep_bits) return fits((int)t.deep_bits) ? (int)t.deep_bits : 0;
    // Several tasks per CU hide each other's per-level latency, so the
    // width is the largest that still gives every CU k tasks at once, for
    // k = 8, 4, then 2, with k sets of visited bits and queues fitting its LDS
    // (road-like |V| = 114,599, S = 2048, ms: k=8 3.46, k=4 3.55, k=2 3.91).
    const uint64_t cus = (uint64_t)std::max(1, c->cus);
    for (uint64_t k : {8u, 4u, 2u}) {
        for (int b : {32, 16, 8, 4, 2, 1}) {
            const size_t need = deep_lds_bytes(V, b) + kDeepStaticLds + 2 * 512 * (b <= 8 ? 4 : 8);
            if ((S + b - 1) / b >= k * cus && k * need <= c->lds_max) return b;
        }
    }
    int best = 0;
    for (int b : {1, 2, 4, 8, 16, 32}) {
        if (!fits(b)) break;
        if (best == 0 || (S + b - 1) / b >= cus) best = b;
    }
    return best;
}

// Logged rows: the tasks that finish in the last part of a launch leave
// their rows' rewrite to a separate pass over the whole device (1/kLogDefer
// of the tasks).
constexpr uint64_t kLogDefer = 4;
uint64_t log_deferred_tasks(uint64_t ntasks) { return ntasks / kLogDefer; }

// Staged rows (level codes in LDS, one source per task) pay off when most
// distances fit their 4-bit codes (below 14); a graph known to need more
// levels stores as it goes, and so does one that looks road-like before its
// first traversal: its codes would take LDS that holds more tasks at once,
// and such a traversal is bound by its levels (road-like |V| = 114,599, 2,048
// sources, first traversal: 3.5 ms direct, 9.6 ms staged at a quarter of
// the tasks per CU).
constexpr uint32_t kStagedMaxLevels = 32;

// LDS split of a deep-family workgroup: visited bits (and staged level
// codes) first, the rest as two frontier queues.  Enough workgroups per CU
// are planned to hold every task at once when LDS and threads allow.

void deep_layout_v(const msbfs_context *c, const msbfs_tuning &t, uint64_t max_v, int ell,
                   uint64_t ntasks, Plan &p, uint32_t levels_seen, bool road_like) {
    const int B = p.bits;
    p.ell = ell;
    const size_t esz = B <= 8 ? 4 : 8;
    const size_t min_q = 2 * 512 * esz;
    // each resident workgroup also holds the kernel's static LDS
    auto dyn_budget = [&](int k) { return c->lds_max / k - kDeepStaticLds; };
    const uint64_t vis_plain = (max_v * B + 31) / 32;
    const uint64_t nib_words = (max_v + 7) / 8;
    // One source per task and a frontier that fits one wave (|V| / levels,
    // from an earlier traversal of this graph): the one-wave kernel
    // (road-like |V| = 114,599, 3,080 levels, ms: S=2048 3.11 vs 3.44,
    // S=256 2.07 vs 2.21; Grid 256x256 (128 per level) S=512 0.82 vs 0.57
    // and sparse G(n,p) S=2048 11.5 vs 7.7 stay on the workgroup kernel).
    const bool wave_rule = B == 1 && ell > 0 && levels_seen > 0 && max_v <= 64ull * levels_seen;
    const bool wave = ell > 0 && (t.deep_threads ? t.deep_threads == 64 : wave_rule);
    p.staged = false;
    if (B == 1 && !wave && t.deep_rows != 2 &&
        (vis_plain + nib_words) * 4 + min_q <= dyn_budget(1))
        p.staged = t.deep_rows == 1 ||
                   (levels_seen ? levels_seen <= kStagedMaxLevels : !road_like);
    const uint64_t vis_words = vis_plain + (p.staged ? nib_words : 0);
    // Workgroup size: the 2048 threads of a CU shared by the tasks it holds
    // at once (capped by LDS), between 128 and 1024.
    if (t.deep_threads) {
        p.threads = (int)t.deep_threads;
    } else if (wave_rule) {
        p.threads = 64;
    } else {
        const uint64_t fit =
            std::max<uint64_t>(1, c->lds_max / (vis_words * 4 + min_q + kDeepStaticLds));
        const uint64_t k = std::min<uint64_t>((ntasks + c->cus - 1) / c->cus, fit);
        // (road-like |V| = 114,599, S = 2048, B = 1, eight tasks per CU, ms:
        // 64 threads 3.57, 128 3.39, 256 3.48; one task per CU: 256 best)
        if (k <= 1)
            p.threads = B >= 4 ? 1024 : 256;
        else
            p.threads = k < 4 ? 512 : (k < 8 ? 256 : 128);
    }
    p.vis_words = (uint32_t)vis_words;
    const int max_by_threads = std::max(1, 2048 / p.threads);
    int per_cu = (int)std::min<uint64_t>((ntasks + c->cus - 1) / c->cus, max_by_threads);
    per_cu = std::max(per_cu, 1);
    while (per_cu > 1 && p.vis_words * 4 + min_q > dyn_budget(per_cu)) --per_cu;
    const size_t per_wg = dyn_budget(per_cu);
    // Logged u32 output (one-wave kernel, one source per task): one LDS
    // counter per bin of 2^kDeepLogBits vertices after the queues; a
    // (vertex in bin, level) entry is below 0xFFFFFFFF (the kernel's "no
    // entry" mark) when level < 2^(32 - kDeepLogBits) - 1, which
    // |V| < 2^(32 - kDeepLogBits) guarantees (level <= |V| - 1).
    // Only with four or more tasks per CU: rewriting a task's row costs
    // about 7 us per bin when the task ends, hidden behind other tasks when
    // there are many (road-like, 114,599 vertices, ms logged / direct:
    // S=2048 (8 per CU) 2.81 / 2.95, S=256 2.14 / 1.89, S=1 1.71 / 1.48).
    const bool logged = p.threads == 64 && B == 1 && ell > 0 && max_v > 0 && t.deep_rows != 2 &&
                        max_v < (1ull << (32 - kDeepLogBits)) &&
                        ntasks >= 4 * (uint64_t)std::max(1, c->cus);
    const uint64_t nbins = logged ? (max_v + (1u << kDeepLogBits) - 1) >> kDeepLogBits : 0;
    size_t qbytes = per_wg > p.vis_words * 4 + nbins * 4 ? per_wg - p.vis_words * 4 - nbins * 4 : 0;
    uint64_t qcap = qbytes / (2 * esz);
    qcap = std::min<uint64_t>(qcap, std::max<uint64_t>((uint64_t)B * max_v, 1));
    qcap = std::min<uint64_t>(qcap, 1u << 20);
    p.qcap = (uint32_t)qcap;
    p.lds = p.vis_words * 4 + 2 * qcap * esz;
    // the task's end rewrites each bin's span through LDS: the visited bits
    // and queues must hold one bin
    p.log_bw = 0;
    if (logged && std::min<uint64_t>(max_v, 1u << kDeepLogBits) * 4 <= p.lds) {
        p.log_bw = kDeepLogBits;
        p.bin_off = (uint32_t)(p.lds / 4);
        p.lds += nbins * 4;
    }
    // the one-wave kernel's 64 sink entries come out of the static reservation
    p.sink_off = (uint32_t)((p.lds + 7) / 8 * 2);
    if (p.threads == 64 && ell > 0) p.lds = p.sink_off * 4 + 64 * esz;
}

// Lattice kernel: the adjacency words first (shared by the workgroup's
// waves), then one region per wave laid out as the one-wave kernel's LDS:
// visited bits, two queues, the log's bin counters (logged u32 output),
// the 64-entry append sink.  As many waves per workgroup (tasks per CU) as
// the tasks fill, up to kLatticeMaxWaves, each wave's queues holding at
// least 64 entries (a level that outgrows them spills to global memory).
bool lattice_layout(const msbfs_context *c, const msbfs_graph *g, uint64_t ntasks, bool logged_ok,
                    Plan &p) {
    if (!g->lat || g->lat_bits == 0) return false;
    const uint64_t V = g->V;
    const uint64_t adj = (g->lat_words + 3ull) / 4 * 4;  // 16-byte aligned regions
    if (adj * 4 >= c->lds_max) return false;
    const uint64_t budget = c->lds_max / 4 - adj;  // words for the regions
    const uint64_t vis = (V + 31) / 32;
    const uint64_t cus = (uint64_t)std::max(1, c->cus);
    const uint64_t want = std::min<uint64_t>(kLatticeMaxWaves, std::max<uint64_t>(1, (ntasks + cus - 1) / cus));
    // logged rows as for the one-wave kernel (four or more tasks per CU); a
    // task's end rewrites each bin through its visited bits and queues
    const bool logged = logged_ok && V < (1ull << (32 - kDeepLogBits)) && ntasks >= 4 * cus;
    for (int lg = logged ? 1 : 0; lg >= 0; --lg) {
        const uint64_t nbins = lg ? (V + (1u << kDeepLogBits) - 1) >> kDeepLogBits : 0;
        for (uint64_t w = want; w >= 1; --w) {
            const uint64_t region = budget / w / 4 * 4;
            const uint64_t fixed = vis + nbins + 64;
            if (region < fixed + 2 * 64) continue;
            uint64_t qcap = std::min<uint64_t>((region - fixed) / 2, std::max<uint64_t>(V, 1));
            if (lg && vis + 2 * qcap < std::min<uint64_t>(V, 1u << kDeepLogBits)) continue;
            p.lat = g->lat_bits;
            p.lat_waves = (uint32_t)w;
            p.lat_words = (uint32_t)adj;
            p.threads = (int)(64 * w);
            p.ell = 2 * g->lat_bits;
            p.vis_words = (uint32_t)vis;
            p.qcap = (uint32_t)qcap;
            p.staged = false;
            p.log_bw = lg ? kDeepLogBits : 0;
            p.bin_off = (uint32_t)(vis + 2 * qcap);
            p.sink_off = (uint32_t)(p.bin_off + nbins);
            p.lat_region = (uint32_t)((p.sink_off + 64 + 3) / 4 * 4);
            p.lds = (size_t)(adj + w * p.lat_region) * 4;
            return true;
        }
    }
    return false;
}

int ensure_zeroed(DevBuf &b, size_t bytes, bool force, hipStream_t st) {
    int r = b.ensure(bytes);
    if (r < 0) return -r;
    if (r == 1 || force) HIP_TRY(hipMemsetAsync(b.p, 0, b.bytes, st));
    return MSBFS_OK;
}

int ensure_plain(DevBuf &b, size_t bytes) {
    int r = b.ensure(bytes);
    return r < 0 ? -r : MSBFS_OK;
}

}  // namespace msbfs

namespace {

// Output element kind for the kernels (0 u32, 1 u16, 2 u8) and bytes per element.
uint32_t dist_kind(int dtype) {
    return dtype == MSBFS_DIST_U16 ? 1u : dtype == MSBFS_DIST_U8 ? 2u : 0u;
}
size_t dist_bytes(int dtype) {
    return dtype == MSBFS_DIST_U16 ? 2 : dtype == MSBFS_DIST_U8 ? 1 : 4;
}


// lattice: whether the deep family may take the lattice kernel (the
// no-allocation entry point keeps to the one-wave and workgroup kernels).
Plan plan_for(const msbfs_engine *e, const msbfs_graph *g, uint64_t S, bool lattice = true) {
    Plan p;
    p.method = e->method;
    if (p.method == MSBFS_METHOD_AUTO) {
        const int ref = msbfs_select_auto_method(g->V, g->E, g->max_degree, S);
        p.method = ref;
        if (ref == MSBFS_METHOD_WORKGROUP_PERSISTENT_BFS) {
            p.family = MSBFS_FAMILY_DEEP;
        } else {
            p.family = MSBFS_FAMILY_WIDE;
            p.policy = MSBFS_DIR_ADAPTIVE;
        }
    } else if (p.method == MSBFS_METHOD_WORKGROUP_PERSISTENT_BFS ||
               p.method == MSBFS_METHOD_NAIVE_SEQUENTIAL) {
        // NaiveSequential (dispatch/naive_seq.rs: one single-source search
        // after another) maps to the deep family with one source per task:
        // independent single-source searches, run side by side.
        p.family = MSBFS_FAMILY_DEEP;
    } else if (p.method == MSBFS_METHOD_BATCHED_DENSE_SPMM) {
        // BatchedDenseSpMM (dispatch/batched_spmm.rs: every vertex's parents
        // combined each level, all sources at once) maps to the wide
        // family's pull, which is that product with early exit.
        p.family = MSBFS_FAMILY_WIDE;
        p.policy = MSBFS_DIR_PULL;
    } else {
        p.family = MSBFS_FAMILY_WIDE;
        p.policy = is_pull_only(p.method)   ? MSBFS_DIR_PULL
                   : is_push_only(p.method) ? MSBFS_DIR_PUSH
                                            : MSBFS_DIR_ADAPTIVE;
    }
    if (p.family == MSBFS_FAMILY_DEEP) {
        msbfs_tuning t = e->tune;
        if (p.method == MSBFS_METHOD_NAIVE_SEQUENTIAL) t.deep_bits = 1;
        p.bits = deep_bits_for(e->ctx, t, g->V, S);
        if (p.bits == 0) {
            // The task's visited bits do not fit in LDS: the device-wide
            // family takes over (the reference likewise falls back to its
            // multi-word pull path, workgroup_persistent.rs:494-512).
            p.family = MSBFS_FAMILY_WIDE;
            p.policy = MSBFS_DIR_ADAPTIVE;
        } else {
            deep_layout_v(e->ctx, t, g->V, g->ell_width, (S + p.bits - 1) / p.bits, p,
                          g->deep_levels.load(std::memory_order_relaxed),
                          g->max_degree <= 8 && g->E <= 4 * g->V);
            // The lattice kernel takes the place the one-wave kernel would
            // take by default (or, when asked, of any one-source-per-task
            // run): same tasks, the adjacency read from LDS instead of the
            // neighbour table.
            if (lattice && g->lat && p.bits == 1 &&
                (t.deep_lattice == 1 || (t.deep_lattice == 0 && t.deep_threads == 0 &&
                                         p.threads == 64 && g->ell_width > 0))) {
                Plan q = p;
                // direct stores by default: with no global load in the
                // level chain, scattered stores do not hold it up, and the
                // binned log's appends and rewrite cost more (road-like
                // |V| = 114,599, 2,048 sources: 2.03 ms direct, 2.26 logged)
                if (lattice_layout(e->ctx, g, S, t.deep_rows == 1, q)) p = q;
            }
        }
    }
    return p;
}

// Launch form of a wide traversal.  Stepped launches (one per phase of each
// level) cost a few microseconds per level in launch gaps but no grid
// barriers, and are faster on low-diameter graphs (R-MAT scale 20, ms
// stepped / cooperative: 2,048 sources 5.07 / 5.28, 64 sources 1.29 /
// 1.30); the cooperative kernel wins when levels are many.
constexpr uint32_t kSteppedMaxLevels = 64;

bool wide_stepped(const msbfs_tuning &t, const msbfs_graph *g) {
    if (t.stepped == 1) return true;
    if (t.stepped == 2) return false;
    const uint32_t known = std::max(g->wide_levels.load(std::memory_order_relaxed),
                                    g->deep_levels.load(std::memory_order_relaxed));
    if (known) return known <= kSteppedMaxLevels;
    return !(g->max_degree <= 8 && g->E <= 4 * g->V);  // the reference's road-like test
}

int check_sources(const msbfs_graph *g, const uint32_t *sources, uint64_t S) {
    // GpuBfsEngine::run_ms_bfs_into argument checks (engine.rs:228-249).
    if (g->V == 0)
        return fail(MSBFS_ERR_SOURCE_OOB, "Source vertex %u at index 0 exceeds graph vertex count 0",
                    sources[0]);
    for (uint64_t i = 0; i < S; ++i)
        if ((uint64_t)sources[i] >= g->V)
            return fail(MSBFS_ERR_SOURCE_OOB,
                        "Source vertex %u at index %llu exceeds graph vertex count %llu",
                        sources[i], (unsigned long long)i, (unsigned long long)g->V);
    return MSBFS_OK;
}

// Enqueues one traversal.  `out` (device) may be null when only the
// reachability words are wanted.
// un: a batch union traversal (seed pairs and output map; wide family,
// no reachability), else null.
int enqueue_run(msbfs_engine *e, const msbfs_graph *g, const uint32_t *sources, uint64_t S,
                int dtype, void *out, uint32_t *reach, uint32_t reach_wpv, hipStream_t st,
                const Plan &plan, const UnionSeeds *un = nullptr) {
    msbfs_context *c = e->ctx;
    int rc;
    const uint64_t V = g->V;
    if (un && (plan.family != MSBFS_FAMILY_WIDE || reach))
        return fail(MSBFS_ERR_INVALID_ARG, "union traversal: wide family and distances only");
    // sources (or seed vertices and bits): host -> pinned -> device
    const uint64_t nsrc_words = un ? 2 * un->nseeds : S;
    if ((rc = e->h_sources.ensure(nsrc_words * 4)) < 0) return -rc;
    if (un) {
        std::memcpy(e->h_sources.p, un->seed_v, un->nseeds * 4);
        std::memcpy((uint32_t *)e->h_sources.p + un->nseeds, un->seed_b, un->nseeds * 4);
    } else {
        std::memcpy(e->h_sources.p, sources, S * 4);
    }
    if ((rc = ensure_plain(e->sources, nsrc_words * 4))) return rc;
    // the flags, then the deep launch's task counters (cleared together)
    if ((rc = ensure_plain(e->flags, (FlagCount + 4) * 4))) return rc;
    if ((rc = e->h_flags.ensure(FlagCount * 4)) < 0) return -rc;
    // the cooperative wide kernel needs the device to itself; stepped wide
    // launches and deep launches share it
    const bool stepped = plan.family == MSBFS_FAMILY_WIDE && wide_stepped(e->tune, g);
    const bool coop = plan.family == MSBFS_FAMILY_WIDE && !stepped;
    LaunchOrder order(c, st, coop);
    if ((rc = order.begin())) return rc;
    HIP_TRY(hipEventRecord(e->ev[0], st));
    HIP_TRY(hipMemcpyAsync(e->sources.p, e->h_sources.p, nsrc_words * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(launch_zero_u32(e->flags.as<uint32_t>(), FlagCount + 4, st));

    e->last = msbfs_run_stats{};
    e->last.method = plan.method;
    e->last.family = plan.family;

    if (plan.family == MSBFS_FAMILY_WIDE) {
        if (!c->coop) return fail(MSBFS_ERR_UNSUPPORTED, "device lacks cooperative launch");
        const uint64_t W = (S + 63) / 64;
        const uint64_t words = W * V;
        const bool force = e->dirty;
        if ((rc = ensure_zeroed(e->seen, words * 8, force, st))) return rc;
        // Frontier slots: with an output, the first `defer` distances keep
        // their words for the final pass (up to 15: the final pass encodes a
        // distance in a nibble with 15 for unreached; within a quarter of the
        // free device memory and 16 GiB), the rest alternate in two slots.
        uint32_t defer = 0;
        if (out) {
            size_t free_b = 0, total_b = 0;
            HIP_TRY(hipMemGetInfo(&free_b, &total_b));
            const uint64_t budget = std::min<uint64_t>(free_b / 4 + e->fr.bytes / 4, 16ull << 30);
            const uint64_t fit = budget / std::max<uint64_t>(words * 8, 1);
            defer = fit >= 4 ? (uint32_t)std::min<uint64_t>(fit - 2, 15) : 0;
        }
        if ((rc = ensure_zeroed(e->fr, (defer + 2) * words * 8, force, st))) return rc;
        if ((rc = ensure_zeroed(e->nx, words * 8, force, st))) return rc;
        if ((rc = ensure_zeroed(e->fbA, (V / 32 + 1) * 4, force, st))) return rc;
        if ((rc = ensure_zeroed(e->fbB, (V / 32 + 1) * 4, force, st))) return rc;
        if ((rc = ensure_zeroed(e->lv, (V + 1) * 4, force, st))) return rc;
        if ((rc = ensure_plain(e->qf0, V * 4))) return rc;
        if ((rc = ensure_plain(e->qf1, V * 4))) return rc;
        if ((rc = ensure_plain(e->qc, V * 4))) return rc;
        const size_t stamp_before = e->stamp.bytes;
        if ((rc = ensure_zeroed(e->stamp, V * 4, force, st))) return rc;
        if (e->stamp.bytes != stamp_before || force) e->tick_base = 1;
        if ((uint64_t)e->tick_base + V + 4 >= 0xFFFFFFFFull) {
            HIP_TRY(hipMemsetAsync(e->stamp.p, 0, e->stamp.bytes, st));
            e->tick_base = 1;
        }
        // level counters (4 x 32 B), then the grid barrier words from byte 256
        static_assert(4 * sizeof(WideCounters) <= 256, "level counters overlap the barrier");
        // ... and the stepped launches' level control after the barrier words
        constexpr size_t kCtlOff = 256 + kGridBarrierWords * 4;
        static_assert(kCtlOff % 16 == 0 && sizeof(WideCtl) % 4 == 0, "level control layout");
        if ((rc = ensure_plain(e->ctr, kCtlOff + sizeof(WideCtl)))) return rc;
        // zero row (one word per source word), read in place of frontier
        // words of parents outside the frontier; never written
        if ((rc = ensure_zeroed(e->zrow, std::max<uint64_t>(W, 64) * 8, false, st))) return rc;
        if ((rc = ensure_plain(e->act, 3 * W * 8))) return rc;  // set up by the seed kernel
        e->dirty = false;

        WideParams p{};
        p.V = (uint32_t)V;
        p.W = (uint32_t)W;
        p.S = S;
        p.E = g->E;
        p.ro = g->ro;
        p.ci = g->ci;
        p.co = g->co;
        p.ri = g->ri;
        p.seen = e->seen.as<unsigned long long>();
        p.fr = e->fr.as<unsigned long long>();
        p.defer = defer;
        p.nx = e->nx.as<unsigned long long>();
        p.qf0 = e->qf0.as<uint32_t>();
        p.qf1 = e->qf1.as<uint32_t>();
        p.qc = e->qc.as<uint32_t>();
        p.stamp = e->stamp.as<uint32_t>();
        p.sources = e->sources.as<uint32_t>();
        p.ctr = e->ctr.as<WideCounters>();
        p.flags = e->flags.as<uint32_t>();
        p.out = out;
        p.reach = reach;
        p.reach_wpv = reach_wpv;
        p.tick_base = e->tick_base;
        p.out_kind = dist_kind(dtype);
        p.policy = (uint32_t)plan.policy;
        p.alpha = e->tune.alpha ? e->tune.alpha : kDefaultAlpha;
        p.beta = e->tune.beta ? e->tune.beta : kDefaultBeta;
        p.rule = e->tune.direction_rule;
        p.heavy_t = kHeavyDegree;
        p.heavy_out = g->heavy_out;
        p.heavy_in = g->heavy_in;
        p.n_heavy_out = g->n_heavy_out;
        p.n_heavy_in = g->n_heavy_in;
        p.fbA = e->fbA.as<uint32_t>();
        p.fbB = e->fbB.as<uint32_t>();
        p.lv = e->lv.as<uint32_t>();
        p.hchunk_out = (const uint4 *)g->hchunk_out;
        p.hchunk_in = (const uint4 *)g->hchunk_in;
        p.n_hchunk_out = g->n_hchunk_out;
        p.n_hchunk_in = g->n_hchunk_in;
        p.light_in = (const uint4 *)g->light_in;
        p.n_light_in = g->n_light_in;
        if (un) {
            p.seed_v = e->sources.as<uint32_t>();
            p.seed_b = e->sources.as<uint32_t>() + un->nseeds;
            p.nseeds = (uint32_t)un->nseeds;
            p.bmap = un->bmap;
            p.vgraph = un->vgraph;
        }
        e->tick_base += (uint32_t)(V + 2);

        // Tile width: lanes per vertex.  With an even word count of 16 or
        // more the deferred pull levels give each lane two adjacent words by
        // default, so a wave holds twice the vertices (each with one scan of
        // its in-list); an explicit tile width keeps one word per lane
        // unless two are asked for.
        auto tile_for = [&](uint64_t k) {
            int t = 1;
            while (t < 64 && (uint64_t)t * k < W) t *= 2;
            return t;
        };
        const uint32_t wpl = e->tune.words_per_lane;
        p.pair = W % 2 == 0 && (wpl == 2 || (wpl == 0 && !e->tune.tile_words && W >= 16));
        int tw = e->tune.tile_words ? (int)e->tune.tile_words : tile_for(p.pair ? 2 : 1);
        if (p.pair && tw < 8) {  // two words per lane need tiles of 8 or more lanes
            p.pair = 0;
            if (!e->tune.tile_words) tw = tile_for(1);
        }
        // light vertices by decreasing in-degree once a vertex's words fill
        // whole cache lines (16 words: 128 bytes), else in vertex order
        if (W >= 16) p.light_in = (const uint4 *)g->light_in_deg;
        const uint32_t tv = 64 / tw;
        uint32_t nb = 64;
        if (W > 64) {
            nb = 1;
            while (nb * 2 <= 64 && (size_t)nb * 2 * (W + 1) * 8 <= 32768) nb *= 2;
        }
        if (nb < tv) nb = tv;
        p.pull_nb = nb;
        p.zrow = e->zrow.as<unsigned long long>();
        p.act = e->act.as<unsigned long long>();
        p.gbar = (uint32_t *)((char *)e->ctr.p + 256);
        const size_t lds = wide_pull_lds(nb, (uint32_t)W);
        if (e->tune.trace_levels) {
            e->trace_cap = e->tune.trace_levels;
            if ((rc = ensure_zeroed(e->trace, (size_t)(e->trace_cap + 1) * 4 * 8, true, st))) return rc;
            p.trace = e->trace.as<unsigned long long>();
            p.trace_cap = e->trace_cap;
        }
        e->trace_valid = e->tune.trace_levels != 0;
        int occ = 0;
        HIP_TRY(wide_max_blocks_per_cu(tw, kWideThreads, lds, &occ));
        if (occ < 1) return fail(MSBFS_ERR_HIP, "wide kernel cannot be resident");
        int per_cu = e->tune.wide_blocks_per_cu ? (int)e->tune.wide_blocks_per_cu : 2;
        per_cu = std::min(per_cu, occ);
        const int blocks = c->cus * per_cu;

        HIP_TRY(launch_zero_u32((uint32_t *)e->ctr.p, (kCtlOff + sizeof(WideCtl)) / 4, st));
        HIP_TRY(launch_wide_seed(p, st));
        HIP_TRY(hipEventRecord(e->ev[1], st));
        if (stepped) {
            // as many levels as this graph has needed before, and a margin;
            // finish_run enqueues more if the frontier is not empty by then
            p.ctl = (WideCtl *)((char *)e->ctr.p + kCtlOff);
            const uint32_t seen = g->wide_levels.load(std::memory_order_relaxed);
            const uint32_t nlev = seen ? seen + 2 : 16;
            HIP_TRY(launch_wide_steps(tw, p, blocks, kWideThreads, lds, nlev, st));
            e->step_p = p;
            e->step_tw = tw;
            e->step_blocks = blocks;
            e->step_lds = lds;
            e->step_levels = nlev;
        } else {
            HIP_TRY(launch_wide(tw, p, blocks, kWideThreads, lds, st));
        }
        {
            // final pass: whole tiles per workgroup, up to four workgroups per CU
            const uint64_t ntile = (V + kFinalVerts - 1) / kFinalVerts;
            // (R-MAT scale 20, 2,048 sources, ms per traversal by workgroups
            // per CU: 1 6.33, 2 6.00, 4 5.99, 8 5.97, 16 5.93)
            const int fblocks = (int)std::max<uint64_t>(1, std::min<uint64_t>(ntile, (uint64_t)c->cus * 16));
            HIP_TRY(launch_wide_final(p, fblocks, st));
        }
        HIP_TRY(hipEventRecord(e->ev[2], st));
        e->last.grid_blocks = (uint32_t)blocks;
        e->last.block_threads = (uint32_t)kWideThreads;
        e->last.workspace_bytes = e->seen.bytes * 2 + e->fr.bytes + e->stamp.bytes * 4;
    } else {
        const int B = plan.bits;
        const uint64_t ntasks = (S + B - 1) / B;
        const int threads = plan.threads;
        // the lattice kernel runs lat_waves tasks per workgroup, each wave
        // with its own spill queues and log sink
        const uint64_t waves = plan.lat ? plan.lat_waves : 1;
        {
            const hipError_t pe = plan.lat ? deep_lattice_prepare(plan.lat, (int)waves, plan.lds)
                                           : deep_prepare(B, plan.ell, threads, plan.lds);
            if (pe != hipSuccess)
                return fail(pe == hipErrorOutOfMemory ? MSBFS_ERR_OOM : MSBFS_ERR_HIP,
                            "deep kernel set-up (%d sources per task, %d threads, %zu bytes of "
                            "LDS): %s", B, plan.threads, plan.lds, hipGetErrorString(pe));
        }
        int occ = 0;
        if (plan.lat)
            HIP_TRY(deep_lattice_blocks_per_cu(plan.lat, (int)waves, plan.lds, &occ));
        else
            HIP_TRY(deep_max_blocks_per_cu(B, plan.ell, threads, plan.lds, &occ));
        if (occ < 1) return fail(MSBFS_ERR_HIP, "deep kernel cannot be resident (lds %zu)", plan.lds);
        if (e->tune.deep_blocks_per_cu) occ = std::min(occ, (int)e->tune.deep_blocks_per_cu);
        const uint64_t resident = (uint64_t)c->cus * occ;
        const int blocks = (int)std::min<uint64_t>((ntasks + waves - 1) / waves, resident);
        const uint64_t cap = std::max<uint64_t>((uint64_t)B * V, (uint64_t)B);
        const uint64_t spill = cap > plan.qcap ? cap - plan.qcap : 1;
        const size_t esz = B <= 8 ? 4 : 8;
        if ((rc = ensure_plain(e->queues, (size_t)blocks * waves * 2 * spill * esz))) return rc;
        if ((rc = ensure_plain(e->tasks, ntasks * sizeof(DeepTask)))) return rc;
        if ((rc = ensure_plain(e->graphs, sizeof(DeepGraph)))) return rc;
        if ((rc = e->h_tasks.ensure(ntasks * sizeof(DeepTask) + sizeof(DeepGraph))) < 0) return -rc;
        // The task table and the graph record depend only on these values;
        // when they match the last upload, the device copies are current.
        const uint64_t key[10] = {S, (uint64_t)B, V, ntasks, (uint64_t)(uintptr_t)g->ro,
                                  (uint64_t)(uintptr_t)g->ci, (uint64_t)(uintptr_t)g->ell,
                                  (uint64_t)plan.ell, (uint64_t)(uintptr_t)e->tasks.p,
                                  (uint64_t)(uintptr_t)e->graphs.p};
        const bool upload = std::memcmp(key, e->tasks_key, sizeof(key)) != 0;
        if (upload) {
        // The pinned staging buffer may still be read by a previous enqueue.
        HIP_TRY(hipStreamSynchronize(st));
        auto *ht = (DeepTask *)e->h_tasks.p;
        for (uint64_t t = 0; t < ntasks; ++t) {
            ht[t].graph = 0;
            ht[t].src_off = t * B;
            ht[t].nsrc = (uint32_t)std::min<uint64_t>(B, S - t * B);
            ht[t].out_off = t * B * V;
            ht[t].reach_bit = t * B;
        }
        auto *hg = (DeepGraph *)((char *)e->h_tasks.p + ntasks * sizeof(DeepTask));
        hg->ro = g->ro;
        hg->ci = g->ci;
        hg->ell = g->ell;
        hg->V = (uint32_t)V;
        hg->ell_width = (uint32_t)plan.ell;
        hg->lat = g->lat;
        hg->lat_pad = g->lat_pad;
        std::memcpy(hg->lat_d, g->lat_d, sizeof(hg->lat_d));
        HIP_TRY(hipMemcpyAsync(e->tasks.p, ht, ntasks * sizeof(DeepTask), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(e->graphs.p, hg, sizeof(DeepGraph), hipMemcpyHostToDevice, st));
        std::memcpy(e->tasks_key, key, sizeof(key));
        }

        DeepParams p{};
        p.tasks = e->tasks.as<DeepTask>();
        p.graphs = e->graphs.as<DeepGraph>();
        p.sources = e->sources.as<uint32_t>();
        p.out = out;
        p.queues = e->queues.p;
        p.cap = spill + plan.qcap;
        p.lds_qcap = plan.qcap;
        p.lds_vis_words = plan.vis_words;
        // rows written through the binned log: u32 output only
        p.log_bw = (out && dtype == MSBFS_DIST_U32) ? plan.log_bw : 0;
        p.lds_bin_off = plan.bin_off;
        p.lds_sink_off = plan.sink_off;
        p.staged = (plan.staged && (dtype == MSBFS_DIST_U32 || dtype == MSBFS_DIST_U16)) ? 1u : 0u;
        p.log_defer_after = (uint32_t)ntasks;
        uint32_t log_nbins = 0;
        p.lat_words = plan.lat_words;
        p.lat_region = plan.lat_region;
        p.lat_waves = (uint32_t)waves;
        if (p.log_bw) {
            if ((rc = ensure_plain(e->log_sink, (size_t)blocks * waves * 64 * 4))) return rc;
            p.log_sink = e->log_sink.as<uint32_t>();
            log_nbins = (uint32_t)((V + (1u << p.log_bw) - 1) >> p.log_bw);
            const uint64_t nlist = log_deferred_tasks(ntasks);
            if ((rc = ensure_plain(e->log_defer, std::max<size_t>(1, nlist * (1 + log_nbins) * 4))))
                return rc;
            p.log_defer = e->log_defer.as<uint32_t>();
            p.log_defer_after = (uint32_t)(ntasks - nlist);
        }
        p.work = e->flags.as<uint32_t>() + FlagCount;  // (cleared with the flags)
        p.flags = e->flags.as<uint32_t>();
        // reachability: each task writes its sources' rows of a source-major
        // bitmap, transposed into the vertex-major words afterwards
        const uint32_t vw_n = (uint32_t)((V + 31) / 32);
        if (reach) {
            if ((rc = ensure_plain(e->reach_t, (size_t)S * vw_n * 4))) return rc;
            p.reach_t = e->reach_t.as<uint32_t>();
            p.reach_vw = vw_n;
        }
        p.ntasks = (uint32_t)ntasks;
        p.out_kind = dist_kind(dtype);
        HIP_TRY(hipEventRecord(e->ev[1], st));
        if (plan.lat)
            HIP_TRY(launch_deep_lattice(plan.lat, p, blocks, plan.lds, st));
        else
            HIP_TRY(launch_deep(B, plan.ell, p, blocks, threads, plan.lds, st));
        if (p.log_bw) HIP_TRY(launch_log_rewrite(p, log_nbins, st));
        if (reach)
            HIP_TRY(launch_reach_transpose(e->reach_t.as<uint32_t>(), vw_n, S, (uint32_t)V, reach,
                                           reach_wpv, st));
        HIP_TRY(hipEventRecord(e->ev[2], st));
        e->last.instances = (uint32_t)ntasks;
        e->last.bits_per_instance = (uint32_t)B;
        e->last.grid_blocks = (uint32_t)blocks;
        e->last.block_threads = (uint32_t)threads;
        e->last.lattice_bits = (uint32_t)plan.lat;
        e->last.workspace_bytes = e->queues.bytes;
    }
    HIP_TRY(hipMemcpyAsync(e->h_flags.p, e->flags.p, FlagCount * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(e->ev[3], st));
    if ((rc = order.end())) return rc;
    e->pending = true;
    e->last_stream = st;
    e->pending_dtype = dtype;
    e->pending_stepped = stepped;
    if (plan.family == MSBFS_FAMILY_WIDE || (plan.bits == 1 && g->ell_width > 0)) {
        msbfs_graph *const gm = const_cast<msbfs_graph *>(g);
        gm->refs.fetch_add(1);
        e->pending_graph = gm;
    }
    return MSBFS_OK;
}

// Waits for the pending run and turns its device flags into a status.
int finish_run(msbfs_engine *e) {
    if (!e->pending) return MSBFS_OK;
    e->pending = false;
    msbfs_graph *const pg = e->pending_graph;
    e->pending_graph = nullptr;
    hipError_t er = hipStreamSynchronize(e->last_stream);
    if (er != hipSuccess) {
        graph_release(pg);
        e->dirty = true;
        return hip_fail(er, "traversal");
    }
    const uint32_t *f = (const uint32_t *)e->h_flags.p;
    // A stepped traversal with a frontier left after its enqueued levels:
    // twice as many more levels (plus the final pass) until it ends.
    while (e->pending_stepped && f[FlagIncomplete] && !f[FlagAbort]) {
        hipStream_t st = e->last_stream;
        uint32_t *const dflags = e->flags.as<uint32_t>();
        int rc = MSBFS_OK;
        e->step_levels = std::min<uint32_t>(2 * e->step_levels, e->step_p.V + 2);
        if (launch_zero_u32(dflags + FlagIncomplete, 1, st) != hipSuccess ||
            launch_wide_steps(e->step_tw, e->step_p, e->step_blocks, kWideThreads, e->step_lds,
                              e->step_levels, st) != hipSuccess ||
            launch_wide_final(e->step_p, (int)std::max<uint64_t>(1, std::min<uint64_t>(
                                  (e->step_p.V + kFinalVerts - 1) / kFinalVerts,
                                  (uint64_t)e->ctx->cus * 16)), st) != hipSuccess ||
            hipEventRecord(e->ev[2], st) != hipSuccess ||
            hipMemcpyAsync(e->h_flags.p, e->flags.p, FlagCount * 4, hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            hipEventRecord(e->ev[3], st) != hipSuccess)
            rc = fail(MSBFS_ERR_HIP, "stepped traversal: continuing launches failed");
        if (rc == MSBFS_OK && (er = hipStreamSynchronize(st)) != hipSuccess)
            rc = hip_fail(er, "traversal");
        if (rc) {
            graph_release(pg);
            e->dirty = true;
            return rc;
        }
    }
    if (pg) {
        // record the level count for later plans on this graph
        if (!f[FlagQueueOverflow] && !f[FlagAbort]) {
            std::atomic<uint32_t> &lv = e->last.family == MSBFS_FAMILY_WIDE ? pg->wide_levels
                                                                            : pg->deep_levels;
            uint32_t cur = lv.load(std::memory_order_relaxed);
            while (f[FlagLevels] > cur &&
                   !lv.compare_exchange_weak(cur, f[FlagLevels], std::memory_order_relaxed)) {
            }
        }
        graph_release(pg);
    }
    e->last.levels = f[FlagLevels];
    e->last.push_levels = f[FlagPushLevels];
    e->last.pull_levels = f[FlagPullLevels];
    float ms = 0;
    if (hipEventElapsedTime(&ms, e->ev[1], e->ev[2]) == hipSuccess) e->last.kernel_ms = ms;
    if (hipEventElapsedTime(&ms, e->ev[0], e->ev[3]) == hipSuccess) e->last.total_ms = ms;
    if (f[FlagQueueOverflow]) {
        e->dirty = true;
        return fail(MSBFS_ERR_HIP, "internal frontier queue overflow");
    }
    if (f[FlagAbort]) {
        e->dirty = true;
        return fail(MSBFS_ERR_HIP, "grid barrier wait timed out: the traversal was abandoned");
    }
    if (f[FlagU16Overflow] && e->pending_dtype == MSBFS_DIST_U16)
        return fail(MSBFS_ERR_DIAMETER_U16,
                    "a distance >= 65535 cannot be stored as u16 (0xFFFF is the unreachable "
                    "sentinel); use u32 output");
    if (f[FlagU16Overflow] && e->pending_dtype == MSBFS_DIST_U8)
        return fail(MSBFS_ERR_DIAMETER_U8,
                    "a distance >= 255 cannot be stored as u8 (0xFF is the unreachable "
                    "sentinel); use u16 or u32 output");
    return MSBFS_OK;
}

// Widens n wire elements (type `from`) into n host elements (type `to`),
// mapping the sentinel to the sentinel, on up to `threads` host threads.
template <typename F, typename T>
void widen_span(const F *src, T *dst, uint64_t n, int threads) {
    constexpr F fs = (F)~(F)0;
    constexpr T ts = (T)~(T)0;
    auto body = [&](uint64_t b, uint64_t e) {
        if constexpr (sizeof(F) == sizeof(T)) {
            std::memcpy(dst + b, src + b, (e - b) * sizeof(T));
        } else {
            for (uint64_t i = b; i < e; ++i) dst[i] = src[i] == fs ? ts : (T)src[i];
        }
    };
    if (threads <= 1 || n < (1u << 20)) {
        body(0, n);
        return;
    }
    std::vector<std::thread> th;
    const uint64_t per = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const uint64_t b = t * per, e = std::min(n, b + per);
        if (b < e) th.emplace_back(body, b, e);
    }
    for (auto &x : th) x.join();
}

void widen_any(const void *src, int from, void *dst, int to, uint64_t n, int threads) {
    if (from == MSBFS_DIST_U8) {
        if (to == MSBFS_DIST_U8) widen_span((const uint8_t *)src, (uint8_t *)dst, n, threads);
        else if (to == MSBFS_DIST_U16) widen_span((const uint8_t *)src, (uint16_t *)dst, n, threads);
        else widen_span((const uint8_t *)src, (uint32_t *)dst, n, threads);
    } else if (from == MSBFS_DIST_U16) {
        if (to == MSBFS_DIST_U16) widen_span((const uint16_t *)src, (uint16_t *)dst, n, threads);
        else widen_span((const uint16_t *)src, (uint32_t *)dst, n, threads);
    } else {
        widen_span((const uint32_t *)src, (uint32_t *)dst, n, threads);
    }
}

// Distances packed `bits` (4 or 12) each in groups of 8 (launch_pack_u16_bits)
// -> T, all ones -> T's sentinel, on up to `threads` host threads.
template <typename T>
void unpack_bits(const uint8_t *src, int bits, T *dst, uint64_t n, int threads) {
    constexpr T ts = (T)~(T)0;
    const uint32_t sent = (1u << bits) - 1u;
    auto body = [&](uint64_t g0, uint64_t g1) {
        for (uint64_t g = g0; g < g1; ++g) {
            uint32_t v[8];
            if (bits == 4) {
                uint32_t r;
                std::memcpy(&r, src + 4 * g, 4);
                for (int k = 0; k < 8; ++k) v[k] = (r >> (4 * k)) & 15u;
            } else {
                uint32_t w[3];
                std::memcpy(w, src + 12 * g, 12);
                v[0] = w[0] & 0xFFFu;
                v[1] = (w[0] >> 12) & 0xFFFu;
                v[2] = (w[0] >> 24) | ((w[1] & 0xFu) << 8);
                v[3] = (w[1] >> 4) & 0xFFFu;
                v[4] = (w[1] >> 16) & 0xFFFu;
                v[5] = (w[1] >> 28) | ((w[2] & 0xFFu) << 4);
                v[6] = (w[2] >> 8) & 0xFFFu;
                v[7] = w[2] >> 20;
            }
            const uint64_t b = 8 * g, m = std::min<uint64_t>(8, n - b);
            for (uint64_t k = 0; k < m; ++k) dst[b + k] = v[k] == sent ? ts : (T)v[k];
        }
    };
    const uint64_t ng = (n + 7) / 8;
    if (threads <= 1 || n < (1u << 20)) {
        body(0, ng);
        return;
    }
    std::vector<std::thread> th;
    const uint64_t per = (ng + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const uint64_t b = t * per, e = std::min(ng, b + per);
        if (b < e) th.emplace_back(body, b, e);
    }
    for (auto &x : th) x.join();
}

// Device -> caller's host buffer through three pinned staging chunks: the
// copy of chunk k+1.. runs on `st` while host threads widen chunk k.  The
// wire holds `bits` per distance: 32, 16 or 8 (plain u32 / u16 / u8) or 4
// or 12 (launch_pack_u16_bits groups of 8).
int transfer_widen(msbfs_engine *e, const void *src, int bits, void *dst, int to, uint64_t n,
                   hipStream_t st) {
    constexpr int NS = 3;
    const bool packed = bits == 4 || bits == 12;
    const size_t tb = dist_bytes(to);
    // elements per chunk: 32 MiB of wire data, whole groups of 8
    const uint64_t chunk = (32ull << 20) * 8 / bits / 8 * 8;
    const uint64_t chunk_bytes = chunk / 8 * bits;
    auto wire_bytes = [&](uint64_t cnt) { return (cnt + 7) / 8 * bits; };
    int rc = e->h_stage.ensure(NS * chunk_bytes);
    if (rc < 0) return -rc;
    for (auto &ev : e->stage_ev)
        if (!ev) HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    const uint64_t nch = (n + chunk - 1) / chunk;
    const int threads = std::min(16, host_threads());
    auto issue = [&](uint64_t k) -> int {
        const uint64_t off = k * chunk, cnt = std::min(chunk, n - off);
        HIP_TRY(hipMemcpyAsync((char *)e->h_stage.p + (k % NS) * chunk_bytes,
                               (const char *)src + off / 8 * bits,
                               packed ? wire_bytes(cnt) : cnt * bits / 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipEventRecord(e->stage_ev[k % NS], st));
        return MSBFS_OK;
    };
    for (uint64_t k = 0; k < std::min<uint64_t>(NS, nch); ++k)
        if ((rc = issue(k))) return rc;
    for (uint64_t k = 0; k < nch; ++k) {
        HIP_TRY(hipEventSynchronize(e->stage_ev[k % NS]));
        const uint64_t off = k * chunk, cnt = std::min(chunk, n - off);
        const uint8_t *w = (const uint8_t *)e->h_stage.p + (k % NS) * chunk_bytes;
        char *d = (char *)dst + off * tb;
        if (!packed)
            widen_any(w, bits == 32 ? MSBFS_DIST_U32 : bits == 16 ? MSBFS_DIST_U16 : MSBFS_DIST_U8, d,
                      to, cnt, threads);
        else if (to == MSBFS_DIST_U32)
            unpack_bits(w, bits, (uint32_t *)d, cnt, threads);
        else if (to == MSBFS_DIST_U16)
            unpack_bits(w, bits, (uint16_t *)d, cnt, threads);
        else
            unpack_bits(w, bits, (uint8_t *)d, cnt, threads);
        if (k + NS < nch && (rc = issue(k + NS))) return rc;
    }
    return MSBFS_OK;
}

int run_common(msbfs_engine *e, const msbfs_graph *g, const uint32_t *sources, uint64_t S,
               int dtype, void *out_dev, uint32_t *reach_dev, uint32_t reach_wpv, hipStream_t st,
               bool sync) {
    const Plan plan = plan_for(e, g, S);
    int rc = finish_run(e);
    if (rc) return rc;
    rc = enqueue_run(e, g, sources, S, dtype, out_dev, reach_dev, reach_wpv, st, plan);
    if (rc) {
        e->dirty = true;
        e->pending = false;
        return rc;
    }
    return sync ? finish_run(e) : MSBFS_OK;
}

}  // namespace

int msbfs::engine_run_union(msbfs_engine *e, const msbfs_graph *g, const UnionSeeds &u, uint64_t S,
                            int dtype, void *out, hipStream_t st) {
    if (!e || !g || S == 0 || u.nseeds == 0) return fail(MSBFS_ERR_INVALID_ARG, "empty union run");
    std::lock_guard<std::mutex> lk(e->mu);
    const Plan plan = plan_for(e, g, S);
    int rc = finish_run(e);
    if (rc) return rc;
    rc = enqueue_run(e, g, nullptr, S, dtype, out, nullptr, 0, st, plan, &u);
    if (rc) {
        e->dirty = true;
        e->pending = false;
    }
    return rc;
}

extern "C" int msbfs_engine_resolve(const msbfs_engine *e, const msbfs_graph *g, uint64_t S,
                                    int *method_out) {
    if (!e || !g) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    const Plan p = plan_for(e, g, S);
    if (method_out) *method_out = p.method;
    return p.family;
}

extern "C" int msbfs_run_ms_bfs_device(msbfs_engine *e, const msbfs_graph *g,
                                       const uint32_t *sources, uint64_t S, int dtype,
                                       void *out_device, uint64_t out_len, void *stream,
                                       int sync) {
    clear_error();
    if (!e || !g) return fail(MSBFS_ERR_INVALID_ARG, "null engine or graph");
    if (dtype != MSBFS_DIST_U32 && dtype != MSBFS_DIST_U16 && dtype != MSBFS_DIST_U8)
        return fail(MSBFS_ERR_INVALID_ARG, "unknown dist type %d", dtype);
    if (S == 0) return MSBFS_OK;
    if (!sources) return fail(MSBFS_ERR_INVALID_ARG, "null sources");
    std::lock_guard<std::mutex> lk(e->mu);
    int rc = check_sources(g, sources, S);
    if (rc) return rc;
    if (out_len < S * g->V)
        return fail(MSBFS_ERR_OUT_TOO_SMALL, "output holds %llu elements, %llu needed",
                    (unsigned long long)out_len, (unsigned long long)(S * g->V));
    if (!out_device) return fail(MSBFS_ERR_INVALID_ARG, "null output");
    HIP_TRY(hipSetDevice(e->ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : e->ctx->stream;
    return run_common(e, g, sources, S, dtype, out_device, nullptr, 0, st, sync != 0);
}

extern "C" int msbfs_engine_check(msbfs_engine *e) {
    if (!e) return fail(MSBFS_ERR_INVALID_ARG, "null engine");
    std::lock_guard<std::mutex> lk(e->mu);
    return finish_run(e);
}

extern "C" int msbfs_engine_level_trace(msbfs_engine *e, msbfs_level_record *out,
                                        uint32_t max_levels, uint32_t *n_levels) {
    if (!e || !n_levels) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    int rc = finish_run(e);
    if (rc) return rc;
    *n_levels = 0;
    if (!e->trace_valid || e->last.family != MSBFS_FAMILY_WIDE) return MSBFS_OK;
    const uint32_t L = std::min(e->last.levels, e->trace_cap);
    std::vector<unsigned long long> raw((size_t)(e->trace_cap + 1) * 4);
    HIP_TRY(hipMemcpy(raw.data(), e->trace.p, raw.size() * 8, hipMemcpyDeviceToHost));
    int khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, e->ctx->device));
    const double us_per_tick = khz > 0 ? 1000.0 / khz : 0.0;
    const unsigned long long t0 = raw[0];
    const uint32_t n = std::min(L, max_levels);
    for (uint32_t l = 0; l < n && out; ++l) {
        msbfs_level_record &r = out[l];
        r.level = l;
        r.direction = (uint32_t)(raw[4 * l + 1] >> 32);
        r.frontier = (uint32_t)raw[4 * l + 1];
        r.reserved = 0;
        r.new_bits = raw[4 * l + 3];
        r.start_us = (double)(raw[4 * l] - t0) * us_per_tick;
        r.mid_us = (double)(raw[4 * l + 2] - t0) * us_per_tick;
        r.end_us = (double)(raw[4 * (l + 1)] - t0) * us_per_tick;
    }
    *n_levels = n;
    return MSBFS_OK;
}

extern "C" int msbfs_engine_last_stats(msbfs_engine *e, msbfs_run_stats *st) {
    if (!e || !st) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    int rc = finish_run(e);
    *st = e->last;
    return rc;
}

extern "C" int msbfs_run_ms_bfs(msbfs_engine *e, const msbfs_graph *g, const uint32_t *sources,
                                uint64_t S, int dtype, void *out, uint64_t out_len) {
    clear_error();
    if (!e || !g) return fail(MSBFS_ERR_INVALID_ARG, "null engine or graph");
    if (dtype != MSBFS_DIST_U32 && dtype != MSBFS_DIST_U16 && dtype != MSBFS_DIST_U8)
        return fail(MSBFS_ERR_INVALID_ARG, "unknown dist type %d", dtype);
    if (S == 0) return MSBFS_OK;
    if (!sources) return fail(MSBFS_ERR_INVALID_ARG, "null sources");
    std::lock_guard<std::mutex> lk(e->mu);
    int rc = check_sources(g, sources, S);
    if (rc) return rc;
    const uint64_t n = S * g->V;
    if (out_len < n)
        return fail(MSBFS_ERR_OUT_TOO_SMALL, "output holds %llu elements, %llu needed",
                    (unsigned long long)out_len, (unsigned long long)n);
    if (!out) return fail(MSBFS_ERR_INVALID_ARG, "null output");
    HIP_TRY(hipSetDevice(e->ctx->device));
    hipStream_t st = e->ctx->stream;
    // Device run in the most compact type that surely holds the result.
    const int dev = dtype == MSBFS_DIST_U8 ? MSBFS_DIST_U8 : MSBFS_DIST_U16;
    if ((rc = ensure_plain(e->out_tmp, n * dist_bytes(dev)))) return rc;
    rc = run_common(e, g, sources, S, dev, e->out_tmp.p, nullptr, 0, st, true);
    if (rc == MSBFS_ERR_DIAMETER_U16 && dtype == MSBFS_DIST_U32) {
        // a distance of 65,535 or more: the run is repeated with u32 on the device
        clear_error();
        if ((rc = ensure_plain(e->out_tmp, n * 4))) return rc;
        rc = run_common(e, g, sources, S, MSBFS_DIST_U32, e->out_tmp.p, nullptr, 0, st, true);
        if (rc) return rc;
        return transfer_widen(e, e->out_tmp.p, 32, out, dtype, n, st);
    }
    if (rc) return rc;
    if (dev == MSBFS_DIST_U16) {
        // the narrowest wire that holds every distance below its all-ones
        // sentinel (levels = largest distance + 1): 4, 8 or 12 bits, else u16
        const uint32_t L = e->last.levels;
        const int bits = L <= 15 ? 4 : L <= 255 ? 8 : L <= 4095 ? 12 : 16;
        if (bits == 8) {
            if ((rc = ensure_plain(e->out_pack, n))) return rc;
            HIP_TRY(launch_pack_u16_u8(e->out_tmp.as<uint16_t>(), e->out_pack.as<uint8_t>(), n, st));
            return transfer_widen(e, e->out_pack.p, 8, out, dtype, n, st);
        }
        if (bits != 16) {
            if ((rc = ensure_plain(e->out_pack, (n + 7) / 8 * bits))) return rc;
            HIP_TRY(launch_pack_u16_bits(e->out_tmp.as<uint16_t>(), e->out_pack.as<uint32_t>(), n, bits, st));
            return transfer_widen(e, e->out_pack.p, bits, out, dtype, n, st);
        }
    }
    return transfer_widen(e, e->out_tmp.p, (int)dist_bytes(dev) * 8, out, dtype, n, st);
}

extern "C" int msbfs_run_ms_reachability(msbfs_engine *e, const msbfs_graph *g,
                                         const uint32_t *sources, uint64_t S, uint32_t *words,
                                         uint64_t words_len) {
    clear_error();
    if (!e || !g) return fail(MSBFS_ERR_INVALID_ARG, "null engine or graph");
    // run_ms_reachability: S == 0 or |V| == 0 -> all-zero bitmask (engine.rs:428-430).
    if (S == 0 || g->V == 0) return MSBFS_OK;
    if (!sources) return fail(MSBFS_ERR_INVALID_ARG, "null sources");
    std::lock_guard<std::mutex> lk(e->mu);
    int rc = check_sources(g, sources, S);
    if (rc) return rc;
    const uint64_t wpv = (S + 31) / 32;
    const uint64_t n = g->V * wpv;
    if (words_len < n)
        return fail(MSBFS_ERR_OUT_TOO_SMALL, "bitmask holds %llu words, %llu needed",
                    (unsigned long long)words_len, (unsigned long long)n);
    if (!words) return fail(MSBFS_ERR_INVALID_ARG, "null output");
    HIP_TRY(hipSetDevice(e->ctx->device));
    if ((rc = ensure_plain(e->reach_tmp, n * 4))) return rc;
    hipStream_t st = e->ctx->stream;
    HIP_TRY(launch_zero_u32(e->reach_tmp.as<uint32_t>(), n, st));
    rc = run_common(e, g, sources, S, MSBFS_DIST_U32, nullptr, e->reach_tmp.as<uint32_t>(),
                    (uint32_t)wpv, st, true);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(words, e->reach_tmp.p, n * 4, hipMemcpyDeviceToHost));
    return MSBFS_OK;
}

// ---------------------------------------------------------------- autotuner

namespace {

// Mean kernel time of two timed runs after one untimed run, or a negative
// value when the configuration cannot run.
float time_config(msbfs_engine *e, const msbfs_graph *g, const uint32_t *sources, uint64_t S,
                  int dtype, void *out, hipStream_t st) {
    float total = 0;
    for (int it = 0; it < 3; ++it) {
        if (run_common(e, g, sources, S, dtype, out, nullptr, 0, st, true) != MSBFS_OK)
            return -1.0f;
        if (it) total += e->last.kernel_ms;
    }
    return total / 2;
}

}  // namespace

extern "C" int msbfs_engine_autotune(msbfs_engine *e, const msbfs_graph *g,
                                     const uint32_t *sources, uint64_t S, int dtype,
                                     void *out_device, uint64_t out_len, msbfs_tuning *best,
                                     int *best_method, float *best_ms) {
    clear_error();
    if (!e || !g) return fail(MSBFS_ERR_INVALID_ARG, "null engine or graph");
    if (dtype != MSBFS_DIST_U32 && dtype != MSBFS_DIST_U16 && dtype != MSBFS_DIST_U8)
        return fail(MSBFS_ERR_INVALID_ARG, "unknown dist type %d", dtype);
    if (S == 0 || !sources) return fail(MSBFS_ERR_INVALID_ARG, "autotune needs sources");
    std::lock_guard<std::mutex> lk(e->mu);
    int rc = check_sources(g, sources, S);
    if (rc) return rc;
    const uint64_t n = S * g->V;
    HIP_TRY(hipSetDevice(e->ctx->device));
    void *out = out_device;
    if (out) {
        if (out_len < n)
            return fail(MSBFS_ERR_OUT_TOO_SMALL, "output holds %llu elements, %llu needed",
                        (unsigned long long)out_len, (unsigned long long)n);
    } else {
        if ((rc = ensure_plain(e->out_tmp, n * dist_bytes(dtype)))) return rc;
        out = e->out_tmp.p;
    }
    if ((rc = finish_run(e))) return rc;
    hipStream_t st = e->ctx->stream;
    const int method0 = e->method;
    const msbfs_tuning base = e->tune;
    msbfs_tuning win_t = base;
    int win_m = method0;
    float win = 1e30f;
    auto consider = [&](int method, const msbfs_tuning &t) {
        e->method = method;
        e->tune = t;
        const float ms = time_config(e, g, sources, S, dtype, out, st);
        if (ms >= 0 && ms < win) {
            win = ms;
            win_t = t;
            win_m = method;
        }
    };
    const bool try_deep = method0 == MSBFS_METHOD_AUTO ||
                          method0 == MSBFS_METHOD_WORKGROUP_PERSISTENT_BFS;
    const bool try_wide = method0 != MSBFS_METHOD_WORKGROUP_PERSISTENT_BFS;
    if (try_deep) {
        const int dm = MSBFS_METHOD_WORKGROUP_PERSISTENT_BFS;
        for (uint32_t bits : {1u, 2u, 4u, 8u, 16u, 32u}) {
            msbfs_tuning t = base;
            t.deep_bits = bits;
            if (deep_bits_for(e->ctx, t, g->V, S) != (int)bits) continue;  // does not fit LDS
            if (bits > 1 && (S + bits / 2 - 1) / (bits / 2) <= 1) break;   // wider than S
            for (uint32_t threads : {64u, 128u, 256u, 512u, 1024u}) {
                t.deep_threads = threads;
                consider(dm, t);
            }
        }
    }
    if (try_wide) {
        const int wm = method0 == MSBFS_METHOD_AUTO ? MSBFS_METHOD_PUSH_PULL_ADAPTIVE : method0;
        // one knob at a time from the defaults, keeping each knob's best
        msbfs_tuning cur = base;
        cur.tile_words = 0;
        cur.words_per_lane = 0;
        cur.wide_blocks_per_cu = 0;
        cur.direction_rule = 0;
        const float w0 = [&] {
            e->method = wm;
            e->tune = cur;
            return time_config(e, g, sources, S, dtype, out, st);
        }();
        float cur_ms = w0 < 0 ? 1e30f : w0;
        if (w0 >= 0 && w0 < win) {
            win = w0;
            win_t = cur;
            win_m = wm;
        }
        auto knob = [&](auto set, const std::vector<uint32_t> &vals) {
            msbfs_tuning keep = cur;
            for (uint32_t v : vals) {
                msbfs_tuning t = cur;
                set(t, v);
                e->method = wm;
                e->tune = t;
                const float ms = time_config(e, g, sources, S, dtype, out, st);
                if (ms >= 0 && ms < cur_ms) {
                    cur_ms = ms;
                    keep = t;
                }
                if (ms >= 0 && ms < win) {
                    win = ms;
                    win_t = t;
                    win_m = wm;
                }
            }
            cur = keep;
        };
        const uint64_t W = (S + 63) / 64;
        uint32_t tw = 1;
        while (tw < 64 && tw < W) tw *= 2;
        std::vector<uint32_t> tws;
        if (tw > 1) tws.push_back(tw / 2);
        if (tw < 64) tws.push_back(tw * 2);
        knob([](msbfs_tuning &t, uint32_t v) { t.words_per_lane = v; }, {1u});
        knob([](msbfs_tuning &t, uint32_t v) { t.tile_words = v; }, tws);
        knob([](msbfs_tuning &t, uint32_t v) { t.wide_blocks_per_cu = v; }, {1u});
        knob([](msbfs_tuning &t, uint32_t v) { t.direction_rule = v; }, {1u});
        knob([](msbfs_tuning &t, uint32_t v) { t.alpha = v; }, {7u, 28u});
        knob([](msbfs_tuning &t, uint32_t v) { t.beta = v; }, {12u, 48u});
    }
    e->method = win_m;
    e->tune = win_t;
    e->tune.trace_levels = base.trace_levels;
    if (win >= 1e30f) {
        e->method = method0;
        e->tune = base;
        return fail(MSBFS_ERR_HIP, "autotune: no candidate ran");
    }
    clear_error();
    if (best) *best = e->tune;
    if (best_method) *best_method = win_m;
    if (best_ms) *best_ms = win;
    return MSBFS_OK;
}

// ------------------------------------------------------ no-allocation variant
//
// msbfs_run_ms_bfs_device_noalloc: the deep family with every workspace in a
// static per-device array, for callers that cannot allocate (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
// Layout of the array (each part 256-byte aligned): the task counter and
// flags, the graph record, the task table, the sources, the one-wave
// kernel's per-workgroup sink words, and the frontier queues' spill space
// (the rest).  Calls on one device share its array and leave their contents
// in it; the counter and flags are cleared by a fill kernel at every call, the
// rest is written before it is read.  Concurrent calls on one device race.

// The array's name is a host-side stand-in: kernels get the per-device
// address that hipGetSymbolAddress returns (msbfs_context::scratch).
static __device__ __align__(256) unsigned char scratch[MSBFS_NOALLOC_SCRATCH_BYTES];

namespace {
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

int msbfs::noalloc_setup(msbfs_context *c) {
    // with c->device current (msbfs_context_create)
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e == hipSuccess ? hipErrorInvalidSymbol : e,
                                               "no-alloc scratch address");
    c->scratch = p;
    return MSBFS_OK;
}

extern "C" int msbfs_run_ms_bfs_device_noalloc(msbfs_engine *e, const msbfs_graph *g,
                                               const uint32_t *sources, uint64_t S, int dtype,
                                               void *out_device, uint64_t out_len, void *stream) {
    clear_error();
    if (!e || !g) return fail(MSBFS_ERR_INVALID_ARG, "null engine or graph");
    if (dtype != MSBFS_DIST_U32 && dtype != MSBFS_DIST_U16 && dtype != MSBFS_DIST_U8)
        return fail(MSBFS_ERR_INVALID_ARG, "unknown dist type %d", dtype);
    if (S == 0) return MSBFS_OK;
    if (!sources) return fail(MSBFS_ERR_INVALID_ARG, "null sources");
    std::lock_guard<std::mutex> lk(e->mu);
    int rc = check_sources(g, sources, S);
    if (rc) return rc;
    if (out_len < S * g->V)
        return fail(MSBFS_ERR_OUT_TOO_SMALL, "output holds %llu elements, %llu needed",
                    (unsigned long long)out_len, (unsigned long long)(S * g->V));
    if (!out_device) return fail(MSBFS_ERR_INVALID_ARG, "null output");
    msbfs_context *c = e->ctx;
    if (!c->scratch) return fail(MSBFS_ERR_HIP, "no-alloc scratch was not set up");
    if (S > MSBFS_NOALLOC_MAX_SOURCES)
        return fail(MSBFS_ERR_OUT_TOO_SMALL,
                    "no-alloc call: %llu sources, at most %u fit its scratch",
                    (unsigned long long)S, (unsigned)MSBFS_NOALLOC_MAX_SOURCES);
    const Plan plan = plan_for(e, g, S, false);
    if (plan.family != MSBFS_FAMILY_DEEP)
        return fail(MSBFS_ERR_UNSUPPORTED,
                    "no-alloc call: this graph and source count route to the wide family, whose "
                    "workspace (source words per vertex) does not fit a fixed scratch; use "
                    "msbfs_run_ms_bfs_device");
    HIP_TRY(hipSetDevice(c->device));
    if ((rc = finish_run(e))) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;

    const int B = plan.bits;
    const uint64_t V = g->V, ntasks = (S + B - 1) / B;
    const int threads = plan.threads;
    {
        const hipError_t pe = deep_prepare(B, plan.ell, threads, plan.lds);
        if (pe != hipSuccess)
            return fail(MSBFS_ERR_HIP, "deep kernel set-up: %s", hipGetErrorString(pe));
    }
    int occ = 0;
    HIP_TRY(deep_max_blocks_per_cu(B, plan.ell, threads, plan.lds, &occ));
    if (occ < 1) return fail(MSBFS_ERR_HIP, "deep kernel cannot be resident");
    const int blocks = (int)std::min<uint64_t>(ntasks, (uint64_t)c->cus * occ);
    const size_t esz = B <= 8 ? 4 : 8;
    // scratch layout
    char *const base = (char *)c->scratch;
    const size_t off_graph = 256;
    const size_t off_tasks = align256(off_graph + sizeof(DeepGraph));
    const size_t off_src = align256(off_tasks + ntasks * sizeof(DeepTask));
    const size_t off_sink = align256(off_src + S * 4);
    // logged rows: the deferred tasks' list and bin counts
    const bool logged = dtype == MSBFS_DIST_U32 && plan.log_bw;
    const uint64_t log_nbins = logged ? (V + (1u << plan.log_bw) - 1) >> plan.log_bw : 0;
    const uint64_t log_nlist = logged ? log_deferred_tasks(ntasks) : 0;
    const size_t off_defer = align256(off_sink + (size_t)blocks * 64 * 4);
    const size_t off_queues = align256(off_defer + log_nlist * (1 + log_nbins) * 4);
    if (off_queues + (size_t)blocks * 2 * esz > MSBFS_NOALLOC_SCRATCH_BYTES)
        return fail(MSBFS_ERR_OUT_TOO_SMALL, "no-alloc call: the task table does not fit its scratch");
    // spill entries per queue: the worst case (every vertex in one level)
    // when it fits, else what the scratch holds; a larger frontier is an error
    const uint64_t worst = (uint64_t)B * V > plan.qcap ? (uint64_t)B * V - plan.qcap : 1;
    const uint64_t room = (MSBFS_NOALLOC_SCRATCH_BYTES - off_queues) / ((size_t)blocks * 2 * esz);
    const uint64_t spill = std::max<uint64_t>(1, std::min(worst, room));

    // task table and graph record through pageable memory (synchronous copies)
    std::vector<DeepTask> ht(ntasks);
    for (uint64_t t = 0; t < ntasks; ++t) {
        ht[t].graph = 0;
        ht[t].src_off = t * B;
        ht[t].nsrc = (uint32_t)std::min<uint64_t>(B, S - t * B);
        ht[t].out_off = t * B * V;
        ht[t].reach_bit = t * B;
    }
    DeepGraph hg{};
    hg.ro = g->ro;
    hg.ci = g->ci;
    hg.ell = g->ell;
    hg.V = (uint32_t)V;
    hg.ell_width = (uint32_t)plan.ell;
    LaunchOrder order(c, st, false);
    if ((rc = order.begin())) return rc;
    HIP_TRY(hipMemcpyAsync(base + off_tasks, ht.data(), ntasks * sizeof(DeepTask),
                           hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(base + off_graph, &hg, sizeof(hg), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(base + off_src, sources, S * 4, hipMemcpyHostToDevice, st));
    uint32_t *const work = (uint32_t *)base, *const flags = work + 8;
    HIP_TRY(launch_zero_u32(work, 8 + FlagCount, st));

    DeepParams p{};
    p.tasks = (const DeepTask *)(base + off_tasks);
    p.graphs = (const DeepGraph *)(base + off_graph);
    p.sources = (const uint32_t *)(base + off_src);
    p.out = out_device;
    p.queues = base + off_queues;
    p.cap = spill + plan.qcap;
    p.lds_qcap = plan.qcap;
    p.lds_vis_words = plan.vis_words;
    p.work = work;
    p.flags = flags;
    p.ntasks = (uint32_t)ntasks;
    p.out_kind = dtype == MSBFS_DIST_U16 ? 1u : dtype == MSBFS_DIST_U8 ? 2u : 0u;
    p.log_bw = (dtype == MSBFS_DIST_U32) ? plan.log_bw : 0;
    p.lds_bin_off = plan.bin_off;
    p.lds_sink_off = plan.sink_off;
    p.log_sink = (uint32_t *)(base + off_sink);
    p.log_defer = (uint32_t *)(base + off_defer);
    p.log_defer_after = (uint32_t)(ntasks - log_nlist);
    p.staged = (plan.staged && (dtype == MSBFS_DIST_U32 || dtype == MSBFS_DIST_U16)) ? 1u : 0u;
    HIP_TRY(hipEventRecord(e->ev[1], st));
    HIP_TRY(launch_deep(B, plan.ell, p, blocks, threads, plan.lds, st));
    if (p.log_bw) HIP_TRY(launch_log_rewrite(p, (uint32_t)log_nbins, st));
    HIP_TRY(hipEventRecord(e->ev[2], st));
    if ((rc = order.end())) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    uint32_t f[FlagCount];
    HIP_TRY(hipMemcpy(f, flags, sizeof(f), hipMemcpyDeviceToHost));
    e->last = msbfs_run_stats{};
    e->last.method = plan.method;
    e->last.family = plan.family;
    e->last.levels = f[FlagLevels];
    e->last.instances = (uint32_t)ntasks;
    e->last.bits_per_instance = (uint32_t)B;
    e->last.grid_blocks = (uint32_t)blocks;
    e->last.block_threads = (uint32_t)threads;
    float ms = 0;
    if (hipEventElapsedTime(&ms, e->ev[1], e->ev[2]) == hipSuccess) e->last.kernel_ms = ms;
    if (f[FlagQueueOverflow])
        return fail(MSBFS_ERR_OUT_TOO_SMALL,
                    "no-alloc call: a frontier level outgrew the scratch's queue space "
                    "(%llu entries per task); use msbfs_run_ms_bfs_device",
                    (unsigned long long)(spill + plan.qcap));
    if (f[FlagU16Overflow] && dtype == MSBFS_DIST_U16)
        return fail(MSBFS_ERR_DIAMETER_U16,
                    "a distance >= 65535 cannot be stored as u16 (0xFFFF is the unreachable "
                    "sentinel); use u32 output");
    if (f[FlagU16Overflow] && dtype == MSBFS_DIST_U8)
        return fail(MSBFS_ERR_DIAMETER_U8,
                    "a distance >= 255 cannot be stored as u8 (0xFF is the unreachable "
                    "sentinel); use u16 or u32 output");
    return MSBFS_OK;
}
