// Batched multi-graph launch: many resident graphs, each with its own
// (uneven) source set, traversed by one persistent deep-family launch.
//
// Every graph's sources are cut into tasks of up to B sources; all tasks of
// all graphs go into one table, largest graphs first, and the persistent
// workgroups pull them from a global counter.  One workgroup runs one task
// at a time with the task's visited bits in LDS, so graphs of any size mix
// freely inside the launch and no device-wide barrier is involved.  A graph
// whose visited bits do not fit one workgroup's LDS goes to the wide family
// instead: one engine (workspace) per such graph, run on the same stream
// after the persistent launch.  A run can be recorded into a hipGraph once
// and replayed (the wide runs stay eager: their level count is only known on
// the device); an autotuner times the candidate plans on the batch itself.
//
// The reference has no batched path (one engine call per graph and source
// set, crates/gpu_bfs/src/engine.rs:219-297).
#include <algorithm>
#include <numeric>

#include "engine_internal.hpp"

using namespace msbfs;

struct msbfs_batch {
    msbfs_engine *eng = nullptr;
    int dtype = MSBFS_DIST_U32;
    uint32_t n = 0;
    std::vector<const msbfs_graph *> graphs;
    std::vector<uint64_t> S, src_off, out_off;
    uint64_t total_sources = 0, total_out = 0, max_v = 0;
    int ell = 0;
    msbfs_tuning tune{};
    Plan plan;
    uint64_t ntasks = 0;
    int blocks = 0;
    hipStream_t stream = nullptr;
    DevBuf d_graphs, d_tasks, d_sources, queues, work, flags;
    PinnedBuf h_flags, h_stage;
    std::vector<uint32_t> h_sources;
    // ev[0], ev[1]: a run's start and end on the caller's stream; ev[2]:
    // the fork to the unions' streams (unions run beside the persistent
    // launch and one another: their output rows are disjoint); ev[3] unused
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *graph_out = nullptr;
    msbfs_run_stats last{};
    bool pending = false;
    bool planned = false;
    std::vector<char> is_wide;               // per graph: traversed by the wide family
    std::vector<uint32_t> wide;              // those graphs
    std::vector<msbfs_engine *> wide_eng;    // one engine per wide graph
    std::vector<msbfs_sweep_entry> sweep;    // the last autotune's candidates
    // Union traversals (batch_union): graphs grouped by source words per
    // vertex, one disjoint union per group, traversed by the wide family.
    struct UnionClass {
        uint32_t W = 0;                      // source words per vertex
        std::vector<uint32_t> gs;            // the batch's graphs in this union
        msbfs_graph *ug = nullptr;           // the union graph
        msbfs_engine *eng = nullptr;         // its workspace
        DevBuf bmap, vgraph;                 // per graph, per union vertex
        std::vector<uint32_t> seed_v, seed_b;
        uint64_t smax = 0;                   // most sources of one graph
        hipStream_t stream = nullptr;        // its own stream (unions overlap one another)
        hipEvent_t done = nullptr;           // its join event
    };
    std::vector<UnionClass> uclass;
    std::vector<char> in_union;   // per graph: traversed in a union (this plan)
    std::vector<char> built_for;  // the membership the unions were built for
    bool union_built = false;
    bool union_mode = false;      // some graph is in a union
    bool all_deep = false;        // batch_union 2: no unions, big graphs on wide_eng
    uint64_t deep_max_v = 0;      // largest graph of the deep launch (this plan)
};

namespace {

// A graph runs in the persistent launch when one source per task fits:
// its visited bits, two minimal queues and the kernel's static LDS.
bool deep_fits(const msbfs_context *c, uint64_t V) {
    return ((V + 31) / 32) * 4 + 2 * 512 * 4 + kDeepStaticLds <= c->lds_max;
}

size_t dist_bytes(int dtype) { return dtype == MSBFS_DIST_U16 ? 2 : 4; }

void drop_graph(msbfs_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_out = nullptr;
}

// Source words per vertex of a graph's union: the next power of two of
// ceil(S / 64).
uint32_t union_words_of(uint64_t S) {
    uint32_t w = 1;
    while ((uint64_t)w * 64 < S) w *= 2;
    return w;
}

// Seeds of the union traversals from the batch's current sources: graph g's
// i-th source is union vertex base_g + source with source bit i.
void union_seeds(msbfs_batch *b) {
    for (auto &u : b->uclass) {
        u.seed_v.clear();
        u.seed_b.clear();
        uint64_t base = 0;
        for (uint32_t g : u.gs) {
            for (uint64_t i = 0; i < b->S[g]; ++i) {
                u.seed_v.push_back((uint32_t)(base + b->h_sources[b->src_off[g] + i]));
                u.seed_b.push_back((uint32_t)i);
            }
            base += b->graphs[g]->V;
        }
    }
}

// Builds the union graphs once: graphs with sources grouped by source words
// per vertex (the next power of two of ceil(S / 64)), each group's CSRs
// joined with their vertex ids shifted by the graph's base, every union
// vertex carrying its graph's source count (the pull's bound) and index.
void union_release(msbfs_batch *b);

// (for the graphs of b->in_union; rebuilt when that membership changes)
int union_build(msbfs_batch *b) {
    if (b->union_built && b->built_for == b->in_union) return MSBFS_OK;
    union_release(b);
    msbfs_context *c = b->eng->ctx;
    std::vector<uint32_t> words_of(b->n, 0);
    std::vector<uint32_t> ws;
    for (uint32_t g = 0; g < b->n; ++g) {
        if (!b->in_union[g]) continue;
        const uint32_t w = union_words_of(b->S[g]);
        words_of[g] = w;
        if (std::find(ws.begin(), ws.end(), w) == ws.end()) ws.push_back(w);
    }
    std::sort(ws.begin(), ws.end());
    std::vector<msbfs_batch::UnionClass> cls(ws.size());
    int rc = MSBFS_OK;
    for (size_t k = 0; k < ws.size() && rc == MSBFS_OK; ++k) {
        auto &u = cls[k];
        u.W = ws[k];
        uint64_t Vu = 0, Eu = 0;
        for (uint32_t g = 0; g < b->n; ++g)
            if (words_of[g] == u.W) {
                u.gs.push_back(g);
                Vu += b->graphs[g]->V;
                Eu += b->graphs[g]->E;
                u.smax = std::max<uint64_t>(u.smax, b->S[g]);
            }
        if (Vu > 0xFFFFFFFEull || Eu > 0xFFFFFFFFull) {
            rc = fail(MSBFS_ERR_UNSUPPORTED, "batch union of %llu vertices and %llu edges exceeds u32 ids",
                      (unsigned long long)Vu, (unsigned long long)Eu);
            break;
        }
        std::vector<uint32_t> ro(Vu + 1), ci(Eu), vbound(Vu), vg(Vu);
        std::vector<BatchGraphMap> map(u.gs.size());
        uint64_t vb = 0, eb = 0;
        for (size_t i = 0; i < u.gs.size() && rc == MSBFS_OK; ++i) {
            const msbfs_graph *G = b->graphs[u.gs[i]];
            hipError_t er = hipMemcpy(ro.data() + vb, G->ro, (G->V + 1) * 4, hipMemcpyDeviceToHost);
            if (er == hipSuccess && G->E)
                er = hipMemcpy(ci.data() + eb, G->ci, G->E * 4, hipMemcpyDeviceToHost);
            if (er != hipSuccess) {
                rc = hip_fail(er, "batch union: graph download");
                break;
            }
            for (uint64_t v = 0; v <= G->V; ++v) ro[vb + v] += (uint32_t)eb;
            for (uint64_t j = 0; j < G->E; ++j) ci[eb + j] += (uint32_t)vb;
            for (uint64_t v = 0; v < G->V; ++v) {
                vbound[vb + v] = (uint32_t)b->S[u.gs[i]];
                vg[vb + v] = (uint32_t)i;
            }
            map[i].base = (uint32_t)vb;
            map[i].V = (uint32_t)G->V;
            map[i].S = (uint32_t)b->S[u.gs[i]];
            map[i].out_off = b->out_off[u.gs[i]];
            vb += G->V;
            eb += G->E;
        }
        if (rc) break;
        if ((rc = graph_create_impl(c, Vu, Eu, ro.data(), ci.data(), vbound.data(), &u.ug))) break;
        if ((rc = ensure_plain(u.bmap, map.size() * sizeof(BatchGraphMap))) ||
            (rc = ensure_plain(u.vgraph, Vu * 4)))
            break;
        hipError_t er = hipMemcpy(u.bmap.p, map.data(), map.size() * sizeof(BatchGraphMap),
                                  hipMemcpyHostToDevice);
        if (er == hipSuccess) er = hipMemcpy(u.vgraph.p, vg.data(), Vu * 4, hipMemcpyHostToDevice);
        if (er != hipSuccess) {
            rc = hip_fail(er, "batch union: map upload");
            break;
        }
        if ((rc = msbfs_engine_create(c, MSBFS_METHOD_PUSH_PULL_ADAPTIVE, &u.eng))) break;
        er = hipStreamCreateWithFlags(&u.stream, hipStreamNonBlocking);
        if (er == hipSuccess) er = hipEventCreateWithFlags(&u.done, hipEventDisableTiming);
        if (er != hipSuccess) rc = hip_fail(er, "batch union stream");
    }
    b->uclass = std::move(cls);
    if (rc) {
        union_release(b);
        return rc;
    }
    union_seeds(b);
    b->union_built = true;
    b->built_for = b->in_union;
    return MSBFS_OK;
}

void union_release(msbfs_batch *b) {
    for (auto &u : b->uclass) {
        if (u.eng) msbfs_engine_destroy(u.eng);  // (finishes its run)
        if (u.ug) graph_release(u.ug);
        if (u.stream) {
            (void)hipStreamSynchronize(u.stream);
            (void)hipStreamDestroy(u.stream);
        }
        if (u.done) (void)hipEventDestroy(u.done);
        u.eng = nullptr;
        u.ug = nullptr;
        u.stream = nullptr;
        u.done = nullptr;
    }
    b->uclass.clear();
    b->union_built = false;
}

// Sources per task for the batch: the widest B whose visited bits for the
// largest graph fit in LDS while the task count still fills the device.
bool in_deep(const msbfs_batch *b, uint32_t g) { return !b->is_wide[g] && !b->in_union[g]; }

int batch_bits(const msbfs_batch *b) {
    const msbfs_context *c = b->eng->ctx;
    auto fits = [&](int B) { return ((b->deep_max_v * B + 31) / 32) * 4 + kDeepStaticLds <= c->lds_max; };
    if (b->tune.deep_bits) return fits((int)b->tune.deep_bits) ? (int)b->tune.deep_bits : 0;
    int best = 0;
    for (int B : {1, 2, 4, 8, 16, 32}) {
        if (!fits(B)) break;
        uint64_t tasks = 0;
        for (uint32_t g = 0; g < b->n; ++g)
            if (in_deep(b, g)) tasks += (b->S[g] + B - 1) / B;
        if (best == 0 || tasks >= 2 * (uint64_t)c->cus) best = B;
        if (B == 8) break;  // wider tasks only when the autotuner asks for them
    }
    return best;
}

// Builds the task table and device buffers for the current tuning: which
// graphs go to union traversals (built or rebuilt here) and which to the
// persistent deep launch.
int batch_plan(msbfs_batch *b) {
    msbfs_context *c = b->eng->ctx;
    drop_graph(b);
    b->all_deep = b->tune.batch_union == 2;
    const uint32_t thr = b->tune.batch_union == 1 ? 1u : (b->tune.union_words ? b->tune.union_words : 2u);
    b->in_union.assign(b->n, 0);
    bool any_union = false;
    for (uint32_t g = 0; g < b->n; ++g) {
        // graphs too large for one workgroup's LDS always join a union
        b->in_union[g] = !b->all_deep && b->S[g] > 0 && b->graphs[g]->V > 0 &&
                         (union_words_of(b->S[g]) >= thr || b->is_wide[g]);
        any_union = any_union || b->in_union[g];
    }
    b->union_mode = any_union;
    int rc;
    uint32_t wmax = 0;
    if (any_union) {
        if ((rc = union_build(b))) return rc;
        msbfs_tuning t = b->tune;
        t.batch_union = 0;
        for (auto &u : b->uclass) {
            if ((rc = msbfs_engine_set_tuning(u.eng, &t))) return rc;
            wmax = std::max(wmax, u.W);
        }
    }
    b->deep_max_v = 0;
    uint64_t deep_sources = 0;
    for (uint32_t g = 0; g < b->n; ++g)
        if (in_deep(b, g)) {
            b->deep_max_v = std::max<uint64_t>(b->deep_max_v, b->graphs[g]->V);
            deep_sources += b->S[g];
        }
    if (deep_sources == 0) {
        // every graph with sources is in a union: no deep launch
        b->plan = Plan{};
        b->plan.family = MSBFS_FAMILY_WIDE;
        b->plan.method = MSBFS_METHOD_PUSH_PULL_ADAPTIVE;
        b->ntasks = 0;
        b->blocks = 0;
        b->last = msbfs_run_stats{};
        b->last.method = b->plan.method;
        b->last.family = MSBFS_FAMILY_WIDE;
        b->last.instances = (uint32_t)b->uclass.size();
        b->last.bits_per_instance = 64 * wmax;
        b->planned = true;
        return MSBFS_OK;
    }
    Plan &p = b->plan;
    p = Plan{};
    p.family = MSBFS_FAMILY_DEEP;
    p.method = MSBFS_METHOD_WORKGROUP_PERSISTENT_BFS;
    p.bits = batch_bits(b);
    if (p.bits == 0)
        return fail(MSBFS_ERR_INVALID_ARG,
                    "batch: %u sources per task of the largest graph (%llu vertices) do not fit "
                    "one workgroup's LDS",
                    b->tune.deep_bits, (unsigned long long)b->max_v);
    for (msbfs_engine *we : b->wide_eng) {
        if ((rc = msbfs_engine_set_tuning(we, &b->tune))) return rc;
    }
    const int B = p.bits;
    std::vector<DeepTask> tasks;
    for (uint32_t g = 0; g < b->n; ++g)
        for (uint64_t k = 0; in_deep(b, g) && k * B < b->S[g]; ++k) {
            DeepTask t{};
            t.graph = g;
            t.src_off = b->src_off[g] + k * B;
            t.nsrc = (uint32_t)std::min<uint64_t>(B, b->S[g] - k * B);
            t.out_off = b->out_off[g] + k * B * b->graphs[g]->V;
            t.reach_bit = 0;
            tasks.push_back(t);
        }
    // Largest graphs first (longest-processing-time order for the queue).
    std::stable_sort(tasks.begin(), tasks.end(), [&](const DeepTask &x, const DeepTask &y) {
        return b->graphs[x.graph]->E > b->graphs[y.graph]->E;
    });
    b->ntasks = tasks.size();
    bool road_like = true;
    for (uint32_t g = 0; g < b->n; ++g)
        if (in_deep(b, g))
            road_like = road_like && b->graphs[g]->max_degree <= 8 && b->graphs[g]->E <= 4 * b->graphs[g]->V;
    deep_layout_v(c, b->tune, b->deep_max_v, b->ell, b->ntasks, p, 0, road_like);
    HIP_TRY(deep_prepare(B, p.ell, p.threads, p.lds));
    int occ = 0;
    HIP_TRY(deep_max_blocks_per_cu(B, p.ell, p.threads, p.lds, &occ));
    if (occ < 1) return fail(MSBFS_ERR_HIP, "batch: deep kernel cannot be resident");
    // With union traversals beside it, the persistent launch takes one
    // workgroup per CU by default and leaves the unions the rest (512
    // G(n,p) graphs, ms per replay by workgroups per CU: 1 5.58, 2 6.52,
    // 3 6.51): its tasks are latency-bound, the unions' levels are not.
    const uint32_t occ_cap = b->tune.deep_blocks_per_cu ? b->tune.deep_blocks_per_cu
                                                        : (b->uclass.empty() ? 0u : 1u);
    if (occ_cap) occ = std::min(occ, (int)occ_cap);
    b->blocks = (int)std::min<uint64_t>(std::max<uint64_t>(b->ntasks, 1), (uint64_t)c->cus * occ);
    const uint64_t cap = std::max<uint64_t>((uint64_t)B * b->deep_max_v, (uint64_t)B);
    const uint64_t spill = cap > p.qcap ? cap - p.qcap : 1;
    const size_t esz = B <= 8 ? 4 : 8;
    if ((rc = ensure_plain(b->queues, (size_t)b->blocks * 2 * spill * esz))) return rc;
    if ((rc = ensure_plain(b->d_tasks, std::max<size_t>(1, tasks.size()) * sizeof(DeepTask))))
        return rc;
    if (!tasks.empty())
        HIP_TRY(hipMemcpy(b->d_tasks.p, tasks.data(), tasks.size() * sizeof(DeepTask),
                          hipMemcpyHostToDevice));
    b->last = msbfs_run_stats{};
    b->last.method = p.method;
    b->last.family = p.family;
    b->last.instances = (uint32_t)(b->ntasks + b->uclass.size());
    b->last.bits_per_instance = std::max<uint32_t>((uint32_t)B, 64 * wmax);
    b->last.grid_blocks = (uint32_t)b->blocks;
    b->last.block_threads = (uint32_t)b->plan.threads;
    b->last.workspace_bytes = b->queues.bytes;
    b->planned = true;
    return MSBFS_OK;
}

DeepParams batch_params(msbfs_batch *b, void *out) {
    const uint32_t B = (uint32_t)b->plan.bits;
    const uint64_t cap = std::max<uint64_t>((uint64_t)B * b->deep_max_v, (uint64_t)B);
    const uint64_t spill = cap > b->plan.qcap ? cap - b->plan.qcap : 1;
    DeepParams p{};
    p.tasks = b->d_tasks.as<DeepTask>();
    p.graphs = b->d_graphs.as<DeepGraph>();
    p.sources = b->d_sources.as<uint32_t>();
    p.out = out;
    p.queues = b->queues.p;
    p.cap = spill + b->plan.qcap;
    p.lds_qcap = b->plan.qcap;
    p.lds_vis_words = b->plan.vis_words;
    p.lds_sink_off = b->plan.sink_off;
    p.staged = b->plan.staged ? 1u : 0u;  // batch outputs are u32 or u16
    p.log_defer_after = (uint32_t)b->ntasks;  // (no logged rows in a batch)
    p.work = b->work.as<uint32_t>();
    p.flags = b->flags.as<uint32_t>();
    p.reach = nullptr;
    p.reach_wpv = 0;
    p.ntasks = (uint32_t)b->ntasks;
    p.out_kind = b->dtype == MSBFS_DIST_U16 ? 1u : 0u;
    return p;
}

// Enqueues one run: clear the task counter and flags with kernels (no
// memset nodes, so the sequence can be captured), the persistent launch,
// and the copy of the flags to the host.
int batch_enqueue(msbfs_batch *b, void *out, hipStream_t st) {
    const DeepParams p = batch_params(b, out);
    HIP_TRY(launch_zero_u32(b->work.as<uint32_t>(), 4, st));
    HIP_TRY(launch_zero_u32(b->flags.as<uint32_t>(), FlagCount, st));
    if (b->ntasks)
        HIP_TRY(launch_deep(b->plan.bits, b->plan.ell, p, b->blocks, b->plan.threads, b->plan.lds,
                            st));
    HIP_TRY(hipMemcpyAsync(b->h_flags.p, b->flags.p, FlagCount * 4, hipMemcpyDeviceToHost, st));
    return MSBFS_OK;
}

// The wide-family graphs, one engine run each on stream `st` (not waited
// for: batch_finish checks them).
int batch_enqueue_wide(msbfs_batch *b, void *out, hipStream_t st) {
    if (!b->all_deep) return MSBFS_OK;  // (big graphs are in unions then)
    for (size_t i = 0; i < b->wide.size(); ++i) {
        const uint32_t g = b->wide[i];
        if (b->S[g] == 0) continue;
        const uint64_t n = b->S[g] * b->graphs[g]->V;
        int rc = msbfs_run_ms_bfs_device(b->wide_eng[i], b->graphs[g],
                                         b->h_sources.data() + b->src_off[g], b->S[g], b->dtype,
                                         (char *)out + b->out_off[g] * dist_bytes(b->dtype), n,
                                         st, 0);
        if (rc) return rc;
    }
    return MSBFS_OK;
}

// The union traversals, one engine run each on the union's own stream,
// forked from `st` at the run's start (ev[2]) and joined back into it (each
// takes the context's launch ordering itself; batch_finish checks them).
int batch_enqueue_union(msbfs_batch *b, void *out, hipStream_t st) {
    int rc = MSBFS_OK;
    for (auto &u : b->uclass) {
        UnionSeeds us;
        us.nseeds = u.seed_v.size();
        us.seed_v = u.seed_v.data();
        us.seed_b = u.seed_b.data();
        us.bmap = u.bmap.as<BatchGraphMap>();
        us.vgraph = u.vgraph.as<uint32_t>();
        hipError_t er = hipStreamWaitEvent(u.stream, b->ev[2], 0);
        if (er != hipSuccess) return hip_fail(er, "union fork");
        rc = engine_run_union(u.eng, u.ug, us, u.smax, b->dtype, out, u.stream);
        er = hipEventRecord(u.done, u.stream);
        if (er == hipSuccess) er = hipStreamWaitEvent(st, u.done, 0);
        if (er != hipSuccess && rc == MSBFS_OK) rc = hip_fail(er, "union join");
        if (rc) return rc;
    }
    return MSBFS_OK;
}

int batch_finish(msbfs_batch *b) {
    if (!b->pending) return MSBFS_OK;
    b->pending = false;
    // every wide and union run is finished (and its errors taken) before the
    // timing event is read
    int other_rc = MSBFS_OK;
    uint32_t other_levels = 0;
    auto take = [&](msbfs_engine *e) {
        int rc = msbfs_engine_check(e);
        if (rc) {
            if (!other_rc) other_rc = rc;
            return;
        }
        msbfs_run_stats ws{};
        if (msbfs_engine_last_stats(e, &ws) == MSBFS_OK) other_levels = std::max(other_levels, ws.levels);
    };
    if (b->all_deep)
        for (size_t i = 0; i < b->wide.size(); ++i)
            if (b->S[b->wide[i]]) take(b->wide_eng[i]);
    if (b->union_mode)
        for (auto &u : b->uclass) take(u.eng);
    HIP_TRY(hipEventSynchronize(b->ev[1]));
    float ms = 0;
    if (hipEventElapsedTime(&ms, b->ev[0], b->ev[1]) == hipSuccess) {
        b->last.kernel_ms = ms;
        b->last.total_ms = ms;
    }
    if (other_rc) return other_rc;
    const uint32_t *f = (const uint32_t *)b->h_flags.p;
    b->last.levels = std::max(f[FlagLevels], other_levels);
    if (f[FlagQueueOverflow]) return fail(MSBFS_ERR_HIP, "batch: internal frontier queue overflow");
    if (f[FlagU16Overflow] && b->dtype == MSBFS_DIST_U16)
        return fail(MSBFS_ERR_DIAMETER_U16, "batch: a distance >= 65535 cannot be stored as u16");
    return MSBFS_OK;
}

// After the persistent launch (eager or replayed): the wide-family runs of
// big graphs (every graph in the deep launch) or the union traversals, each
// taking the context's launch ordering itself, then the end event.
int batch_enqueue_rest(msbfs_batch *b, void *out, hipStream_t st, int sync) {
    b->pending = true;  // (batch_finish collects whatever was started)
    int rc = batch_enqueue_wide(b, out, st);
    if (rc == MSBFS_OK && b->union_mode) rc = batch_enqueue_union(b, out, st);
    if (rc) {
        (void)hipEventRecord(b->ev[1], st);
        (void)batch_finish(b);
        return rc;
    }
    HIP_TRY(hipEventRecord(b->ev[1], st));
    return sync ? batch_finish(b) : MSBFS_OK;
}

int batch_check_out(const msbfs_batch *b, const void *out, uint64_t out_len) {
    if (!out && b->total_out) return fail(MSBFS_ERR_INVALID_ARG, "null output");
    if (out_len < b->total_out)
        return fail(MSBFS_ERR_OUT_TOO_SMALL, "batch output holds %llu elements, %llu needed",
                    (unsigned long long)out_len, (unsigned long long)b->total_out);
    return MSBFS_OK;
}

}  // namespace

extern "C" int msbfs_batch_create(msbfs_engine *e, uint32_t n, msbfs_graph *const *graphs,
                                  const uint64_t *counts, const uint32_t *sources, int dtype,
                                  msbfs_batch **out) {
    clear_error();
    if (!e || !out || (n && (!graphs || !counts)))
        return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    if (dtype != MSBFS_DIST_U32 && dtype != MSBFS_DIST_U16)
        return fail(MSBFS_ERR_INVALID_ARG, "unknown dist type %d", dtype);
    auto *b = new msbfs_batch;
    e->refs.fetch_add(1);
    b->eng = e;
    b->dtype = dtype;
    b->n = n;
    b->tune = e->tune;
    b->S.resize(n);
    b->src_off.resize(n);
    b->out_off.resize(n);
    for (uint32_t g = 0; g < n; ++g) {
        const msbfs_graph *G = graphs[g];
        if (!G) {
            msbfs_batch_destroy(b);
            return fail(MSBFS_ERR_INVALID_ARG, "graph %u is null", g);
        }
        if (G->ctx != e->ctx) {
            msbfs_batch_destroy(b);
            return fail(MSBFS_ERR_INVALID_ARG, "graph %u belongs to another context", g);
        }
        const_cast<msbfs_graph *>(G)->refs.fetch_add(1);
        b->graphs.push_back(G);
        b->S[g] = counts[g];
        b->src_off[g] = b->total_sources;
        b->out_off[g] = b->total_out;
        b->total_sources += counts[g];
        b->total_out += counts[g] * G->V;
        const bool wide = !deep_fits(e->ctx, G->V);
        b->is_wide.push_back(wide ? 1 : 0);
        if (wide)
            b->wide.push_back(g);
        else
            b->max_v = std::max<uint64_t>(b->max_v, G->V);
    }
    // One neighbour-table width for the whole launch: used only when every
    // graph in it has a table of the same width.
    int ell = -1;
    for (uint32_t g = 0; g < n; ++g) {
        if (b->is_wide[g]) continue;
        if (ell < 0)
            ell = b->graphs[g]->ell_width;
        else if (ell != b->graphs[g]->ell_width)
            ell = 0;
    }
    b->ell = std::max(ell, 0);
    for (size_t i = 0; i < b->wide.size(); ++i) {
        msbfs_engine *we = nullptr;
        int rc = msbfs_engine_create(e->ctx, MSBFS_METHOD_PUSH_PULL_ADAPTIVE, &we);
        if (rc) {
            msbfs_batch_destroy(b);
            return rc;
        }
        b->wide_eng.push_back(we);
    }
    if (b->total_sources && !sources) {
        msbfs_batch_destroy(b);
        return fail(MSBFS_ERR_INVALID_ARG, "null sources");
    }
    for (uint32_t g = 0; g < n; ++g) {
        const msbfs_graph *G = b->graphs[g];
        for (uint64_t i = 0; i < b->S[g]; ++i) {
            const uint32_t s = sources[b->src_off[g] + i];
            if ((uint64_t)s >= G->V) {
                msbfs_batch_destroy(b);
                return fail(MSBFS_ERR_SOURCE_OOB,
                            "graph %u: Source vertex %u at index %llu exceeds graph vertex count "
                            "%llu",
                            g, s, (unsigned long long)i, (unsigned long long)G->V);
            }
        }
    }
    b->h_sources.assign(sources, sources + b->total_sources);
    msbfs_context *c = e->ctx;
    hipError_t er = hipSetDevice(c->device);
    if (er == hipSuccess) er = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (int i = 0; i < 4 && er == hipSuccess; ++i) er = hipEventCreate(&b->ev[i]);
    if (er != hipSuccess) {
        msbfs_batch_destroy(b);
        return hip_fail(er, "batch stream");
    }
    std::vector<DeepGraph> tab(std::max<uint32_t>(n, 1));
    for (uint32_t g = 0; g < n; ++g) {
        tab[g].ro = b->graphs[g]->ro;
        tab[g].ci = b->graphs[g]->ci;
        tab[g].ell = b->ell && !b->is_wide[g] ? b->graphs[g]->ell : nullptr;
        tab[g].V = (uint32_t)b->graphs[g]->V;
        tab[g].ell_width = (uint32_t)b->ell;
    }
    int rc;
    if ((rc = ensure_plain(b->d_graphs, tab.size() * sizeof(DeepGraph))) ||
        (rc = ensure_plain(b->d_sources, std::max<uint64_t>(b->total_sources, 1) * 4)) ||
        (rc = ensure_plain(b->work, 16)) || (rc = ensure_plain(b->flags, FlagCount * 4))) {
        msbfs_batch_destroy(b);
        return rc;
    }
    if ((rc = b->h_flags.ensure(FlagCount * 4)) < 0) {
        msbfs_batch_destroy(b);
        return -rc;
    }
    er = hipMemcpy(b->d_graphs.p, tab.data(), tab.size() * sizeof(DeepGraph), hipMemcpyHostToDevice);
    if (er == hipSuccess && b->total_sources)
        er = hipMemcpy(b->d_sources.p, sources, b->total_sources * 4, hipMemcpyHostToDevice);
    if (er != hipSuccess) {
        msbfs_batch_destroy(b);
        return hip_fail(er, "batch upload");
    }
    if ((rc = batch_plan(b))) {
        msbfs_batch_destroy(b);
        return rc;
    }
    *out = b;
    return MSBFS_OK;
}

extern "C" int msbfs_batch_destroy(msbfs_batch *b) {
    if (!b) return MSBFS_OK;
    (void)hipSetDevice(b->eng->ctx->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (msbfs_engine *we : b->wide_eng) msbfs_engine_destroy(we);
    union_release(b);
    drop_graph(b);
    for (auto &ev : b->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (const msbfs_graph *g : b->graphs) graph_release(const_cast<msbfs_graph *>(g));
    msbfs_engine *e = b->eng;
    delete b;
    engine_release(e);
    return MSBFS_OK;
}

extern "C" int msbfs_batch_layout(const msbfs_batch *b, uint64_t *total, uint64_t *offsets) {
    if (!b) return fail(MSBFS_ERR_INVALID_ARG, "null batch");
    if (total) *total = b->total_out;
    if (offsets)
        for (uint32_t g = 0; g < b->n; ++g) offsets[g] = b->out_off[g];
    return MSBFS_OK;
}

extern "C" int msbfs_batch_set_sources(msbfs_batch *b, const uint32_t *sources) {
    clear_error();
    if (!b || (b->total_sources && !sources)) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    for (uint32_t g = 0; g < b->n; ++g)
        for (uint64_t i = 0; i < b->S[g]; ++i)
            if ((uint64_t)sources[b->src_off[g] + i] >= b->graphs[g]->V)
                return fail(MSBFS_ERR_SOURCE_OOB,
                            "graph %u: Source vertex %u at index %llu exceeds graph vertex count "
                            "%llu",
                            g, sources[b->src_off[g] + i], (unsigned long long)i,
                            (unsigned long long)b->graphs[g]->V);
    HIP_TRY(hipSetDevice(b->eng->ctx->device));
    int rc = batch_finish(b);
    if (rc) return rc;
    b->h_sources.assign(sources, sources + b->total_sources);
    if (b->total_sources)
        HIP_TRY(hipMemcpy(b->d_sources.p, sources, b->total_sources * 4, hipMemcpyHostToDevice));
    if (b->union_built) union_seeds(b);
    return MSBFS_OK;
}

extern "C" int msbfs_batch_run(msbfs_batch *b, void *out, uint64_t out_len, void *stream,
                               int sync) {
    clear_error();
    if (!b) return fail(MSBFS_ERR_INVALID_ARG, "null batch");
    int rc = batch_check_out(b, out, out_len);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(b->eng->ctx->device));
    if ((rc = batch_finish(b))) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : b->stream;
    {
        LaunchOrder order(b->eng->ctx, st, false);
        if ((rc = order.begin())) return rc;
        HIP_TRY(hipEventRecord(b->ev[0], st));
        HIP_TRY(hipEventRecord(b->ev[2], st));
        if ((rc = batch_enqueue(b, out, st))) return rc;
        if ((rc = order.end())) return rc;
    }
    return batch_enqueue_rest(b, out, st, sync);
}

extern "C" int msbfs_batch_capture(msbfs_batch *b, void *out, uint64_t out_len, void *stream) {
    clear_error();
    if (!b) return fail(MSBFS_ERR_INVALID_ARG, "null batch");
    int rc = batch_check_out(b, out, out_len);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(b->eng->ctx->device));
    if ((rc = batch_finish(b))) return rc;
    drop_graph(b);
    // (a union traversal's launches depend on its level count, known only on
    // the device, and on per-run level stamps, so unions are not recorded:
    // replays enqueue them eagerly after the recorded launch, as the wide
    // runs of big graphs)
    hipStream_t st = stream ? (hipStream_t)stream : b->stream;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    rc = batch_enqueue(b, out, st);
    hipGraph_t graph = nullptr;
    hipError_t er = hipStreamEndCapture(st, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (er != hipSuccess) return hip_fail(er, "hipStreamEndCapture");
    b->graph = graph;
    er = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (er != hipSuccess) {
        drop_graph(b);
        return hip_fail(er, "hipGraphInstantiate");
    }
    b->graph_out = out;
    return MSBFS_OK;
}

extern "C" int msbfs_batch_replay(msbfs_batch *b, void *stream, int sync) {
    clear_error();
    if (!b) return fail(MSBFS_ERR_INVALID_ARG, "null batch");
    if (!b->exec) return fail(MSBFS_ERR_INVALID_ARG, "batch has no captured graph");
    HIP_TRY(hipSetDevice(b->eng->ctx->device));
    int rc = batch_finish(b);
    if (rc) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : b->stream;
    {
        LaunchOrder order(b->eng->ctx, st, false);
        if ((rc = order.begin())) return rc;
        HIP_TRY(hipEventRecord(b->ev[0], st));
        HIP_TRY(hipEventRecord(b->ev[2], st));
        HIP_TRY(hipGraphLaunch(b->exec, st));
        if ((rc = order.end())) return rc;
    }
    return batch_enqueue_rest(b, b->graph_out, st, sync);
}

extern "C" int msbfs_batch_check(msbfs_batch *b) {
    if (!b) return fail(MSBFS_ERR_INVALID_ARG, "null batch");
    return batch_finish(b);
}

extern "C" int msbfs_batch_stats(msbfs_batch *b, msbfs_run_stats *st) {
    if (!b || !st) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    int rc = batch_finish(b);
    *st = b->last;
    return rc;
}

extern "C" int msbfs_batch_set_tuning(msbfs_batch *b, const msbfs_tuning *t) {
    if (!b || !t) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    int rc = msbfs_tuning_validate(t);
    if (rc) return rc;
    rc = batch_finish(b);
    if (rc) return rc;
    b->tune = *t;
    return batch_plan(b);
}

extern "C" int msbfs_batch_get_tuning(const msbfs_batch *b, msbfs_tuning *t) {
    if (!b || !t) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    *t = b->tune;
    return MSBFS_OK;
}

extern "C" int msbfs_batch_families(const msbfs_batch *b, int32_t *families) {
    if (!b || (b->n && !families)) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    for (uint32_t g = 0; g < b->n; ++g)
        families[g] = (b->in_union[g] || b->is_wide[g]) ? MSBFS_FAMILY_WIDE : MSBFS_FAMILY_DEEP;
    return MSBFS_OK;
}

namespace {

// Plans and times one candidate: a warm-up run, then the mean of two.
msbfs_sweep_entry try_tuning(msbfs_batch *b, const msbfs_tuning &t, void *out, uint64_t out_len) {
    msbfs_sweep_entry r{};
    r.tuning = t;
    r.ms = -1.0f;
    b->tune = t;
    r.status = batch_plan(b);
    float ms = 0;
    for (int it = 0; it < 3 && r.status == MSBFS_OK; ++it) {
        r.status = msbfs_batch_run(b, out, out_len, nullptr, 1);
        if (it > 0) ms += b->last.kernel_ms;
    }
    if (r.status == MSBFS_OK) r.ms = ms / 2;
    b->sweep.push_back(r);
    return r;
}

}  // namespace

// Autotuner.  Deep part (one persistent launch over per-source tasks): every
// feasible (sources per task, workgroup size) pair, and for one source per
// task staged and direct rows.  Union part: the union traversals with the
// default wide knobs.  Wide knobs (when the union traversals win, or some
// graph of the deep launch takes the wide family): from the best plan so
// far, one knob at a time -- wave tile, words per lane, push/pull
// thresholds, launch form -- each value kept when it is faster.
extern "C" int msbfs_batch_autotune(msbfs_batch *b, void *out, uint64_t out_len,
                                    msbfs_tuning *best, float *best_ms) {
    clear_error();
    if (!b) return fail(MSBFS_ERR_INVALID_ARG, "null batch");
    int rc = batch_check_out(b, out, out_len);
    if (rc) return rc;
    if ((rc = batch_finish(b))) return rc;
    b->sweep.clear();
    const msbfs_tuning base = b->tune;
    msbfs_tuning winner = base;
    float win_ms = 1e30f;
    auto consider = [&](const msbfs_tuning &t) {
        const msbfs_sweep_entry r = try_tuning(b, t, out, out_len);
        if (r.status == MSBFS_OK && r.ms < win_ms) {
            win_ms = r.ms;
            winner = t;
        }
    };
    const bool any_deep = b->wide.size() < b->n;
    for (uint32_t bits : {1u, 2u, 4u, 8u, 16u, 32u}) {
        if (!any_deep) break;
        if (((b->max_v * bits + 31) / 32) * 4 + kDeepStaticLds > b->eng->ctx->lds_max) break;
        for (uint32_t threads : {64u, 128u, 256u, 512u, 1024u})
            for (uint32_t rows : {1u, 2u}) {
                // staged rows: one source per task, workgroup kernel
                if ((bits > 1 || (threads == 64 && b->ell > 0)) && rows == 1) continue;
                msbfs_tuning t = base;
                t.batch_union = 2;
                t.deep_bits = bits;
                t.deep_threads = threads;
                t.deep_rows = bits == 1 ? rows : 0;
                consider(t);
            }
    }
    if (b->total_sources) {
        // every graph in unions, then the graphs of at least 2, 4 or 8
        // source words in unions and the rest in the best deep launch
        const msbfs_tuning deep_best = winner;
        msbfs_tuning t = base;
        t.batch_union = 1;
        t.deep_bits = t.deep_threads = t.deep_rows = 0;
        consider(t);
        for (uint32_t words : {2u, 4u, 8u}) {
            bool any = false;
            for (uint32_t g = 0; g < b->n; ++g) any = any || (b->S[g] && union_words_of(b->S[g]) >= words);
            if (!any) break;
            msbfs_tuning h = deep_best;
            h.batch_union = 0;
            h.union_words = words;
            consider(h);
        }
    }
    if (winner.batch_union != 2 || !b->wide.empty()) {
        if (win_ms >= 1e30f) consider(winner);
        auto knob = [&](uint32_t msbfs_tuning::*field, std::initializer_list<uint32_t> vals) {
            // from the best so far, each value of one knob
            const msbfs_tuning from = winner;
            for (uint32_t v : vals) {
                if (from.*field == v) continue;
                msbfs_tuning t = from;
                t.*field = v;
                consider(t);
            }
        };
        // tiles up to twice the widest wide graph's default (lanes past the
        // graph's source words stay idle)
        uint64_t wmax = 1;
        if (winner.batch_union != 2)
            for (uint32_t g = 0; g < b->n; ++g) {
                if (b->S[g]) wmax = std::max<uint64_t>(wmax, union_words_of(b->S[g]));
            }
        else
            for (uint32_t g : b->wide) wmax = std::max<uint64_t>(wmax, (b->S[g] + 63) / 64);
        uint32_t tw0 = 1;
        while (tw0 < 64 && tw0 < wmax) tw0 *= 2;
        std::vector<uint32_t> tws;
        for (uint32_t tw = 1; tw <= std::min(64u, 2 * tw0); tw *= 2) tws.push_back(tw);
        {
            const msbfs_tuning from = winner;
            for (uint32_t v : tws) {
                msbfs_tuning t = from;
                t.tile_words = v;
                consider(t);
            }
        }
        knob(&msbfs_tuning::words_per_lane, {1u, 2u});
        knob(&msbfs_tuning::alpha, {4u, 8u, 28u});
        knob(&msbfs_tuning::beta, {12u, 48u});
        knob(&msbfs_tuning::stepped, {1u, 2u});
        if (winner.batch_union == 0) {
            // beside the unions the deep launch's best shape differs from
            // its best alone (512 G(n,p) graphs, ms per replay: 640 threads
            // x 1 workgroup per CU 5.45, 512 x 1 5.55, 768 x 1 5.63, 1024 x
            // 1 6.13, 256 x 2 5.70, 512 x 3 6.52)
            knob(&msbfs_tuning::deep_threads, {256u, 512u, 640u, 768u, 1024u});
            knob(&msbfs_tuning::deep_blocks_per_cu, {1u, 2u, 4u});
        }
    }
    b->tune = winner;
    if ((rc = batch_plan(b))) return rc;
    if (win_ms >= 1e30f) return fail(MSBFS_ERR_HIP, "autotune: no candidate ran");
    if (best) *best = winner;
    if (best_ms) *best_ms = win_ms;
    clear_error();
    return MSBFS_OK;
}

extern "C" int msbfs_batch_sweep(const msbfs_batch *b, msbfs_sweep_entry *entries, uint32_t cap,
                                 uint32_t *n) {
    if (!b || !n || (cap && !entries)) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    *n = (uint32_t)b->sweep.size();
    for (uint32_t i = 0; i < cap && i < *n; ++i) entries[i] = b->sweep[i];
    return MSBFS_OK;
}
