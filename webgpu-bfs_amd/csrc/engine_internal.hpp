// Internal host-side types of libmsbfs shared by the engine (engine.cpp)
// and the batched and serving front ends (batch.cpp, server.cpp): device
// buffers, the context / graph / engine objects and the run plan.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace msbfs {

inline int hip_fail(hipError_t e, const char *what) {
    const int code = (e == hipErrorOutOfMemory) ? MSBFS_ERR_OOM : MSBFS_ERR_HIP;
    return fail(code, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr)                                      \
    do {                                                   \
        hipError_t _e = (expr);                            \
        if (_e != hipSuccess) return hip_fail(_e, #expr);  \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // Grows the buffer to at least `need` bytes.  Returns 1 when it was
    // (re)allocated (contents undefined), 0 when kept, <0 on failure.
    int ensure(size_t need) {
        if (need == 0) need = 16;
        if (bytes >= need) return 0;
        release();
        size_t want = need + std::min<size_t>(need / 8, (size_t)1 << 30);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            return -hip_fail(e, "hipMalloc workspace");
        }
        bytes = want;
        return 1;
    }
    template <class T>
    T *as() const { return (T *)p; }
};

struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    int ensure(size_t need) {
        if (need == 0) need = 16;
        if (bytes >= need) return 0;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipHostMalloc(&p, need, hipHostMallocDefault);
        if (e != hipSuccess) return -hip_fail(e, "hipHostMalloc");
        bytes = need;
        return 1;
    }
};

constexpr uint32_t kDefaultAlpha = 14;  // push_pull.rs:29-35
constexpr uint32_t kDefaultBeta = 24;

inline bool is_pull_only(int m) {
    return m == MSBFS_METHOD_BIT_PARALLEL_64_PULL || m == MSBFS_METHOD_MULTI_WORD_BIT_PARALLEL ||
           m == MSBFS_METHOD_INDIRECT_MACRO_BATCH;
}
inline bool is_push_only(int m) {
    return m == MSBFS_METHOD_BIT_PARALLEL_32 || m == MSBFS_METHOD_BIT_PARALLEL_64 ||
           m == MSBFS_METHOD_BOOLEAN_REACHABILITY || m == MSBFS_METHOD_SPARSE_QUEUE ||
           m == MSBFS_METHOD_SPARSE_BIT_PARALLEL_64 ||
           m == MSBFS_METHOD_MULTI_WORD_BIT_PARALLEL_PUSH;
}

}  // namespace msbfs

// Context, graphs and engines are reference counted: a graph or engine
// holds its context, a batch holds its engine and graphs, so the public
// destroy calls may come in any order.
struct msbfs_context {
    std::atomic<int> refs{1};
    int device = 0;
    // Launch ordering.  A cooperative (grid-barrier) traversal needs all its
    // workgroups resident at once, so it must not share the device with
    // other traversals of this library: it waits on the last launch of every
    // other stream, and every later launch waits on it.  All of this is
    // expressed with stream events; the host never blocks.
    std::mutex order_mu;
    hipEvent_t coop_done = nullptr;
    bool coop_valid = false;
    std::vector<std::pair<hipStream_t, hipEvent_t>> stream_done;
    hipDeviceProp_t prop{};
    hipStream_t stream = nullptr;
    int cus = 0;
    size_t lds_max = 0;
    int coop = 0;
    // device address of this GPU's copy of the no-alloc scratch array
    void *scratch = nullptr;
};

struct msbfs_graph {
    std::atomic<int> refs{1};
    msbfs_context *ctx = nullptr;
    uint64_t V = 0, E = 0;
    uint32_t max_degree = 0;
    int symmetric = 0;
    uint32_t *ro = nullptr, *ci = nullptr, *co = nullptr, *ri = nullptr;
    uint32_t *ell = nullptr;  // V x ell_width neighbour slots, 0xFFFFFFFF padded
    int ell_width = 0;        // 4 or 8 when max degree allows, else 0
    // lattice adjacency (DeepGraph::lat): lat_bits bits per vertex (2 or 4;
    // 0 when the graph is not a lattice), lat_words u32 words on the device
    uint32_t *lat = nullptr;
    int lat_bits = 0;
    uint32_t lat_words = 0;
    uint32_t lat_d[4] = {0, 0, 0, 0};
    uint32_t lat_pad = 0;     // zero fields before vertex 0's (>= the largest offset)
    uint32_t *heavy_out = nullptr, *heavy_in = nullptr;  // degree > kHeavyDegree
    uint32_t n_heavy_out = 0, n_heavy_in = 0;
    uint32_t *hchunk_out = nullptr, *hchunk_in = nullptr;  // uint4 (v, e0, e1, 0) items
    uint32_t n_hchunk_out = 0, n_hchunk_in = 0;
    // uint4 (v, e0, e1, 0): vertices with 1..kHeavyDegree in-edges, increasing v
    uint32_t *light_in = nullptr;
    uint32_t *light_in_deg = nullptr;  // the same items by decreasing in-degree
    uint32_t n_light_in = 0;
    uint64_t bytes = 0;
    // Most BFS levels a deep traversal with one source per task has needed
    // on this graph (0: none yet); |V| / levels estimates its frontier size.
    std::atomic<uint32_t> deep_levels{0};
    // Most levels a wide traversal has needed on this graph (0: none yet):
    // how many levels a stepped launch enqueues up front.
    std::atomic<uint32_t> wide_levels{0};
};

struct msbfs_engine {
    std::atomic<int> refs{1};
    msbfs_context *ctx = nullptr;
    int method = MSBFS_METHOD_AUTO;
    msbfs_tuning tune{};
    std::mutex mu;
    // wide workspace (kept all-zero between runs by the kernels, except the
    // frontier slots `fr`, whose rows are written whole before they are read)
    msbfs::DevBuf seen, fr, nx, qf0, qf1, qc, stamp, ctr, fbA, fbB, lv, zrow, act;
    uint64_t wide_words = 0;  // entries allocated per state array
    uint32_t tick_base = 1;
    bool dirty = true;
    // deep workspace
    msbfs::DevBuf queues, tasks, graphs, log_sink, log_defer;
    // shared
    msbfs::DevBuf flags, sources, out_tmp, out_pack, reach_tmp, reach_t, trace;
    // host-output path: pinned staging chunks and their copy-done events
    msbfs::PinnedBuf h_stage;
    hipEvent_t stage_ev[3] = {nullptr, nullptr, nullptr};
    uint32_t trace_cap = 0;
    bool trace_valid = false;
    msbfs::PinnedBuf h_sources, h_flags, h_tasks;
    // what the device task table and graph record hold (skip re-uploading them)
    uint64_t tasks_key[10] = {};
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    msbfs_run_stats last{};
    hipStream_t last_stream = nullptr;
    bool pending = false;
    int pending_dtype = 0;
    msbfs_graph *pending_graph = nullptr;  // held until its level count is recorded
    // stepped wide launch of the pending run: its parameters, for enqueueing
    // more levels when the frontier outlasts those enqueued
    bool pending_stepped = false;
    msbfs::WideParams step_p{};
    int step_tw = 0, step_blocks = 0;
    size_t step_lds = 0;
    uint32_t step_levels = 0;
};


namespace msbfs {

struct Plan {
    int method = 0;
    int family = MSBFS_FAMILY_NONE;
    int policy = MSBFS_DIR_ADAPTIVE;
    int bits = 0;       // deep: sources per task
    int ell = 0;        // deep: neighbour-table width (0: CSR)
    int threads = 0;    // deep: workgroup size
    uint32_t vis_words = 0;  // deep: LDS words of visited bits
    uint32_t qcap = 0;  // deep: LDS entries per frontier queue
    uint32_t log_bw = 0;   // deep: vertex bin bits of the logged u32 output (0: none)
    uint32_t bin_off = 0;  // deep: LDS word offset of the log's bin counters
    uint32_t sink_off = 0; // deep one-wave kernel: LDS word offset of its append sink
    bool staged = false;   // deep: u32/u16 rows staged in LDS as level codes
    size_t lds = 0;     // deep: dynamic LDS bytes per workgroup
    int lat = 0;           // deep lattice kernel: adjacency bits per vertex (0: not used)
    uint32_t lat_waves = 0;   // deep lattice kernel: waves (tasks at once) per workgroup
    uint32_t lat_region = 0;  // deep lattice kernel: LDS words per wave
    uint32_t lat_words = 0;   // deep lattice kernel: LDS words of the adjacency
};

void context_release(msbfs_context *c);
// Looks up this device's address of the no-alloc scratch array (with the
// context's device current) and keeps it in the context.
int noalloc_setup(msbfs_context *c);
// Brackets the enqueue of one traversal on stream `st` (coop = it contains a
// cooperative launch).  The context's ordering lock is held from begin() to
// the guard's end of life, so no other traversal is enqueued in between.
class LaunchOrder {
  public:
    LaunchOrder(msbfs_context *c, hipStream_t st, bool coop)
        : c_(c), st_(st), coop_(coop), lk_(c->order_mu) {}
    int begin();  // stream waits that keep a cooperative kernel alone
    int end();    // records this traversal's completion event
  private:
    msbfs_context *c_;
    hipStream_t st_;
    bool coop_;
    std::unique_lock<std::mutex> lk_;
};
void graph_release(msbfs_graph *g);
int graph_create_impl(msbfs_context *c, uint64_t V, uint64_t E, const uint32_t *ro,
                      const uint32_t *ci, const uint32_t *vbound, msbfs_graph **out);
// A batch union traversal (msbfs_batch): seeds as (vertex, source bit)
// pairs, and the union's per-graph output map (device arrays).
struct UnionSeeds {
    uint64_t nseeds = 0;
    const uint32_t *seed_v = nullptr, *seed_b = nullptr;  // host
    const BatchGraphMap *bmap = nullptr;                   // device
    const uint32_t *vgraph = nullptr;                      // device
};
// Enqueues one wide traversal of a union graph on `st` (not waited for:
// msbfs_engine_check or the next call finishes it).  S: the largest source
// count of the union's graphs.
int engine_run_union(msbfs_engine *e, const msbfs_graph *g, const UnionSeeds &u, uint64_t S,
                     int dtype, void *out, hipStream_t st);
void engine_release(msbfs_engine *e);

int ensure_zeroed(DevBuf &b, size_t bytes, bool force, hipStream_t st);
int ensure_plain(DevBuf &b, size_t bytes);
// LDS split of a deep-family workgroup for graphs of up to max_v vertices.
// levels_seen: the graphs' deep level count from an earlier traversal (0:
// none yet); road_like: every graph passes the reference's road-like test
// (max degree <= 8 and |E| <= 4|V|), taken as a sign of many levels.
void deep_layout_v(const msbfs_context *c, const msbfs_tuning &t, uint64_t max_v, int ell,
                   uint64_t ntasks, Plan &p, uint32_t levels_seen, bool road_like);
int deep_bits_for(const msbfs_context *c, const msbfs_tuning &t, uint64_t V, uint64_t S);
// LDS split of the lattice kernel for one graph (adjacency, then one region
// per wave); false when it does not fit.
bool lattice_layout(const msbfs_context *c, const msbfs_graph *g, uint64_t ntasks, bool logged_ok,
                    Plan &p);

}  // namespace msbfs
