// Serving loop: resident graphs, continuous batching of BFS queries.
//
// Each resident graph is a lane with several launch slots, each an engine
// workspace and a HIP stream.  Queries (a graph and 1..512 sources) wait in
// the lane's FIFO.  Whenever a lane has a free slot, the scheduler thread
// admits the longest prefix of the FIFO whose distance outputs fit in the
// output pool (u16/u32 x k x |V| bytes each) and whose sources fit the
// launch cap, packs their sources back to back into the launch's 64-bit
// source words (so queries share words), and enqueues one traversal whose
// output rows land directly in the admitted queries' pool slices.  A query
// that arrives while the lane's traversals run therefore starts as soon as
// a slot and its output memory are free, not when the running traversals
// end: its kernels take the workgroup slots the running ones give up (a
// persistent deep launch retires workgroups as its task queue drains) and
// run beside them.  Queries arriving while every slot is busy accumulate
// and are packed into the next launch.  Wide-family lanes run stepped
// launches (no grid barrier, no co-resident grid needed), so they share the
// device with the other lanes too.  Results stay in device memory until the
// client releases them, which returns their bytes to the pool and lets
// waiting queries in.
//
// The reference has no serving path; its engine serves one call at a time
// (crates/gpu_bfs/src/engine.rs:61-63 serialises calls on pool mutexes).
#include <chrono>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>

#include "engine_internal.hpp"

using namespace msbfs;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over one device buffer, with coalescing frees.
class Pool {
public:
    void init(uint64_t size) {
        free_.clear();
        free_[0] = size;
        size_ = size;
        used_ = 0;
    }
    bool alloc(uint64_t n, uint64_t *off) {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                *off = it->first;
                const uint64_t rest = it->second - n, start = it->first + n;
                free_.erase(it);
                if (rest) free_[start] = rest;
                used_ += n;
                return true;
            }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t size() const { return size_; }
    uint64_t used() const { return used_; }

private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t size_ = 0, used_ = 0;
};

struct Query {
    uint64_t id = 0;
    uint32_t lane = 0;
    std::vector<uint32_t> sources;
    uint64_t bytes = 0;       // output bytes (k x |V| x b_d)
    uint64_t off = 0;         // pool offset while admitted
    int state = 0;            // 0 pending, 1 running, 2 done, 3 failed
    int status = MSBFS_OK;
    std::string error;
    Clock::time_point t_submit, t_start, t_done;
    uint64_t launch = 0;
};

// One launch slot of a lane: its own engine workspace, stream and
// completion event, and the queries of the traversal it runs.
struct Slot {
    msbfs_engine *eng = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::vector<uint64_t> running;
    bool inflight = false;
    std::vector<uint32_t> packed;
};

struct Lane {
    const msbfs_graph *g = nullptr;
    std::vector<Slot> slots;
    std::deque<uint64_t> pending;
    uint64_t launches = 0, sources = 0;
};

}  // namespace

struct msbfs_server {
    msbfs_context *ctx = nullptr;
    msbfs_server_config cfg{};
    std::vector<Lane> lanes;
    DevBuf pool_buf;
    Pool pool;
    std::map<uint64_t, Query> queries;
    uint64_t next_id = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool running = false, stopping = false;
    uint64_t completed = 0, failed = 0, launches = 0, sources_done = 0;
    uint64_t overlapped = 0, max_in_flight = 0, in_flight = 0;
    Clock::time_point t_first_submit, t_last_done;
    bool any_submit = false;
    std::vector<double> latencies_ms;
    std::deque<uint64_t> completions;  // finished (done or failed), not yet reported
    bool log = false;                  // MSBFS_SERVER_LOG=1: one stderr line per launch/retire
    Clock::time_point t0 = Clock::now();
};

namespace {

size_t elem_bytes(const msbfs_server *s) { return s->cfg.dist_type == MSBFS_DIST_U16 ? 2 : 4; }

// Admits and launches work on every free slot of every lane; called with
// the lock held.  A launch takes the longest FIFO prefix that fits the source
// and query caps (one query when packing is off) and whose outputs fit one
// block of the pool; the block is carved into the queries' slices in order,
// so query i's rows follow query i-1's and one packed traversal fills them
// all.
// Beside a running traversal of the same lane a new launch waits for half
// a full pack, or until its oldest query has waited admit_wait_us: at low
// load queries start at once, under saturation packs stay large (packed
// sources share source words; concurrent small packs of one graph would
// each pay for a whole traversal).
bool worth_launching(const msbfs_server *s, const Lane &L, Clock::time_point now) {
    bool busy = false;
    for (const Slot &sl : L.slots) busy = busy || sl.inflight;
    if (!busy || !s->cfg.pack) return true;
    uint64_t nsrc = 0;
    for (uint64_t id : L.pending) {
        nsrc += s->queries.at(id).sources.size();
        if (2 * nsrc >= s->cfg.max_sources_per_launch) return true;
    }
    const Query &head = s->queries.at(L.pending.front());
    return now - head.t_submit >= std::chrono::microseconds(s->cfg.admit_wait_us);
}

void schedule(msbfs_server *s) {
    for (uint32_t li = 0; li < s->lanes.size(); ++li) {
        Lane &L = s->lanes[li];
        for (Slot &sl : L.slots) {
            if (L.pending.empty()) break;
            if (sl.inflight) continue;
            if (!worth_launching(s, L, Clock::now())) break;
            std::vector<uint64_t> take;
            uint64_t nsrc = 0, total = 0;
            for (uint64_t id : L.pending) {
                const Query &q = s->queries[id];
                const uint64_t k = q.sources.size();
                if (!take.empty() && (nsrc + k > s->cfg.max_sources_per_launch ||
                                      take.size() >= s->cfg.max_queries_per_launch ||
                                      !s->cfg.pack))
                    break;
                take.push_back(id);
                nsrc += k;
                total += q.bytes;
            }
            uint64_t base = 0;
            while (!take.empty() && !s->pool.alloc(total, &base)) {
                const Query &q = s->queries[take.back()];
                total -= q.bytes;
                nsrc -= q.sources.size();
                take.pop_back();
            }
            if (take.empty()) break;  // the head query waits for released memory
            L.pending.erase(L.pending.begin(), L.pending.begin() + take.size());
            sl.packed.clear();
            const Clock::time_point now = Clock::now();
            uint64_t off = base;
            for (uint64_t id : take) {
                Query &q = s->queries[id];
                q.off = off;
                off += q.bytes;
                q.state = 1;
                q.t_start = now;
                q.launch = s->launches;
                sl.packed.insert(sl.packed.end(), q.sources.begin(), q.sources.end());
            }
            int rc = msbfs_run_ms_bfs_device(sl.eng, L.g, sl.packed.data(), sl.packed.size(),
                                             s->cfg.dist_type, (char *)s->pool_buf.p + base,
                                             sl.packed.size() * L.g->V, sl.stream, 0);
            if (rc == MSBFS_OK && hipEventRecord(sl.done, sl.stream) != hipSuccess)
                rc = MSBFS_ERR_HIP;
            if (rc != MSBFS_OK) {
                const std::string msg = msbfs_last_error();
                for (uint64_t id : take) {
                    Query &q = s->queries[id];
                    q.state = 3;
                    q.status = rc;
                    q.error = msg;
                    q.t_done = Clock::now();
                    s->pool.release(q.off, q.bytes);
                    ++s->failed;
                    s->completions.push_back(id);
                }
                s->cv_done.notify_all();
                continue;
            }
            if (s->in_flight) ++s->overlapped;
            ++s->in_flight;
            s->max_in_flight = std::max(s->max_in_flight, s->in_flight);
            if (s->log) {
                int m = 0;
                const int fam = msbfs_engine_resolve(sl.eng, L.g, sl.packed.size(), &m);
                fprintf(stderr, "[msbfs-server] %.3f ms launch #%llu lane %u slot %zu: %zu queries, "
                        "%zu sources, family %d, %llu in flight, pool %.2f/%.2f GiB\n",
                        std::chrono::duration<double, std::milli>(Clock::now() - s->t0).count(),
                        (unsigned long long)s->launches, li, (size_t)(&sl - L.slots.data()),
                        take.size(), sl.packed.size(), fam, (unsigned long long)s->in_flight,
                        s->pool.used() / 1073741824.0, s->pool.size() / 1073741824.0);
                fflush(stderr);
            }
            sl.running = take;
            sl.inflight = true;
            L.launches++;
            L.sources += nsrc;
            s->launches++;
        }
    }
}

// Retires finished launches; called with the lock held.
bool reap(msbfs_server *s) {
    bool any = false;
    for (Lane &L : s->lanes) {
        for (Slot &sl : L.slots) {
            if (!sl.inflight) continue;
            if (hipEventQuery(sl.done) != hipSuccess) continue;
            const int rc = msbfs_engine_check(sl.eng);
            const std::string msg = rc ? msbfs_last_error() : "";
            const Clock::time_point now = Clock::now();
            for (uint64_t id : sl.running) {
                Query &q = s->queries[id];
                q.t_done = now;
                s->completions.push_back(id);
                if (rc == MSBFS_OK) {
                    q.state = 2;
                    ++s->completed;
                    s->sources_done += q.sources.size();
                    s->latencies_ms.push_back(
                        std::chrono::duration<double, std::milli>(q.t_done - q.t_submit).count());
                } else {
                    q.state = 3;
                    q.status = rc;
                    q.error = msg;
                    s->pool.release(q.off, q.bytes);
                    ++s->failed;
                }
            }
            s->t_last_done = now;
            if (s->log) {
                fprintf(stderr, "[msbfs-server] %.3f ms retire lane %zu slot %zu: %zu queries, status %d\n",
                        std::chrono::duration<double, std::milli>(now - s->t0).count(),
                        (size_t)(&L - s->lanes.data()), (size_t)(&sl - L.slots.data()),
                        sl.running.size(), rc);
                fflush(stderr);
            }
            sl.running.clear();
            sl.inflight = false;
            --s->in_flight;
            any = true;
        }
    }
    if (any) s->cv_done.notify_all();
    return any;
}

void worker_main(msbfs_server *s) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stopping) {
        reap(s);
        schedule(s);
        const bool busy = s->in_flight != 0;
        if (busy) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            lk.lock();
        } else {
            s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
    // drain
    for (;;) {
        if (s->in_flight == 0) break;
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(50));
        lk.lock();
        reap(s);
    }
}

}  // namespace

extern "C" int msbfs_server_create(msbfs_context *c, const msbfs_server_config *cfg,
                                   msbfs_server **out) {
    clear_error();
    if (!c || !out) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    auto *s = new msbfs_server;
    s->ctx = c;
    c->refs.fetch_add(1);
    if (cfg) s->cfg = *cfg;
    const char *lg = getenv("MSBFS_SERVER_LOG");
    s->log = lg && lg[0] == '1';
    if (!s->cfg.output_pool_bytes) s->cfg.output_pool_bytes = 16ull << 30;
    if (!s->cfg.max_sources_per_launch) s->cfg.max_sources_per_launch = 2048;
    if (!s->cfg.max_queries_per_launch) s->cfg.max_queries_per_launch = 512;
    if (!s->cfg.slots_per_lane) s->cfg.slots_per_lane = 4;
    if (!s->cfg.admit_wait_us) s->cfg.admit_wait_us = 2000;
    if (s->cfg.slots_per_lane > 16) {
        msbfs_server_destroy(s);
        return fail(MSBFS_ERR_INVALID_ARG, "slots_per_lane must be <= 16");
    }
    if (s->cfg.dist_type != MSBFS_DIST_U32 && s->cfg.dist_type != MSBFS_DIST_U16) {
        msbfs_server_destroy(s);
        return fail(MSBFS_ERR_INVALID_ARG, "unknown dist type %d", s->cfg.dist_type);
    }
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) {
        msbfs_server_destroy(s);
        return hip_fail(e, "hipSetDevice");
    }
    int r = s->pool_buf.ensure(s->cfg.output_pool_bytes);
    if (r < 0) {
        msbfs_server_destroy(s);
        return -r;
    }
    s->pool.init(s->pool_buf.bytes);
    *out = s;
    return MSBFS_OK;
}

extern "C" int msbfs_server_add_graph(msbfs_server *s, msbfs_graph *g, int method,
                                      uint32_t *lane_id) {
    clear_error();
    if (!s || !g) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->running) return fail(MSBFS_ERR_INVALID_ARG, "server already started");
    Lane L;
    L.g = g;
    auto drop = [](Slot &sl) {
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.eng) msbfs_engine_destroy(sl.eng);
    };
    for (uint32_t k = 0; k < s->cfg.slots_per_lane; ++k) {
        Slot sl;
        int rc = msbfs_engine_create(s->ctx, method, &sl.eng);
        if (rc) {
            for (Slot &o : L.slots) drop(o);
            return rc;
        }
        // wide-family traversals as stepped launches: no co-resident grid,
        // so they run beside the other slots' and lanes' traversals
        msbfs_tuning t{};
        msbfs_engine_get_tuning(sl.eng, &t);
        t.stepped = 1;
        msbfs_engine_set_tuning(sl.eng, &t);
        hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e != hipSuccess) {
            drop(sl);
            for (Slot &o : L.slots) drop(o);
            return hip_fail(e, "lane stream");
        }
        L.slots.push_back(sl);
    }
    g->refs.fetch_add(1);
    s->lanes.push_back(std::move(L));
    if (lane_id) *lane_id = (uint32_t)(s->lanes.size() - 1);
    return MSBFS_OK;
}

extern "C" int msbfs_server_set_tuning(msbfs_server *s, uint32_t lane, const msbfs_tuning *t) {
    if (!s || !t || lane >= s->lanes.size()) return fail(MSBFS_ERR_INVALID_ARG, "bad argument");
    int rc = msbfs_tuning_validate(t);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    for (Slot &sl : s->lanes[lane].slots)
        if ((rc = msbfs_engine_set_tuning(sl.eng, t))) return rc;
    return MSBFS_OK;
}

extern "C" int msbfs_server_start(msbfs_server *s) {
    if (!s) return fail(MSBFS_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->running) return MSBFS_OK;
    // Every slot runs one traversal of the widest launch it can get (the
    // source cap, or what the output pool holds) before the first query, so
    // its workspace is sized once: a workspace that grows while serving
    // frees (a device-wide synchronisation) and allocates gigabytes under
    // the scheduler's lock, stalling every lane.
    HIP_TRY(hipSetDevice(s->ctx->device));
    for (Lane &L : s->lanes) {
        if (L.g->V == 0) continue;
        const uint64_t fit = s->pool.size() / (L.g->V * elem_bytes(s));
        const uint64_t k = std::min<uint64_t>(s->cfg.max_sources_per_launch, fit);
        if (k == 0) continue;
        std::vector<uint32_t> warm(k, 0u);
        for (Slot &sl : L.slots) {
            int rc = msbfs_run_ms_bfs_device(sl.eng, L.g, warm.data(), k, s->cfg.dist_type,
                                             s->pool_buf.p, k * L.g->V, sl.stream, 1);
            if (rc) return rc;
        }
    }
    s->running = true;
    s->stopping = false;
    s->worker = std::thread(worker_main, s);
    return MSBFS_OK;
}

extern "C" int msbfs_server_submit(msbfs_server *s, uint32_t lane, const uint32_t *sources,
                                   uint32_t k, uint64_t *query_id) {
    clear_error();
    if (!s || !query_id || (k && !sources)) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    if (k == 0 || k > 512) return fail(MSBFS_ERR_INVALID_ARG, "a query has 1..512 sources, got %u", k);
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->running || s->stopping) return fail(MSBFS_ERR_BUSY, "server is not running");
    if (lane >= s->lanes.size()) return fail(MSBFS_ERR_INVALID_ARG, "unknown graph lane %u", lane);
    const msbfs_graph *g = s->lanes[lane].g;
    for (uint32_t i = 0; i < k; ++i)
        if ((uint64_t)sources[i] >= g->V)
            return fail(MSBFS_ERR_SOURCE_OOB,
                        "Source vertex %u at index %u exceeds graph vertex count %llu", sources[i],
                        i, (unsigned long long)g->V);
    Query q;
    q.id = s->next_id++;
    q.lane = lane;
    q.sources.assign(sources, sources + k);
    q.bytes = (uint64_t)k * g->V * elem_bytes(s);
    if (q.bytes > s->pool.size())
        return fail(MSBFS_ERR_BUSY, "query output (%llu bytes) exceeds the output pool",
                    (unsigned long long)q.bytes);
    q.t_submit = Clock::now();
    if (!s->any_submit) {
        s->any_submit = true;
        s->t_first_submit = q.t_submit;
    }
    s->lanes[lane].pending.push_back(q.id);
    *query_id = q.id;
    s->queries.emplace(q.id, std::move(q));
    s->cv_work.notify_one();
    return MSBFS_OK;
}

extern "C" int msbfs_server_wait(msbfs_server *s, uint64_t id, double timeout_s) {
    if (!s) return fail(MSBFS_ERR_INVALID_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->queries.find(id);
    if (it == s->queries.end()) return fail(MSBFS_ERR_INVALID_ARG, "unknown query %llu", (unsigned long long)id);
    const auto deadline = Clock::now() + std::chrono::duration<double>(timeout_s);
    while (it->second.state < 2) {
        if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && it->second.state < 2)
            return fail(MSBFS_ERR_BUSY, "query %llu not finished", (unsigned long long)id);
    }
    if (it->second.state == 3) return fail(it->second.status, "%s", it->second.error.c_str());
    return MSBFS_OK;
}

extern "C" int msbfs_server_next_completed(msbfs_server *s, double timeout_s, uint64_t *query_id) {
    if (!s || !query_id) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    const auto deadline = Clock::now() + std::chrono::duration<double>(timeout_s);
    while (s->completions.empty()) {
        if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && s->completions.empty())
            return fail(MSBFS_ERR_BUSY, "no query completed within %.3f s", timeout_s);
    }
    *query_id = s->completions.front();
    s->completions.pop_front();
    return MSBFS_OK;
}

extern "C" int msbfs_server_query_info(msbfs_server *s, uint64_t id, msbfs_query_info *info) {
    if (!s || !info) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->queries.find(id);
    if (it == s->queries.end()) return fail(MSBFS_ERR_INVALID_ARG, "unknown query");
    const Query &q = it->second;
    info->state = q.state;
    info->status = q.status;
    info->lane = q.lane;
    info->num_sources = (uint32_t)q.sources.size();
    info->launch = q.launch;
    info->device_ptr = q.state == 2 ? (char *)s->pool_buf.p + q.off : nullptr;
    info->elements = (uint64_t)q.sources.size() * s->lanes[q.lane].g->V;
    auto ms = [](Clock::time_point a, Clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    info->queue_ms = q.state >= 1 ? ms(q.t_submit, q.t_start) : 0.0;
    info->latency_ms = q.state >= 2 ? ms(q.t_submit, q.t_done) : 0.0;
    return MSBFS_OK;
}

extern "C" int msbfs_server_copy_result(msbfs_server *s, uint64_t id, void *host, uint64_t len) {
    if (!s || !host) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    void *src = nullptr;
    uint64_t n = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->queries.find(id);
        if (it == s->queries.end() || it->second.state != 2)
            return fail(MSBFS_ERR_INVALID_ARG, "query %llu has no result", (unsigned long long)id);
        const Query &q = it->second;
        n = (uint64_t)q.sources.size() * s->lanes[q.lane].g->V;
        src = (char *)s->pool_buf.p + q.off;
    }
    if (len < n) return fail(MSBFS_ERR_OUT_TOO_SMALL, "result has %llu elements", (unsigned long long)n);
    HIP_TRY(hipMemcpy(host, src, n * elem_bytes(s), hipMemcpyDeviceToHost));
    return MSBFS_OK;
}

extern "C" int msbfs_server_release(msbfs_server *s, uint64_t id) {
    if (!s) return fail(MSBFS_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->queries.find(id);
    if (it == s->queries.end()) return fail(MSBFS_ERR_INVALID_ARG, "unknown query");
    if (it->second.state == 2) s->pool.release(it->second.off, it->second.bytes);
    if (it->second.state >= 2) {
        s->queries.erase(it);
        s->cv_work.notify_one();
        return MSBFS_OK;
    }
    return fail(MSBFS_ERR_BUSY, "query still pending or running");
}

extern "C" int msbfs_server_get_stats(msbfs_server *s, msbfs_server_stats *st) {
    if (!s || !st) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(st, 0, sizeof(*st));
    st->completed = s->completed;
    st->failed = s->failed;
    st->launches = s->launches;
    st->sources = s->sources_done;
    st->pool_bytes = s->pool.size();
    st->pool_used = s->pool.used();
    uint64_t pend = 0;
    for (const Lane &L : s->lanes) pend += L.pending.size();
    st->pending = pend;
    if (s->any_submit && s->completed)
        st->span_s = std::chrono::duration<double>(s->t_last_done - s->t_first_submit).count();
    std::vector<double> v = s->latencies_ms;
    st->overlapped_launches = s->overlapped;
    st->max_in_flight = s->max_in_flight;
    if (!v.empty()) {
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            size_t i = (size_t)std::min<double>(v.size() - 1, std::floor(p * (v.size() - 1) + 0.5));
            return v[i];
        };
        st->p50_ms = pct(0.50);
        st->p99_ms = pct(0.99);
        st->max_ms = v.back();
        double sum = 0;
        for (double x : v) sum += x;
        st->mean_ms = sum / v.size();
    }
    return MSBFS_OK;
}

extern "C" int msbfs_server_reset_stats(msbfs_server *s) {
    if (!s) return fail(MSBFS_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->completed = s->failed = s->launches = s->sources_done = 0;
    s->overlapped = 0;
    s->max_in_flight = s->in_flight;
    s->latencies_ms.clear();
    s->any_submit = false;
    return MSBFS_OK;
}

extern "C" int msbfs_server_stop(msbfs_server *s) {
    if (!s) return fail(MSBFS_ERR_INVALID_ARG, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->running) return MSBFS_OK;
        s->stopping = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    std::lock_guard<std::mutex> lk(s->mu);
    s->running = false;
    // queries never admitted fail
    for (Lane &L : s->lanes) {
        for (uint64_t id : L.pending) {
            Query &q = s->queries[id];
            q.state = 3;
            q.status = MSBFS_ERR_BUSY;
            q.error = "server stopped";
        }
        L.pending.clear();
    }
    s->cv_done.notify_all();
    return MSBFS_OK;
}

extern "C" int msbfs_server_destroy(msbfs_server *s) {
    if (!s) return MSBFS_OK;
    msbfs_server_stop(s);
    (void)hipSetDevice(s->ctx->device);
    for (Lane &L : s->lanes) {
        for (Slot &sl : L.slots) {
            if (sl.stream) (void)hipStreamSynchronize(sl.stream);
            if (sl.done) (void)hipEventDestroy(sl.done);
            if (sl.stream) (void)hipStreamDestroy(sl.stream);
            if (sl.eng) engine_release(sl.eng);
        }
        if (L.g) graph_release(const_cast<msbfs_graph *>(L.g));
    }
    s->lanes.clear();
    s->pool_buf.release();
    msbfs_context *c = s->ctx;
    delete s;
    context_release(c);
    return MSBFS_OK;
}
