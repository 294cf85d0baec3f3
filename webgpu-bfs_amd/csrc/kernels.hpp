// Device-side data structures and launch wrappers shared by the kernels
// (kernels.hip) and the host engine (engine.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace msbfs {

// ---------------------------------------------------------------- deep family
// One task = one group of up to B sources of one graph.  A workgroup keeps
// the task's visited bits (B bits per vertex) in LDS and runs every BFS
// level of the task with workgroup barriers only.
struct DeepTask {
    uint32_t graph;    // index into the graph table
    uint32_t nsrc;     // sources in this task (<= B)
    uint64_t src_off;  // index of the task's first source in `sources`
    uint64_t out_off;  // element offset of the task's first output row
    uint64_t reach_bit;  // global source index of bit 0 (reachability output)
};

struct DeepGraph {
    const uint32_t *ro;
    const uint32_t *ci;
    const uint32_t *ell;   // V * ell_width neighbour slots (0xFFFFFFFF = empty) or null
    uint32_t V;
    uint32_t ell_width;    // 0 (use CSR), 4 or 8
    // Lattice adjacency (graphs whose every edge joins vertices whose ids
    // differ by one of at most four offsets): LAT bits per vertex, bit k of
    // u set when the edge (u, u + lat_d[k]) exists; the edge (u, u - lat_d[k])
    // is bit k of vertex u - lat_d[k].  Unused offsets are 0.  Null when the
    // graph is not such a lattice.
    // The fields are stored after lat_pad zero fields (lat_pad >= every
    // offset), so vertex u's field is field lat_pad + u and a lower
    // neighbour's lookup needs no bounds test.
    const uint32_t *lat;
    uint32_t lat_d[4];
    uint32_t lat_pad;
};

struct DeepParams {
    const DeepTask *tasks;
    const DeepGraph *graphs;
    const uint32_t *sources;
    void *out;             // u32 or u16 distances, or null
    void *queues;          // per block: 2 * (cap - lds_qcap) spill entries (global)
    uint64_t cap;          // entries per queue (LDS part + global spill)
    uint32_t lds_qcap;     // entries per queue held in LDS
    uint32_t lds_vis_words;  // LDS words reserved for visited bits
    uint32_t *work;        // [0]: next task; zeroed before launch
    uint32_t *flags;       // see Flag* below
    uint32_t *reach;       // unused by the deep kernels (see reach_t)
    uint32_t reach_wpv;
    uint32_t *reach_t;     // reachability, source-major: a row of reach_vw words per source, or null
    uint32_t reach_vw;     // words per row: ceil(|V| / 32)
    uint32_t ntasks;
    uint32_t out_kind;      // output elements: 0 u32, 1 u16, 2 u8
    // u32 rows written through a binned log (one-wave kernel, one source per
    // task): vertices are binned by v >> log_bw; 0 = direct stores.
    uint32_t log_bw;
    uint32_t lds_bin_off;   // LDS word offset of the per-bin log counters
    uint32_t *log_sink;     // 64 words per workgroup: stores of lanes without a log entry
    uint32_t lds_sink_off;  // one-wave kernel: LDS word offset of its 64 sink entries
    // u32 / u16 rows staged in LDS as 4-bit level codes after the visited
    // bits (workgroup kernel, one source per task); lds_vis_words covers both
    uint32_t staged;
    // Logged rows of the tasks that finish last are rewritten by a separate
    // pass: the (log_defer_after + 1)-th task to finish and later ones list
    // themselves in log_defer[0 .. ntasks - log_defer_after) and their bins'
    // entry counts after that list (nbins per listed task); work[1] counts
    // finished tasks, work[2] listed ones.
    uint32_t log_defer_after;
    uint32_t *log_defer;
    // Lattice kernel (deep_wave_kernel with the adjacency in LDS): the
    // adjacency words come first in LDS, shared by the workgroup's waves,
    // then one region of lat_region words per wave, laid out as the one-wave
    // kernel's LDS (the offsets above are relative to it).
    uint32_t lat_words;
    uint32_t lat_region;
    uint32_t lat_waves;
};
// Vertex bins of the logged output: 4,096 vertices (16 KiB of u32 in LDS).
constexpr uint32_t kDeepLogBits = 12;

// ---------------------------------------------------------------- wide family
struct WideCounters {
    uint32_t next_count;   // vertices appended to the next frontier queue
    uint32_t cand_count;   // push: candidate vertices touched this level
    uint32_t pad0, pad1;
    unsigned long long active_bits;  // newly set (source, vertex) bits
    unsigned long long deg_sum;      // out-degree sum of the next frontier
};

// Union traversal of a batch: many graphs as one disjoint union, graph g's
// vertices the union's [base, base + V), its sources the source bits 0 ..
// S - 1 (bits are per graph: a vertex only ever gets bits of its own graph),
// its rows out[out_off + s * V + (v - base)].
struct BatchGraphMap {
    uint32_t base, V, S, pad;
    uint64_t out_off;
};

struct WideParams {
    uint32_t V, W;          // vertices, 64-bit source words
    uint64_t S, E;
    const uint32_t *ro, *ci;  // CSR (push)
    const uint32_t *co, *ri;  // CSC (pull); alias CSR when symmetric
    unsigned long long *seen, *nx;  // vertex-major [V][W]
    // Frontier words by distance, `defer + 2` slots of V*W words (defer <= 15): distance d < defer
    // keeps slot d until the final pass writes the output from them; later
    // distances alternate between slots defer and defer+1.
    unsigned long long *fr;
    uint32_t defer;
    uint32_t *qf0, *qf1, *qc, *stamp;           // [V]
    const uint32_t *sources;
    WideCounters *ctr;      // ring of 4
    uint32_t *flags;
    void *out;
    uint32_t *reach;
    uint32_t reach_wpv;
    uint32_t tick_base;
    uint32_t out_kind;      // output elements: 0 u32, 1 u16, 2 u8
    uint32_t policy;        // msbfs_direction_policy
    uint32_t alpha, beta;
    uint32_t rule;          // 0 measured frontier size / degree sum, 1 reference estimate
    uint32_t heavy_t;       // degree above which a vertex is handled by a whole wave
    const uint32_t *heavy_out;  // vertices with out-degree > heavy_t (CSR)
    const uint32_t *heavy_in;   // vertices with in-degree > heavy_t (CSC)
    uint32_t n_heavy_out, n_heavy_in;
    uint32_t *fbA, *fbB;    // frontier bitmaps (1 bit per vertex), ping-pong
    uint32_t *lv;           // per vertex: bit d set when it was in the frontier of kept level d
    const uint4 *hchunk_out;  // (vertex, first edge, end edge, 0): heavy out-lists in chunks
    const uint4 *hchunk_in;   // same for heavy in-lists (CSC)
    uint32_t n_hchunk_out, n_hchunk_in;
    const uint4 *light_in;  // (vertex, first in-edge, end, 0) of in-degree 1..heavy_t, increasing
    uint32_t n_light_in;
    uint32_t pair;  // deferred pull levels: two adjacent source words per lane (W even, TW >= 8)
    uint32_t pull_nb;       // vertices per workgroup block in pull levels
    const unsigned long long *zrow;  // >= W zero words: gather target of parents not in the frontier
    unsigned long long *act;  // ring of 3 x W words: sources whose frontier is not empty
    uint32_t *gbar;         // grid barrier words (kGridBarrierWords, zero at launch)
    uint32_t trace_cap;     // levels the trace buffer holds
    unsigned long long *trace;  // per level: start clock, dir<<32|frontier, mid clock, new bits
    struct WideCtl *ctl;    // stepped launches: level control (null for the cooperative kernel)
    // Batch union traversal (null otherwise): seed i is vertex seed_v[i]
    // with source bit seed_b[i] (else source i is sources[i] with bit i);
    // vertex v belongs to bmap[vgraph[v]], whose rows take its distances.
    // light_in / hchunk_in items then carry in .w the vertex's source count
    // (its graph's S): bits at or past it are never needed (0: no bound).
    const uint32_t *seed_v, *seed_b;
    uint32_t nseeds;
    const BatchGraphMap *bmap;
    const uint32_t *vgraph;
};

// Level control of a stepped wide traversal (one launch per phase of each
// level instead of one cooperative kernel): written by wide_ctl_kernel
// before every level, read by the phase kernels.
struct WideCtl {
    uint32_t L;             // level the phase kernels run
    uint32_t next;          // level wide_ctl_kernel decides next
    uint32_t dir;           // DirPush / DirPull of level L
    uint32_t n_front;       // frontier vertices of level L
    uint32_t done;          // no frontier left: the remaining launches return at once
    uint32_t push_levels, pull_levels, pad;
};

// Degree above which the wide family gives a vertex's edge list to a whole
// wave instead of one lane (R-MAT scale 20, ms at 2,048 / 64 sources: 32
// 5.87 / 1.45, 64 5.50 / 1.25, 128 5.51 / 1.24, 256 5.49 / 1.20; with the
// degree-sorted light list: 128 5.49 / 1.25, 256 5.17 / 1.20, 512 5.03 / 1.53).
constexpr uint32_t kHeavyDegree = 256;
// Edges per work item when a heavy vertex's list is split across waves:
// in-lists (pull; R-MAT scale 20, 2,048 sources, ms: 2048 5.99, 1024 5.66,
// 512 5.53, 256 5.57; 64 sources within 2%) and out-lists (push; shorter,
// so a hub source's edges spread over many waves).
constexpr uint32_t kHeavyChunk = 512;
constexpr uint32_t kHeavyChunkOut = 256;

// Grid barrier of the wide kernel: top counter, 8 group counters, the
// generation and the abort flag, each on its own 128-byte line.
constexpr uint32_t kGridBarrierWords = 32 * 11;

// flags[] layout (device, u32)
enum : int {
    FlagU16Overflow = 0,
    FlagQueueOverflow = 1,
    FlagLevels = 2,
    FlagPushLevels = 3,
    FlagPullLevels = 4,
    FlagAbort = 5,          // a grid barrier wait timed out
    FlagIncomplete = 6,     // stepped launch: more levels than were enqueued
    FlagCount = 8
};

// The rewrite of the listed tasks' logged rows (after launch_deep).
hipError_t launch_log_rewrite(const DeepParams &p, uint32_t nbins, hipStream_t st);
hipError_t launch_deep(int bits, int ell, const DeepParams &p, int blocks, int threads,
                       size_t lds_bytes, hipStream_t st);
hipError_t deep_max_blocks_per_cu(int bits, int ell, int threads, size_t lds_bytes, int *out);
// The lattice kernel (one source per task, the graph's lattice adjacency
// in LDS): lat = bits per vertex of the adjacency (2 or 4), a workgroup of
// p.lat_waves waves, each running its own tasks.
constexpr int kLatticeMaxWaves = 8;
hipError_t launch_deep_lattice(int lat, const DeepParams &p, int blocks, size_t lds_bytes,
                               hipStream_t st);
hipError_t deep_lattice_prepare(int lat, int waves, size_t lds_bytes);
hipError_t deep_lattice_blocks_per_cu(int lat, int waves, size_t lds_bytes, int *out);
// LDS the deep kernel declares statically (task slot, level counters, claim
// sink), reserved out of the per-workgroup budget; deep_prepare fails if the
// compiled kernel needs more.
constexpr size_t kDeepStaticLds = 1024;
hipError_t deep_prepare(int bits, int ell, int threads, size_t lds_bytes);

hipError_t launch_wide(int tile_words, const WideParams &p, int blocks, int threads, size_t lds,
                       hipStream_t st);
// Stepped wide traversal: `levels` x (control, phase A, phase B) launches on
// `st`, then one more control launch; no grid barrier, so the launches need
// no co-resident grid and may share the device with other work.
hipError_t launch_wide_steps(int tile_words, const WideParams &p, int blocks, int threads,
                             size_t lds, uint32_t levels, hipStream_t st);
hipError_t wide_max_blocks_per_cu(int tile_words, int threads, size_t lds, int *out);
hipError_t launch_wide_seed(const WideParams &p, hipStream_t st);
hipError_t launch_reach_transpose(const uint32_t *rt, uint32_t vw_n, uint64_t S, uint32_t V,
                                  uint32_t *reach, uint32_t wpv, hipStream_t st);
// Final pass of a wide traversal (distances of the kept levels, sentinels,
// reachability words, workspace clean-up); launched after launch_wide.
hipError_t launch_wide_final(const WideParams &p, int blocks, hipStream_t st);

hipError_t launch_zero_u32(uint32_t *ptr, uint64_t n, hipStream_t st);
// u16 distances (all < 255 or 0xFFFF) -> u8 with 0xFF for unreached.
hipError_t launch_pack_u16_u8(const uint16_t *in, uint8_t *out, uint64_t n, hipStream_t st);
// u16 distances (all < 2^bits - 1, or 0xFFFF) -> `bits` (4 or 12) per
// distance in groups of 8 (bits bytes per group), the sentinel all ones.
hipError_t launch_pack_u16_bits(const uint16_t *in, uint32_t *out, uint64_t n, int bits, hipStream_t st);

constexpr int kWideThreads = 512;

// Final pass of a wide traversal: tiles of kFinalVerts vertices x kFinalWords
// source words; per (vertex, word) eight nibble words (u32) and the mask of
// source bits it stores (u64) are staged in LDS (static, 20 KiB).
constexpr uint32_t kFinalVerts = 128, kFinalWords = 4;
// Dynamic LDS of the wide kernel: a pull block's new words (nb rows of W+1)
// and the workgroup's active-source words (W).
inline size_t wide_pull_lds(uint32_t nb, uint32_t W) { return ((size_t)nb * (W + 1) + W) * 8; }

}  // namespace msbfs
