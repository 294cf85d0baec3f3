/*
 * msbfs.h -- C ABI of the MI355X multi-source BFS engine (libmsbfs.so).
 *
 * This is the drop-in boundary that replaces the reference's wgpu runtime
 * (crates/gpu_bfs/src/{context,buffer,pipeline}.rs) and WGSL pipelines
 * (the WGSL files under crates/shaders/src/) underneath the `gpu_bfs` engine API.  All
 * signatures take plain pointers and sizes; device memory and streams are
 * passed as opaque `void*` (a hipStream_t may be NULL for the library's own
 * stream).  Every entry point returns an msbfs_status; the message of the
 * last failure on the calling thread is available from msbfs_last_error().
 *
 * Distances use the reference's layout and sentinels
 * (crates/core/src/lib.rs:21-39): dist[s*|V| + v], 0 at the source itself,
 * 0xFFFFFFFF (u32) or 0xFFFF (u16) when unreachable.
 */
#ifndef MSBFS_H
#define MSBFS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The acceptance rules follow the reference engine:
 * S == 0 -> OK with empty output (crates/gpu_bfs/src/engine.rs:228-231);
 * |V| == 0 with S > 0 -> MSBFS_ERR_SOURCE_OOB (engine.rs:233-238);
 * a source >= |V| -> MSBFS_ERR_SOURCE_OOB (engine.rs:240-249). */
typedef enum msbfs_status {
    MSBFS_OK = 0,
    MSBFS_ERR_INVALID_ARG = 1,
    MSBFS_ERR_SOURCE_OOB = 2,
    MSBFS_ERR_EMPTY_GRAPH = 3,
    MSBFS_ERR_OUT_TOO_SMALL = 4,
    MSBFS_ERR_HIP = 5,
    MSBFS_ERR_OOM = 6,
    MSBFS_ERR_DIAMETER_U16 = 7, /* a distance >= 0xFFFF cannot be stored as u16 */
    MSBFS_ERR_INVALID_GRAPH = 8,
    MSBFS_ERR_UNSUPPORTED = 9,
    MSBFS_ERR_NO_DEVICE = 10,
    MSBFS_ERR_BUSY = 11, /* serving: admission refused, retry later */
    MSBFS_ERR_PARSE = 12, /* Matrix Market text rejected; msbfs_mtx_last_error has the detail */
    MSBFS_ERR_DIAMETER_U8 = 13 /* a distance >= 0xFF cannot be stored as u8 */
} msbfs_status;

/* Method selector: same names and order as GpuBfsMethod
 * (crates/gpu_bfs/src/engine.rs:33-54).  Each method maps onto one of this
 * library's traversal families; all of them return identical distances. */
typedef enum msbfs_method {
    MSBFS_METHOD_NAIVE_SEQUENTIAL = 0,      /* deep family, one source per task */
    MSBFS_METHOD_BATCHED_DENSE_SPMM = 1,    /* wide family, pull only */
    MSBFS_METHOD_BIT_PARALLEL_32 = 2,       /* wide family, push only */
    MSBFS_METHOD_BIT_PARALLEL_64 = 3,       /* wide family, push only */
    MSBFS_METHOD_BIT_PARALLEL_64_PULL = 4,  /* wide family, pull only */
    MSBFS_METHOD_BOOLEAN_REACHABILITY = 5,  /* wide family, push only */
    MSBFS_METHOD_PUSH_PULL_ADAPTIVE = 6,    /* wide family, direction switch */
    MSBFS_METHOD_SPARSE_QUEUE = 7,          /* wide family, push only */
    MSBFS_METHOD_SPARSE_BIT_PARALLEL_64 = 8,/* wide family, push only */
    MSBFS_METHOD_MULTI_WORD_BIT_PARALLEL = 9,/* wide family, pull only */
    MSBFS_METHOD_INDIRECT_MACRO_BATCH = 10, /* wide family, pull only */
    MSBFS_METHOD_INDIRECT_SPARSE_QUEUE_PUSH = 11, /* wide family, direction switch */
    MSBFS_METHOD_MULTI_WORD_BIT_PARALLEL_PUSH = 12, /* wide family, push only */
    MSBFS_METHOD_WORKGROUP_PERSISTENT_BFS = 13,     /* deep family (LDS-resident) */
    MSBFS_METHOD_AUTO = 14
} msbfs_method;

/* Output element types.  Unreached pairs hold the all-ones sentinel of the type
 * (u32::MAX, 0xFFFF, 0xFF), so u16 holds distances up to 65,534 and u8 up to 254; a
 * larger distance is an error (MSBFS_ERR_DIAMETER_U16 / _U8), never a silent alias of
 * the sentinel.  u8 is the compact wire format of pack_distances_u32_to_u8.wgsl:18-53. */
typedef enum msbfs_dist_type {
    MSBFS_DIST_U32 = 0,
    MSBFS_DIST_U16 = 1,
    MSBFS_DIST_U8 = 2
} msbfs_dist_type;

/* Traversal families (what a method resolves to). */
typedef enum msbfs_family {
    MSBFS_FAMILY_NONE = 0,
    MSBFS_FAMILY_DEEP = 1, /* per-workgroup instances, visited bits in LDS */
    MSBFS_FAMILY_WIDE = 2  /* device-wide level-synchronous push/pull */
} msbfs_family;

typedef enum msbfs_direction_policy {
    MSBFS_DIR_ADAPTIVE = 0,  /* push/pull switch (alpha, beta) */
    MSBFS_DIR_PUSH = 1,
    MSBFS_DIR_PULL = 2
} msbfs_direction_policy;

typedef struct msbfs_context msbfs_context;
typedef struct msbfs_graph msbfs_graph;
typedef struct msbfs_engine msbfs_engine;
typedef struct msbfs_batch msbfs_batch;
typedef struct msbfs_server msbfs_server;

typedef struct msbfs_device_info {
    char name[64];
    char arch[32];
    int32_t compute_units;
    int32_t lds_bytes_per_block;
    uint64_t total_mem_bytes;
    int32_t cooperative_launch;
    int32_t reserved;
} msbfs_device_info;

typedef struct msbfs_graph_info {
    uint64_t num_vertices;
    uint64_t num_edges;
    uint32_t max_degree;
    uint32_t symmetric;   /* 1 when CSC aliases CSR (exact check) */
    uint64_t device_bytes;
} msbfs_graph_info;

/* Tuning knobs (the autotuner's search space).  Zero means "default". */
typedef struct msbfs_tuning {
    uint32_t alpha;          /* push->pull when m_f > |E|/alpha (push_pull.rs:22-36), default 14 */
    uint32_t beta;           /* pull->push when n_f < |V|/beta, default 24 */
    uint32_t tile_words;     /* wide family wave tile: 64/tile_words vertices x tile_words words (1,2,4,8) */
    uint32_t words_per_lane; /* wide family, pull levels with deferred distances: adjacent source
                                words per lane, 1 or 2 (two need an even word count and a tile of 8
                                or more lanes, else one is used); 0 = auto: 2 when the word count
                                is even and >= 16 and tile_words is 0 */
    uint32_t deep_bits;      /* deep family: sources per instance (1,2,4,8,16,32); 0 = auto */
    uint32_t deep_threads;   /* deep family workgroup size (a multiple of 64, <= 1024; 64 with a
                                neighbour table: the one-wave kernel) */
    uint32_t wide_blocks_per_cu; /* wide family persistent blocks per CU */
    uint32_t direction_rule; /* 0: measured frontier degree sum and size (default);
                                1: the reference's estimate n_f = ceil(bits/S), m_f = n_f*|E|/|V| */
    uint32_t trace_levels;   /* wide family: record up to this many levels (0 = off) */
    uint32_t stepped;        /* wide family launch: 1 stepped, one launch per phase of each level
                                (no co-resident grid needed, so other traversals may run beside
                                it; the serving loop's mode); 2 one cooperative kernel for all
                                levels (grid barriers, the device to itself); 0 auto: cooperative
                                for graphs known to need more than 64 levels, or, before any run,
                                graphs the reference's router calls road-like (max degree <= 8 and
                                |E| <= 4|V|), stepped otherwise */
    uint32_t deep_rows;      /* deep family workgroup kernel, one source per task, u32/u16 output:
                                1 staged, the row kept in LDS as a 4-bit code per vertex and written
                                whole when the task ends (distances >= 14 stored as found); 2 every
                                distance stored as found; 0 auto: staged when the graph's deep
                                traversals are known to need <= 32 levels, or, before the first
                                one, when it is not road-like (max degree <= 8 and |E| <= 4|V|),
                                and the codes fit in LDS.  One-wave kernel, u32 output: 2 stores
                                every distance as found instead of through the binned log (used
                                with four or more tasks per CU).  Lattice kernel, u32 output: 1
                                writes through the binned log, else every distance is stored as
                                found */
    uint32_t deep_lattice;   /* deep family, one source per task, graphs whose every edge joins ids
                                that differ by one of at most four offsets (grids, grid-derived
                                road models): the lattice kernel keeps the whole adjacency in LDS
                                (2 or 4 bits per vertex) shared by a workgroup's waves, one task per
                                wave.  0 auto: in place of the one-wave kernel; 1 for every
                                one-source-per-task run on such a graph; 2 never */
    uint32_t batch_union;    /* batches: union traversals -- graphs grouped by source words per
                                vertex (the next power of two of ceil(S/64)), each group's graphs
                                joined into one disjoint union whose sources share the source bits
                                (bit i is source i of every graph: bits never cross graphs), each
                                union traversed by the wide family, bit-parallel -- against one
                                persistent deep-family launch over per-source tasks.  1 every graph
                                in a union; 2 every graph in the deep launch; 0 graphs of at least
                                union_words source words in unions, the rest in the deep launch */
    uint32_t union_words;    /* with batch_union 0: the fewest source words per vertex (1, 2, 4, ...)
                                that puts a graph into a union; 0 auto: 2 (512 G(n,p) graphs of
                                1K-64K vertices, ms per traversal of the graphs of 1, 2 and 4 words:
                                unions 5.9, 1.8 and 2.0 against about 2.5, 2.1 and 3.6 in the deep
                                launch) */
    uint32_t deep_blocks_per_cu; /* deep family persistent launches (one graph's and a batch's): at
                                most this many workgroups per CU (<= 32); 0 auto: as many as fit,
                                except a batch whose union traversals run beside its deep launch,
                                which takes 1 and leaves the unions the rest of each CU (512 G(n,p)
                                graphs: 5.58 ms per replay against 6.51 with three) */
} msbfs_tuning;

/* One BFS level of the last wide-family run (recorded when trace_levels > 0). */
typedef struct msbfs_level_record {
    uint32_t level;
    uint32_t direction;      /* 0 push, 1 pull */
    uint32_t frontier;       /* vertices in the level's frontier */
    uint32_t reserved;
    uint64_t new_bits;       /* (source, vertex) pairs first reached by the level */
    double start_us;         /* device clock, relative to level 0 */
    double mid_us;           /* after expand (push) or gather (pull) */
    double end_us;
} msbfs_level_record;

typedef struct msbfs_run_stats {
    int32_t method;          /* method actually run (after Auto) */
    int32_t family;          /* msbfs_family */
    uint32_t levels;         /* BFS levels executed */
    uint32_t push_levels;
    uint32_t pull_levels;
    uint32_t instances;      /* deep family instances */
    uint32_t bits_per_instance;
    uint32_t grid_blocks;
    uint32_t block_threads;  /* workgroup size of the traversal kernel (deep family: 64 is the
                                one-wave kernel) */
    uint32_t lattice_bits;   /* deep family: bits per vertex of the LDS lattice adjacency the run
                                used (0: none) */
    float kernel_ms;         /* traversal kernel(s) device time, hipEvents */
    float total_ms;          /* whole call device time, hipEvents */
    uint64_t workspace_bytes;
} msbfs_run_stats;

/* ---- errors / version ---- */
const char *msbfs_last_error(void);
const char *msbfs_version(void);
const char *msbfs_status_string(int status);

/* ---- device context: GpuContext::init_blocking
 * (crates/gpu_bfs/src/context.rs:174-388) ---- */
int msbfs_context_create(int device, msbfs_context **out);
int msbfs_context_destroy(msbfs_context *ctx);
int msbfs_context_info(const msbfs_context *ctx, msbfs_device_info *info);
int msbfs_context_synchronize(msbfs_context *ctx);

/* ---- graph residency: replaces PushPullGraphBuffers
 * (crates/gpu_bfs/src/buffer.rs:96-219).  The CSR is validated like
 * CsrGraph::validate (crates/core/src/graph.rs:226-278, except that neighbour
 * lists need not be sorted), uploaded once, and its transpose is built on
 * the host (graph.rs:90-126) unless the graph is exactly symmetric, in which
 * case the CSC aliases the CSR.  Handles are explicit: nothing is keyed on
 * (|V|,|E|). ---- */
int msbfs_graph_create(msbfs_context *ctx, uint64_t num_vertices, uint64_t num_edges,
                       const uint32_t *row_offsets, const uint32_t *col_indices,
                       msbfs_graph **out);
int msbfs_graph_destroy(msbfs_graph *g);
int msbfs_graph_get_info(const msbfs_graph *g, msbfs_graph_info *info);

/* ---- engine: GpuBfsEngine::new(context, method)
 * (crates/gpu_bfs/src/engine.rs:68) ---- */
int msbfs_engine_create(msbfs_context *ctx, int method, msbfs_engine **out);
int msbfs_engine_destroy(msbfs_engine *e);
int msbfs_engine_method(const msbfs_engine *e);
/* PushPullConfig { alpha, beta } (crates/gpu_bfs/src/dispatch/push_pull.rs:22-36) */
int msbfs_engine_set_push_pull(msbfs_engine *e, uint32_t alpha, uint32_t beta);
/* Checks a tuning record without applying it (MSBFS_ERR_INVALID_ARG with a
 * message for a field out of range).  Every set_tuning entry point (engine,
 * batch, server lane) applies the same checks.  No reference counterpart:
 * the reference's knobs are compile-time shader constants. */
int msbfs_tuning_validate(const msbfs_tuning *t);
int msbfs_engine_set_tuning(msbfs_engine *e, const msbfs_tuning *t);
int msbfs_engine_get_tuning(const msbfs_engine *e, msbfs_tuning *t);
int msbfs_engine_last_stats(msbfs_engine *e, msbfs_run_stats *st);
/* Per-level trace of the last wide-family run: writes up to max_levels
 * records and the number recorded.  The reference has no such facility (its
 * timing is wall clock around whole calls, crates/cli/src/bench.rs:338-370). */
int msbfs_engine_level_trace(msbfs_engine *e, msbfs_level_record *out, uint32_t max_levels,
                             uint32_t *n_levels);

/* Measured routing (the topology-aware replacement of the reference's fixed
 * Auto table, engine.rs:157-198): runs candidate configurations on this graph
 * and these sources and keeps the fastest in the engine.  With method Auto both
 * families compete and the engine's method becomes the winner's
 * (WorkgroupPersistentBfs or PushPullAdaptive); an explicit method is tuned
 * within its family.  Deep candidates: every sources-per-task width whose bits
 * fit LDS x workgroup sizes 64 (one wave per task)/128/256/512/1024.  Wide candidates, one knob at a time
 * from the defaults: tile width (auto, half, double), blocks per CU (1, 2),
 * direction rule (measured, reference), alpha (7, 14, 28), beta (12, 24, 48).
 * Each candidate runs once untimed and twice timed (kernel time).  Output goes to
 * out_device (S*|V| elements of dist_type) or, when it is NULL, to an internal
 * buffer; all candidates produce the same distances.  Writes the winner's
 * tuning, method and mean kernel ms. */
int msbfs_engine_autotune(msbfs_engine *e, const msbfs_graph *g, const uint32_t *sources,
                          uint64_t num_sources, int dist_type, void *out_device,
                          uint64_t out_len, msbfs_tuning *best, int *best_method,
                          float *best_ms);

/* GpuBfsEngine::select_auto_method (crates/gpu_bfs/src/engine.rs:157-198):
 * the reference's decision table, returned as an msbfs_method. */
int msbfs_select_auto_method(uint64_t num_vertices, uint64_t num_edges, uint32_t max_degree,
                             uint64_t num_sources);
/* GpuBfsEngine::resolve_method (engine.rs:201-206) followed by this
 * library's family choice; returns an msbfs_family. */
int msbfs_engine_resolve(const msbfs_engine *e, const msbfs_graph *g, uint64_t num_sources,
                         int *method_out);

/* ---- traversal ----
 * BfsEngine::run_ms_bfs / GpuBfsEngine::run_ms_bfs_into
 * (crates/core/src/lib.rs:21-39, crates/gpu_bfs/src/engine.rs:219-297).
 * `out` is host memory of out_len elements (u32, u16 or u8 by dist_type);
 * out_len must be >= S*|V|.  The distances cross PCIe in the most compact
 * type that holds them (the reference's u8/u16 packing and host unpack,
 * indirect_sparse_push.rs:1580-1967, 2922-3017): computed on the device as u16
 * (u32 only if a distance reaches 65,535), packed to u8 on the device when
 * every distance is below 255, and widened into `out` by host threads while
 * later chunks are still being copied. */
int msbfs_run_ms_bfs(msbfs_engine *e, const msbfs_graph *g, const uint32_t *sources,
                     uint64_t num_sources, int dist_type, void *out, uint64_t out_len);
/* Same, with the output left in device memory `out_device` and work
 * enqueued on `stream` (NULL = library stream).  Returns after enqueueing
 * when `sync` is 0; errors found by the kernels (u16 overflow) are then
 * reported by msbfs_engine_check(). */
int msbfs_run_ms_bfs_device(msbfs_engine *e, const msbfs_graph *g, const uint32_t *sources,
                            uint64_t num_sources, int dist_type, void *out_device,
                            uint64_t out_len, void *stream, int sync);
int msbfs_engine_check(msbfs_engine *e);

/* No-allocation variant of msbfs_run_ms_bfs_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device or pinned memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as msbfs_run_ms_bfs_device with
 * sync = 1 (host sources, device output, synchronous), for source sets the
 * deep family runs; every workspace lives in a static per-device array of
 * MSBFS_NOALLOC_SCRATCH_BYTES, set up by msbfs_context_create.  Rejects
 * with a message: more than MSBFS_NOALLOC_MAX_SOURCES sources
 * (MSBFS_ERR_OUT_TOO_SMALL), graphs and source counts that route to the
 * wide family (MSBFS_ERR_UNSUPPORTED), and a frontier level larger than the
 * array's queue space (MSBFS_ERR_OUT_TOO_SMALL, after the traversal).  Calls
 * on one GPU share its array: calls that overlap on one GPU race. */
#define MSBFS_NOALLOC_SCRATCH_BYTES (64u << 20)
#define MSBFS_NOALLOC_MAX_SOURCES 65536u
int msbfs_run_ms_bfs_device_noalloc(msbfs_engine *e, const msbfs_graph *g,
                                    const uint32_t *sources, uint64_t num_sources, int dist_type,
                                    void *out_device, uint64_t out_len, void *stream);

/* GpuBfsEngine::run_ms_reachability (crates/gpu_bfs/src/engine.rs:420-479):
 * vertex-major bitmask words[v*ceil(S/32) + s/32], bit s%32
 * (ReachabilityBitmask, crates/core/src/graph.rs:446-646). */
int msbfs_run_ms_reachability(msbfs_engine *e, const msbfs_graph *g, const uint32_t *sources,
                              uint64_t num_sources, uint32_t *words, uint64_t words_len);

/* ---- batched launch ----
 * Many resident graphs, each with its own source list, traversed by one
 * persistent launch.  No counterpart in the reference, which makes one
 * engine call per graph (crates/gpu_bfs/src/engine.rs:219-297).  Sources
 * are concatenated graph by graph (source_counts[g] entries for graph g);
 * graph g's distances occupy the output elements
 * [offsets[g], offsets[g] + source_counts[g]*|V_g|) in the usual
 * dist[s*|V_g| + v] layout.  A graph whose visited bits (one per vertex) do
 * not fit one workgroup's LDS is traversed by the wide family instead, one
 * engine run per such graph on the same stream after the persistent launch
 * (outside a captured hipGraph: a replay launches them eagerly). */
int msbfs_batch_create(msbfs_engine *e, uint32_t num_graphs, msbfs_graph *const *graphs,
                       const uint64_t *source_counts, const uint32_t *sources, int dist_type,
                       msbfs_batch **out);
int msbfs_batch_destroy(msbfs_batch *b);
int msbfs_batch_layout(const msbfs_batch *b, uint64_t *total_elements, uint64_t *offsets);
/* Replaces the sources (same counts); a captured graph sees the new ones. */
int msbfs_batch_set_sources(msbfs_batch *b, const uint32_t *sources);
int msbfs_batch_run(msbfs_batch *b, void *out_device, uint64_t out_len, void *stream, int sync);
/* Records one run (output pointer fixed) into a hipGraph; replay launches it. */
int msbfs_batch_capture(msbfs_batch *b, void *out_device, uint64_t out_len, void *stream);
int msbfs_batch_replay(msbfs_batch *b, void *stream, int sync);
int msbfs_batch_check(msbfs_batch *b);
int msbfs_batch_stats(msbfs_batch *b, msbfs_run_stats *st);
int msbfs_batch_set_tuning(msbfs_batch *b, const msbfs_tuning *t);
/* The batch's current tuning (after msbfs_batch_autotune: the winner's). */
int msbfs_batch_get_tuning(const msbfs_batch *b, msbfs_tuning *t);
/* Family of each graph of the batch (msbfs_family), num_graphs entries. */
int msbfs_batch_families(const msbfs_batch *b, int32_t *families);
/* Autotuner: times on the batch itself every feasible deep_bits x
 * deep_threads (64..1024) pair, staged and direct rows for one source per
 * task, and, when some graph takes the wide family, its wave tile
 * (tile_words), words per lane, push/pull thresholds (alpha, beta) and
 * launch form, one knob at a time from the best so far; keeps the fastest
 * and returns it and its time. */
int msbfs_batch_autotune(msbfs_batch *b, void *out_device, uint64_t out_len, msbfs_tuning *best,
                         float *best_ms);
/* One candidate of the last autotune. */
typedef struct msbfs_sweep_entry {
    msbfs_tuning tuning;
    float ms;        /* mean of two timed runs after one warm-up; < 0 when it did not run */
    int32_t status;  /* MSBFS_OK, or the error the candidate's plan or run returned */
} msbfs_sweep_entry;
/* Copies up to `cap` entries of the last autotune's sweep (in the order they
 * ran) and sets *n to the number there are. */
int msbfs_batch_sweep(const msbfs_batch *b, msbfs_sweep_entry *entries, uint32_t cap, uint32_t *n);

/* ---- serving ----
 * Resident graphs ("lanes") and a scheduler thread that packs queries of
 * 1..512 sources into shared 64-bit source words, admitting them by the
 * device memory their distance outputs need.  Continuous batching: each
 * lane has several launch slots (engine + stream); a query arriving while
 * the lane's traversals run is launched as soon as a slot and output memory
 * are free, without waiting for the running traversals, and the traversals
 * of every lane and slot run side by side on the device (wide-family lanes
 * use stepped launches, which need no co-resident grid).  Queries that
 * arrive while every slot is busy are packed into the next launch.  Results
 * stay in device memory until released.  No counterpart in the reference
 * (its engine serves one call at a time, crates/gpu_bfs/src/engine.rs:61-63). */
typedef struct msbfs_server_config {
    uint64_t output_pool_bytes;       /* device memory for results; 0 = 16 GiB */
    uint32_t max_sources_per_launch;  /* 0 = 2048 */
    uint32_t max_queries_per_launch;  /* 0 = 512 */
    int32_t dist_type;                /* MSBFS_DIST_U32 or MSBFS_DIST_U16 */
    int32_t pack;                     /* 1: many queries per traversal; 0: one per traversal */
    uint32_t slots_per_lane;          /* traversals of one lane in flight at once; 0 = 4 */
    uint32_t admit_wait_us;           /* beside a running traversal of the same lane, a launch
                                         waits for half of max_sources_per_launch or until its
                                         oldest query has waited this long; 0 = 2000 */
} msbfs_server_config;

typedef struct msbfs_query_info {
    int32_t state;        /* 0 pending, 1 running, 2 done, 3 failed */
    int32_t status;       /* msbfs_status of a failed query */
    uint32_t lane;
    uint32_t num_sources;
    uint64_t launch;      /* sequence number of the traversal that ran it */
    void *device_ptr;     /* dist[s*|V| + v] for its k sources, while held */
    uint64_t elements;
    double queue_ms;      /* submit -> launch */
    double latency_ms;    /* submit -> result ready */
} msbfs_query_info;

typedef struct msbfs_server_stats {
    uint64_t completed, failed, launches, sources, pending;
    uint64_t pool_bytes, pool_used;
    double span_s;        /* first submit -> last completion */
    double p50_ms, p99_ms, mean_ms, max_ms;
    uint64_t overlapped_launches; /* launches started while another traversal was in flight */
    uint64_t max_in_flight;       /* most traversals in flight at once (all lanes) */
} msbfs_server_stats;

int msbfs_server_create(msbfs_context *ctx, const msbfs_server_config *cfg, msbfs_server **out);
int msbfs_server_add_graph(msbfs_server *s, msbfs_graph *g, int method, uint32_t *lane_id);
int msbfs_server_set_tuning(msbfs_server *s, uint32_t lane, const msbfs_tuning *t);
int msbfs_server_start(msbfs_server *s);
int msbfs_server_submit(msbfs_server *s, uint32_t lane, const uint32_t *sources, uint32_t k,
                        uint64_t *query_id);
int msbfs_server_wait(msbfs_server *s, uint64_t query_id, double timeout_s);
/* Completion queue: the oldest finished (done or failed) query not yet
 * returned by this call.  Clients that hold results should consume them in
 * this order and release them, or the output pool can fill with results
 * while an older query waits for admission. */
int msbfs_server_next_completed(msbfs_server *s, double timeout_s, uint64_t *query_id);
int msbfs_server_query_info(msbfs_server *s, uint64_t query_id, msbfs_query_info *info);
int msbfs_server_copy_result(msbfs_server *s, uint64_t query_id, void *host_out, uint64_t out_len);
int msbfs_server_release(msbfs_server *s, uint64_t query_id);
int msbfs_server_get_stats(msbfs_server *s, msbfs_server_stats *st);
int msbfs_server_reset_stats(msbfs_server *s);
int msbfs_server_stop(msbfs_server *s);
int msbfs_server_destroy(msbfs_server *s);

/* ---- host-side graph utilities (crates/core) ----
 * Library-owned host CSR produced by the generators and converters below;
 * read it with msbfs_host_graph_get and free it with msbfs_host_graph_free. */
typedef struct msbfs_host_graph msbfs_host_graph;
int msbfs_host_graph_get(const msbfs_host_graph *h, uint64_t *num_vertices, uint64_t *num_edges,
                         const uint32_t **row_offsets, const uint32_t **col_indices);
void msbfs_host_graph_free(msbfs_host_graph *h);

/* CsrGraph::from_edge_list (crates/core/src/graph.rs:51-80). */
int msbfs_csr_from_edge_list(uint64_t num_vertices, uint64_t num_pairs, const uint32_t *us,
                             const uint32_t *vs, msbfs_host_graph **out);
/* csr_from_undirected_pairs (crates/core/src/generators.rs:547-556). */
int msbfs_csr_from_undirected_pairs(uint64_t num_vertices, uint64_t num_pairs,
                                    const uint32_t *us, const uint32_t *vs,
                                    msbfs_host_graph **out);
/* CsrGraph::to_csc (graph.rs:90-126): output arrays of V+1 and E entries. */
int msbfs_csr_to_csc(uint64_t num_vertices, uint64_t num_edges, const uint32_t *row_offsets,
                     const uint32_t *col_indices, uint32_t *col_offsets, uint32_t *row_indices);
/* CsrGraph::is_symmetric (graph.rs:129-132): 1 / 0, or a negative status. */
int msbfs_csr_is_symmetric(uint64_t num_vertices, uint64_t num_edges,
                           const uint32_t *row_offsets, const uint32_t *col_indices);
/* Lattice test of the deep family's LDS adjacency (no counterpart in the reference): for a
 * symmetric graph whose every edge joins ids that differ by one of at most four offsets, writes
 * them ascending to offsets[0..k) (the rest 0) and returns k >= 1; returns 0 for any other graph
 * (offsets all 0), or a negative status. */
int msbfs_csr_lattice_offsets(uint64_t num_vertices, uint64_t num_edges,
                              const uint32_t *row_offsets, const uint32_t *col_indices,
                              uint32_t *offsets);
/* CsrGraph::validate (graph.rs:226-278): MSBFS_OK or MSBFS_ERR_INVALID_GRAPH. */
int msbfs_csr_validate(uint64_t num_vertices, uint64_t num_edges, const uint32_t *row_offsets,
                       const uint32_t *col_indices);

/* ---- Matrix Market reader (crates/core/src/mtx.rs:124-287, CsrGraph::from_mtx_reader) ----
 * Coordinate files only; values ignored; self loops dropped; symmetric, skew-symmetric and
 * hermitian files mirrored; |V| = max(rows, cols); result is the canonical CSR of
 * CsrGraph::from_edge_list (graph.rs:51-80).  On rejected text the call returns
 * MSBFS_ERR_PARSE, msbfs_last_error() holds the reference's message (error.rs:45-81) and
 * msbfs_mtx_last_error() the variant and its fields. */
typedef enum msbfs_mtx_error_kind {
    MSBFS_MTX_OK = 0,
    MSBFS_MTX_IO = 1,                    /* MtxError::Io */
    MSBFS_MTX_INVALID_HEADER = 2,        /* MtxError::InvalidHeader */
    MSBFS_MTX_UNSUPPORTED_FORMAT = 3,    /* MtxError::UnsupportedFormat ("array") */
    MSBFS_MTX_INVALID_DIMENSIONS = 4,    /* MtxError::InvalidDimensions */
    MSBFS_MTX_INVALID_COORDINATE_LINE = 5, /* MtxError::InvalidCoordinateLine */
    MSBFS_MTX_ZERO_INDEX = 6,            /* MtxError::ZeroIndex */
    MSBFS_MTX_INDEX_OUT_OF_BOUNDS = 7    /* MtxError::IndexOutOfBounds */
} msbfs_mtx_error_kind;

typedef struct msbfs_mtx_error {
    int32_t kind;          /* msbfs_mtx_error_kind */
    uint32_t reserved;
    uint64_t line_number;  /* 1-based line of the file */
    uint64_t row, col;     /* ZeroIndex / IndexOutOfBounds */
    uint64_t max_rows, max_cols; /* IndexOutOfBounds */
} msbfs_mtx_error;

/* Parses `len` bytes of Matrix Market text (MtxReader::parse). */
int msbfs_mtx_parse(const char *text, uint64_t len, msbfs_host_graph **out);
/* Reads and parses a file (CsrGraph::from_mtx_reader over a BufReader). */
int msbfs_mtx_read(const char *path, msbfs_host_graph **out);
/* Detail of this thread's last failed msbfs_mtx_parse / msbfs_mtx_read. */
int msbfs_mtx_last_error(msbfs_mtx_error *out);

/* Generators.  Grid2D follows crates/core/src/generators.rs:28-80 exactly.
 * The random generators restate the reference's algorithms
 * (R-MAT generators.rs:277-342, G(n,p) Batagelj-Brandes generators.rs:139-191)
 * over this library's own seeded xoshiro256** streams, so their outputs are
 * reproducible here but are not the reference's rand 0.8 streams. */
int msbfs_gen_grid2d(uint64_t width, uint64_t height, msbfs_host_graph **out);
/* Road-like: `width` columns, num_vertices vertices laid out row-major on
 * the grid (last row partial); a seeded Kruskal spanning tree over the grid
 * edges plus `extra_edges` distinct non-tree grid edges. */
int msbfs_gen_road_like(uint64_t num_vertices, uint64_t width, uint64_t extra_edges,
                        uint64_t seed, msbfs_host_graph **out);
int msbfs_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, double d,
                   uint64_t seed, int directed, msbfs_host_graph **out);
int msbfs_gen_erdos_renyi_np(uint64_t num_vertices, double p, uint64_t seed,
                             msbfs_host_graph **out);
/* Uniform sources without replacement, sorted ascending; when
 * non_isolated != 0 only vertices of out-degree > 0 are drawn. */
int msbfs_sample_sources(const uint32_t *row_offsets, uint64_t num_vertices, uint64_t count,
                         uint64_t seed, int non_isolated, uint32_t *out);

#ifdef __cplusplus
}
#endif

#endif /* MSBFS_H */
