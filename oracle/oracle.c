/*
 * oracle.c -- CPU restatement of the reference's multi-source BFS ground truth.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: only
 * the test suite, the smoke check and the CPU-baseline leg of the benchmark
 * load it.  Nothing in the shipped library calls, links or executes it.
 *
 * Every function restates one function of the reference (likr-sandbox/webgpu-bfs,
 * Rust) and cites the file:line it follows.  All arithmetic is integer.
 *
 *   orc_csr_from_edge_list      crates/core/src/graph.rs:51-80
 *   orc_csr_to_csc              crates/core/src/graph.rs:90-126
 *   orc_bfs_single_source_into  crates/cpu_bfs/src/fifo.rs:10-46
 *   orc_ms_bfs_fifo             crates/cpu_bfs/src/fifo.rs:70-102
 *   orc_ms_bfs_parallel         crates/cpu_bfs/src/rayon_ms.rs:11-54
 *   orc_ms_bfs_bit_parallel     crates/cpu_bfs/src/bit_parallel.rs:11-130
 *   orc_traversed_edges         crates/cli/src/bench.rs:478-491
 *   orc_reachability_from_dist  crates/core/src/graph.rs:571-645
 *   orc_grid_manhattan          crates/core/src/generators.rs:98-127
 *   orc_pack_u16 / orc_unpack_u16
 *                               crates/shaders/src/pack_distances_u32_to_u16.wgsl:15-38,
 *                               crates/gpu_bfs/src/dispatch/indirect_sparse_push.rs:2922-2938
 *   orc_pack_u8 / orc_unpack_u8 crates/shaders/src/pack_distances_u32_to_u8.wgsl:18-53,
 *                               crates/gpu_bfs/src/dispatch/indirect_sparse_push.rs:2982-3017
 *
 * Status codes: 0 ok, 2 source out of bounds (the reference's
 * "Source vertex {s} at index {i} exceeds graph vertex count {V}" bail).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <stdatomic.h>

#define ORC_OK 0
#define ORC_ERR_SOURCE_OOB 2
#define ORC_INF 0xFFFFFFFFu

/* ---- CSR construction (graph.rs:51-80): drop self loops and out-of-range
 * endpoints, sort by (u, v), dedup, prefix-sum the row counts. ---- */
static int cmp_pair(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* Returns the number of kept edges; row_offsets has V+1 entries and cols
 * has room for n entries. */
size_t orc_csr_from_edge_list(size_t V, size_t n, const uint32_t *us, const uint32_t *vs,
                              uint32_t *row_offsets, uint32_t *cols) {
    uint64_t *keys = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    size_t m = 0;
    for (size_t i = 0; i < n; ++i) {
        uint32_t u = us[i], v = vs[i];
        if (u != v && (size_t)u < V && (size_t)v < V)
            keys[m++] = ((uint64_t)u << 32) | v;
    }
    qsort(keys, m, sizeof(uint64_t), cmp_pair);
    size_t k = 0;
    for (size_t i = 0; i < m; ++i)
        if (k == 0 || keys[i] != keys[k - 1]) keys[k++] = keys[i];
    memset(row_offsets, 0, (V + 1) * sizeof(uint32_t));
    for (size_t i = 0; i < k; ++i) row_offsets[(keys[i] >> 32) + 1] += 1;
    for (size_t i = 0; i < V; ++i) row_offsets[i + 1] += row_offsets[i];
    for (size_t i = 0; i < k; ++i) cols[i] = (uint32_t)(keys[i] & 0xFFFFFFFFu);
    free(keys);
    return k;
}

/* ---- Transpose by counting sort (graph.rs:90-126); predecessor lists come
 * out sorted because rows are scanned in increasing u. ---- */
void orc_csr_to_csc(size_t V, size_t E, const uint32_t *ro, const uint32_t *ci,
                    uint32_t *col_offsets, uint32_t *row_indices) {
    uint32_t *indeg = (uint32_t *)calloc(V ? V : 1, sizeof(uint32_t));
    for (size_t e = 0; e < E; ++e) indeg[ci[e]] += 1;
    col_offsets[0] = 0;
    for (size_t i = 0; i < V; ++i) col_offsets[i + 1] = col_offsets[i] + indeg[i];
    uint32_t *pos = indeg; /* reuse */
    for (size_t i = 0; i < V; ++i) pos[i] = col_offsets[i];
    for (size_t u = 0; u < V; ++u)
        for (uint32_t e = ro[u]; e < ro[u + 1]; ++e) row_indices[pos[ci[e]]++] = (uint32_t)u;
    free(indeg);
}

/* ---- Single-source FIFO BFS (fifo.rs:10-46).  dist must hold V entries
 * pre-filled with ORC_INF; queue must hold V entries. ---- */
void orc_bfs_single_source_into(size_t V, const uint32_t *ro, const uint32_t *ci,
                                uint32_t source, uint32_t *dist, uint32_t *queue) {
    if (V == 0) return;
    dist[source] = 0;
    size_t head = 0, tail = 0;
    queue[tail++] = source;
    while (head < tail) {
        uint32_t u = queue[head++];
        uint32_t nd = dist[u] + 1;
        for (uint32_t e = ro[u]; e < ro[u + 1]; ++e) {
            uint32_t v = ci[e];
            if (dist[v] == ORC_INF) {
                dist[v] = nd;
                queue[tail++] = v;
            }
        }
    }
}

static int validate_sources(size_t V, const uint32_t *sources, size_t S, size_t *bad_index) {
    for (size_t i = 0; i < S; ++i)
        if ((size_t)sources[i] >= V) {
            if (bad_index) *bad_index = i;
            return ORC_ERR_SOURCE_OOB;
        }
    return ORC_OK;
}

/* ---- Sequential multi-source FIFO (fifo.rs:70-102): dist[s*V+v]. ---- */
int orc_ms_bfs_fifo(size_t V, const uint32_t *ro, const uint32_t *ci,
                    const uint32_t *sources, size_t S, uint32_t *dist) {
    if (S == 0 || V == 0) return ORC_OK;
    if (validate_sources(V, sources, S, NULL)) return ORC_ERR_SOURCE_OOB;
    for (size_t i = 0; i < S * V; ++i) dist[i] = ORC_INF;
    uint32_t *queue = (uint32_t *)malloc(V * sizeof(uint32_t));
    for (size_t s = 0; s < S; ++s)
        orc_bfs_single_source_into(V, ro, ci, sources[s], dist + s * V, queue);
    free(queue);
    return ORC_OK;
}

/* ---- Parallel over sources (rayon_ms.rs:11-54): the output is pre-filled
 * with ORC_INF, each task owns one source row and its own queue.  Tasks are
 * handed out dynamically from an atomic counter (Rayon's work stealing gives
 * the same disjoint-row schedule). ---- */
typedef struct {
    size_t V;
    const uint32_t *ro, *ci, *sources;
    size_t S;
    uint32_t *dist;
    atomic_size_t next;
    int fill_phase; /* 1: pre-fill rows, 0: traverse */
} par_job;

static void *par_worker(void *arg) {
    par_job *j = (par_job *)arg;
    uint32_t *queue = j->fill_phase ? NULL : (uint32_t *)malloc(j->V * sizeof(uint32_t));
    for (;;) {
        size_t s = atomic_fetch_add(&j->next, 1);
        if (s >= j->S) break;
        uint32_t *row = j->dist + s * j->V;
        if (j->fill_phase) {
            for (size_t v = 0; v < j->V; ++v) row[v] = ORC_INF;
        } else {
            orc_bfs_single_source_into(j->V, j->ro, j->ci, j->sources[s], row, queue);
        }
    }
    free(queue);
    return NULL;
}

int orc_ms_bfs_parallel(size_t V, const uint32_t *ro, const uint32_t *ci,
                        const uint32_t *sources, size_t S, uint32_t *dist, int nthreads) {
    if (S == 0 || V == 0) return ORC_OK;
    if (validate_sources(V, sources, S, NULL)) return ORC_ERR_SOURCE_OOB;
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc((size_t)nthreads * sizeof(pthread_t));
    par_job job;
    job.V = V; job.ro = ro; job.ci = ci; job.sources = sources; job.S = S; job.dist = dist;
    for (int phase = 1; phase >= 0; --phase) {
        job.fill_phase = phase;
        atomic_store(&job.next, 0);
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, par_worker, &job);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    free(th);
    return ORC_OK;
}

/* ---- 64-source bit-parallel batch (bit_parallel.rs:11-97). ---- */
static void bit_parallel_batch_64(size_t V, const uint32_t *ro, const uint32_t *ci,
                                  const uint32_t *batch, size_t bs, size_t batch_start,
                                  uint32_t *dist, uint64_t *seen, uint64_t *frontier,
                                  uint64_t *next_frontier, uint32_t *active, uint32_t *next_active) {
    memset(seen, 0, V * sizeof(uint64_t));
    memset(frontier, 0, V * sizeof(uint64_t));
    memset(next_frontier, 0, V * sizeof(uint64_t));
    size_t na = 0, nn = 0;
    for (size_t k = 0; k < bs; ++k) {
        uint32_t s = batch[k];
        uint64_t mask = 1ull << k;
        dist[(batch_start + k) * V + s] = 0;
        if (frontier[s] == 0) active[na++] = s;
        frontier[s] |= mask;
        seen[s] |= mask;
    }
    uint32_t cur = 0;
    while (na) {
        uint32_t nd = cur + 1;
        for (size_t i = 0; i < na; ++i) {
            uint32_t u = active[i];
            uint64_t f = frontier[u];
            frontier[u] = 0;
            for (uint32_t e = ro[u]; e < ro[u + 1]; ++e) {
                uint32_t v = ci[e];
                if (next_frontier[v] == 0) next_active[nn++] = v;
                next_frontier[v] |= f;
            }
        }
        na = 0;
        for (size_t i = 0; i < nn; ++i) {
            uint32_t v = next_active[i];
            uint64_t nf = next_frontier[v];
            next_frontier[v] = 0;
            uint64_t nv = nf & ~seen[v];
            if (nv) {
                seen[v] |= nv;
                frontier[v] = nv;
                active[na++] = v;
                while (nv) {
                    int b = __builtin_ctzll(nv);
                    dist[(batch_start + (size_t)b) * V + v] = nd;
                    nv &= nv - 1;
                }
            }
        }
        nn = 0;
        cur += 1;
    }
}

/* ms_bfs_bit_parallel (bit_parallel.rs:101-130): chunks of 64 sources. */
int orc_ms_bfs_bit_parallel(size_t V, const uint32_t *ro, const uint32_t *ci,
                            const uint32_t *sources, size_t S, uint32_t *dist) {
    if (S == 0 || V == 0) return ORC_OK;
    if (validate_sources(V, sources, S, NULL)) return ORC_ERR_SOURCE_OOB;
    for (size_t i = 0; i < S * V; ++i) dist[i] = ORC_INF;
    uint64_t *seen = (uint64_t *)malloc(V * sizeof(uint64_t));
    uint64_t *fr = (uint64_t *)malloc(V * sizeof(uint64_t));
    uint64_t *nf = (uint64_t *)malloc(V * sizeof(uint64_t));
    /* next_active may receive one entry per edge endpoint touched before the
     * dedup by next_frontier==0, so V entries suffice. */
    uint32_t *act = (uint32_t *)malloc(V * sizeof(uint32_t));
    uint32_t *nact = (uint32_t *)malloc(V * sizeof(uint32_t));
    for (size_t b = 0; b < S; b += 64) {
        size_t bs = S - b < 64 ? S - b : 64;
        bit_parallel_batch_64(V, ro, ci, sources + b, bs, b, dist, seen, fr, nf, act, nact);
    }
    free(seen); free(fr); free(nf); free(act); free(nact);
    return ORC_OK;
}

/* ---- Traversed-edge count (bench.rs:478-491): sum over sources of the
 * out-degrees of every reached vertex. ---- */
uint64_t orc_traversed_edges(size_t V, const uint32_t *ro, size_t S, const uint32_t *dist) {
    uint64_t te = 0;
    for (size_t s = 0; s < S; ++s)
        for (size_t u = 0; u < V; ++u)
            if (dist[s * V + u] != ORC_INF) te += ro[u + 1] - ro[u];
    return te;
}

/* ---- Reachability bitmask from canonical distances (graph.rs:571-645):
 * words[v*wpv + s/32] bit (s%32) set iff dist[s*V+v] != INF. ---- */
void orc_reachability_from_dist(size_t V, size_t S, const uint32_t *dist, uint32_t *words) {
    size_t wpv = S ? (S + 31) / 32 : 0;
    memset(words, 0, V * wpv * sizeof(uint32_t));
    for (size_t s = 0; s < S; ++s)
        for (size_t v = 0; v < V; ++v)
            if (dist[s * V + v] != ORC_INF) words[v * wpv + s / 32] |= 1u << (s % 32);
}

/* ---- Grid2D Manhattan oracle (generators.rs:98-127). ---- */
void orc_grid_manhattan(size_t width, size_t height, const uint32_t *sources, size_t S,
                        uint32_t *dist) {
    size_t V = width * height;
    for (size_t s = 0; s < S; ++s) {
        long sx = (long)(sources[s] % width), sy = (long)(sources[s] / width);
        for (size_t v = 0; v < V; ++v) {
            long vx = (long)(v % width), vy = (long)(v / width);
            long dx = sx > vx ? sx - vx : vx - sx, dy = sy > vy ? sy - vy : vy - sy;
            dist[s * V + v] = (uint32_t)(dx + dy);
        }
    }
}

/* ---- Output codecs.  u16: min(d, 0xFFFF) packed two per u32, low half
 * first; INF maps to 0xFFFF and back to INF.  u8: min(d, 0xFF), four per
 * u32, INF <-> 0xFF. ---- */
void orc_pack_u16(size_t n, const uint32_t *d, uint32_t *out) {
    size_t nw = (n + 1) / 2;
    for (size_t w = 0; w < nw; ++w) {
        uint32_t lo = d[2 * w] > 0xFFFFu ? 0xFFFFu : d[2 * w];
        uint32_t hi = (2 * w + 1 < n) ? (d[2 * w + 1] > 0xFFFFu ? 0xFFFFu : d[2 * w + 1]) : 0xFFFFu;
        out[w] = lo | (hi << 16);
    }
}

void orc_unpack_u16(size_t n, const uint32_t *words, uint32_t *d) {
    for (size_t i = 0; i < n; ++i) {
        uint32_t x = (words[i / 2] >> (16 * (i & 1))) & 0xFFFFu;
        d[i] = x == 0xFFFFu ? ORC_INF : x;
    }
}

void orc_pack_u8(size_t n, const uint32_t *d, uint32_t *out) {
    size_t nw = (n + 3) / 4;
    for (size_t w = 0; w < nw; ++w) {
        uint32_t word = 0;
        for (size_t k = 0; k < 4; ++k) {
            size_t i = 4 * w + k;
            uint32_t x = i < n ? (d[i] > 0xFFu ? 0xFFu : d[i]) : 0xFFu;
            word |= x << (8 * k);
        }
        out[w] = word;
    }
}

void orc_unpack_u8(size_t n, const uint32_t *words, uint32_t *d) {
    for (size_t i = 0; i < n; ++i) {
        uint32_t x = (words[i / 4] >> (8 * (i & 3))) & 0xFFu;
        d[i] = x == 0xFFu ? ORC_INF : x;
    }
}
