// Hand-written gfx950 (CDNA4, wave64) kernels of the multi-source BFS.
//
// All arithmetic is integer: source sets are bit masks, a BFS level ORs the
// frontier masks of a vertex's parents into its children and keeps the bits
// not seen before.  No MFMA: the boolean semiring product is a gather/scatter
// over the graph, bounded by memory traffic, not a dense contraction.
//
// Two families:
//  * deep  -- one workgroup per group of B <= 32 sources.  The group's
//             visited bits (B per vertex) live in LDS; the frontier is a
//             queue of (vertex, new-bits) entries in global memory, filled
//             with wave-aggregated appends (ballot + popcount prefix, one LDS
//             atomic per wave).  Every level of the group runs inside the
//             workgroup with s_barrier only; workgroups pull groups from a
//             global counter (persistent, no grid barrier).  Suited to
//             high-diameter graphs, where a device-wide barrier per level
//             would dominate.
//  * wide  -- one cooperative persistent kernel for all S sources at once,
//             64 sources per 64-bit word, state in HBM word-major ([w][v]).
//             Per level it either pushes (CSR, atomicOr of 64-bit words,
//             candidate queue deduplicated by a level stamp, then a filter
//             pass) or pulls (CSC, per (vertex, word) gather of parent words
//             with early exit once every still-needed bit is found), chosen
//             on the device from the previous level's counters.  Levels are
//             separated by grid barriers; the host is not involved.
//
// The reference computes the same distances with WGSL kernels driven level
// by level from the host (crates/shaders/src/bit_parallel_64_push.wgsl:15-48,
// bit_parallel_64.wgsl:18-78, bit_parallel_64_pull.wgsl:27-111,
// sparse_bp_push.wgsl:23-70, workgroup_persistent_bfs.wgsl:39-196).
#include <hip/hip_runtime.h>

#include "kernels.hpp"

#include <algorithm>
#include <type_traits>


namespace msbfs {

typedef unsigned long long u64;

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// acc + the number of bits of m set for lanes below this one (v_mbcnt pair:
// two instructions in place of masking and counting each half).
__device__ __forceinline__ uint32_t mbcnt64(u64 m, uint32_t acc) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, acc));
}

__device__ __forceinline__ u64 wave_or64(u64 x) {
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo |= (uint32_t)__shfl_xor((int)lo, off);
        hi |= (uint32_t)__shfl_xor((int)hi, off);
    }
    return ((u64)hi << 32) | lo;
}

__device__ __forceinline__ u64 wave_sum64(u64 x) {
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t olo = (uint32_t)__shfl_xor((int)lo, off);
        uint32_t ohi = (uint32_t)__shfl_xor((int)hi, off);
        u64 s = ((u64)hi << 32 | lo) + ((u64)ohi << 32 | olo);
        lo = (uint32_t)s;
        hi = (uint32_t)(s >> 32);
    }
    return ((u64)hi << 32) | lo;
}

// Distance stores.  NT = true streams the store past L2 (non-temporal): used
// for the wide family's row-coalesced stores, which fill whole lines and are
// never read back.  Scattered single-entry stores stay cached so that L2 can
// merge the partial lines they write.
// Output element kinds (OutKind): 0 u32, 1 u16, 2 u8.  A distance that the
// kind cannot hold below its sentinel raises FlagU16Overflow (the flag means
// "distance does not fit the output type") and is stored clamped.
template <bool NT, typename T>
__device__ __forceinline__ void store_elem(void *out, uint64_t idx, T v) {
    if (NT)
        __builtin_nontemporal_store(v, (T *)out + idx);
    else
        ((T *)out)[idx] = v;
}

template <bool NT = false>
__device__ __forceinline__ void store_dist(void *out, uint32_t kind, uint64_t idx, uint32_t d,
                                           uint32_t *flags) {
    if (kind == 1) {
        if (d >= 0xFFFFu) {
            atomicOr(&flags[FlagU16Overflow], 1u);
            d = 0xFFFEu;
        }
        store_elem<NT>(out, idx, (uint16_t)d);
    } else if (kind == 2) {
        if (d >= 0xFFu) {
            atomicOr(&flags[FlagU16Overflow], 1u);
            d = 0xFEu;
        }
        store_elem<NT>(out, idx, (uint8_t)d);
    } else {
        store_elem<NT>(out, idx, d);
    }
}

template <bool NT = false>
__device__ __forceinline__ void store_sentinel(void *out, uint32_t kind, uint64_t idx) {
    if (kind == 1)
        store_elem<NT>(out, idx, (uint16_t)0xFFFFu);
    else if (kind == 2)
        store_elem<NT>(out, idx, (uint8_t)0xFFu);
    else
        store_elem<NT>(out, idx, 0xFFFFFFFFu);
}

__global__ void zero_u32_kernel(uint32_t *p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

hipError_t launch_zero_u32(uint32_t *ptr, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, ptr, n);
    return hipGetLastError();
}

// Eight entries per thread: one 16-byte load, one 8-byte store.
__global__ void pack_u16_u8_kernel(const uint16_t *in, uint8_t *out, uint64_t n) {
    const uint64_t n8 = n / 8;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n8;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 x = ((const uint4 *)in)[i];
        const uint32_t w[4] = {x.x, x.y, x.z, x.w};
        u64 r = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t lo = w[k] & 0xFFFFu, hi = w[k] >> 16;
            r |= (u64)(lo == 0xFFFFu ? 0xFFu : lo) << (16 * k);
            r |= (u64)(hi == 0xFFFFu ? 0xFFu : hi) << (16 * k + 8);
        }
        ((u64 *)out)[i] = r;
    }
    if (blockIdx.x == 0 && threadIdx.x < n % 8) {
        const uint64_t j = n8 * 8 + threadIdx.x;
        const uint32_t v = in[j];
        out[j] = (uint8_t)(v == 0xFFFFu ? 0xFFu : v);
    }
}

// u16 distances (all < 2^W - 1, or 0xFFFF) packed W bits each, W = 4 or 12,
// in groups of 8 (W bytes per group, element k at bits [W k, W k + W) of the
// group's little-endian bit string), the sentinel as all ones; a ragged last
// group is padded with sentinels.
template <int W>
__global__ void pack_u16_bits_kernel(const uint16_t *in, uint32_t *out, uint64_t n) {
    static_assert(W == 4 || W == 12, "4 or 12 bits per distance");
    constexpr uint32_t SENT = (1u << W) - 1u;
    const uint64_t ng = (n + 7) / 8, full = n / 8;
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < ng;
         g += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t v[8];
        if (g < full) {
            const uint4 x = ((const uint4 *)in)[g];
            const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                v[2 * k] = w[k] & 0xFFFFu;
                v[2 * k + 1] = w[k] >> 16;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = g * 8 + k < n ? (uint32_t)in[g * 8 + k] : 0xFFFFu;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = v[k] == 0xFFFFu ? SENT : v[k];
        if constexpr (W == 4) {
            uint32_t r = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) r |= v[k] << (4 * k);
            out[g] = r;
        } else {
            out[3 * g] = v[0] | v[1] << 12 | v[2] << 24;
            out[3 * g + 1] = v[2] >> 8 | v[3] << 4 | v[4] << 16 | v[5] << 28;
            out[3 * g + 2] = v[5] >> 4 | v[6] << 8 | v[7] << 20;
        }
    }
}

hipError_t launch_pack_u16_bits(const uint16_t *in, uint32_t *out, uint64_t n, int bits, hipStream_t st) {
    const uint64_t ng = (n + 7) / 8;
    if (ng == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((ng + 255) / 256, 8192);
    if (bits == 4)
        hipLaunchKernelGGL(pack_u16_bits_kernel<4>, dim3((unsigned)blocks), dim3(256), 0, st, in, out, n);
    else if (bits == 12)
        hipLaunchKernelGGL(pack_u16_bits_kernel<12>, dim3((unsigned)blocks), dim3(256), 0, st, in, out, n);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_pack_u16_u8(const uint16_t *in, uint8_t *out, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n / 8 + 255) / 256;
    blocks = std::max<uint64_t>(1, std::min<uint64_t>(blocks, 8192));
    hipLaunchKernelGGL(pack_u16_u8_kernel, dim3((unsigned)blocks), dim3(256), 0, st, in, out, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------- deep family

template <int B>
struct DeepEntry;
template <>
struct DeepEntry<1> { typedef uint32_t T; };
template <>
struct DeepEntry<2> { typedef uint32_t T; };
template <>
struct DeepEntry<4> { typedef uint32_t T; };
template <>
struct DeepEntry<8> { typedef uint32_t T; };
template <>
struct DeepEntry<16> { typedef u64 T; };
template <>
struct DeepEntry<32> { typedef u64 T; };

// Entries pack (vertex, bits).  B <= 8: v << 8 | bits (needs V < 2^24);
// B >= 16: bits << 32 | v.
__device__ __forceinline__ uint32_t pack_entry(uint32_t v, uint32_t m, uint32_t *) {
    return (v << 8) | m;
}
__device__ __forceinline__ u64 pack_entry(uint32_t v, uint32_t m, u64 *) {
    return ((u64)m << 32) | v;
}
__device__ __forceinline__ void unpack_entry(uint32_t x, uint32_t &v, uint32_t &m) {
    v = x >> 8;
    m = x & 0xFFu;
}
__device__ __forceinline__ void unpack_entry(u64 x, uint32_t &v, uint32_t &m) {
    v = (uint32_t)x;
    m = (uint32_t)(x >> 32);
}

// A frontier queue: the first `lcap` entries live in LDS, the rest spill to
// this block's region of global memory.
// Entries are read through address-space-qualified pointers, so the LDS
// part is a ds_read (waits on LDS traffic only) and the spill part a global
// load, never one flat load that waits on both.
template <typename Entry>
struct DeepQueue {
    Entry *lds;
    Entry *glb;
    uint32_t lcap;
    uint64_t gcap;
    __device__ __forceinline__ Entry get(uint32_t i) const {
        if (i < lcap) return ((__attribute__((address_space(3))) Entry *)lds)[i];
        return ((__attribute__((address_space(1))) Entry *)glb)[i - lcap];
    }
    __device__ __forceinline__ bool put(uint32_t i, Entry x) const {
        if (i < lcap) {
            lds[i] = x;
            return true;
        }
        if ((uint64_t)(i - lcap) < gcap) {
            glb[i - lcap] = x;
            return true;
        }
        return false;
    }
};

// Distance d (already clamped for u16) for every source bit of `bits` at
// vertex v.
// Rows of one task (< 2^29 entries: |V| < 2^24, <= 32 sources) are indexed
// with 32-bit offsets from the task's first row.
template <int B, uint32_t KIND>
__device__ __forceinline__ void deep_write_bits(void *out_task, uint32_t V, uint32_t v,
                                                uint32_t bits, uint32_t d) {
    while (bits) {
        const uint32_t off = (uint32_t)__builtin_ctz(bits) * V + v;
        if constexpr (KIND == 1)
            ((__attribute__((address_space(1))) uint16_t *)out_task)[off] = (uint16_t)d;
        else if constexpr (KIND == 2)
            ((__attribute__((address_space(1))) uint8_t *)out_task)[off] = (uint8_t)d;
        else
            ((__attribute__((address_space(1))) uint32_t *)out_task)[off] = d;
        bits &= bits - 1;
    }
}

// One workgroup per task (a group of <= B sources of one graph).  Visited
// bits in LDS; frontier queues in LDS with global spill.  With an ELL
// neighbour table (graphs of degree <= ELL) a queue entry costs one vector
// load of its neighbour slots, its claims are independent LDS atomics, and
// the wave appends all its new entries with one LDS atomic.
//
// DT = 5 / 6 (u32 / u16 rows, one source per task) stages the task's row in
// LDS: a 4-bit code per vertex, distance + 1 for distances below 14 and 15
// for the rest, which are stored directly as found.  The code is set with
// one LDS atomic when the entry is expanded, in place of a scattered store
// into the row, and the whole row is written with coalesced stores when the
// task ends (sentinels included; codes of 15 are skipped).  Scattered
// stores leave lines partly written in L2 over several levels and load the
// memory system the neighbour-list loads wait on (512 G(n,p) graphs of
// 1K-64K vertices: 12.70 ms with them, 9.12 ms with none at all).
template <int B, int ELL, int DT>
__global__ void __launch_bounds__(1024) deep_bfs_kernel(DeepParams p) {
    // DT: 0 = no distances (reachability only), 1 = u16, 2 = u32, 3 = u8,
    // 5 = u32 staged, 6 = u16 staged.
    constexpr bool OUT = DT != 0;
    constexpr bool STG = DT == 5 || DT == 6;
    static_assert(!STG || B == 1, "staged rows hold one source per task");
    constexpr bool U16 = DT == 1 || DT == 6;
    constexpr uint32_t KIND = U16 ? 1u : DT == 3 ? 2u : 0u;  // output element kind
    constexpr uint32_t ESZ = U16 ? 2u : DT == 3 ? 1u : 4u;
    constexpr uint32_t DMAX = U16 ? 0xFFFEu : DT == 3 ? 0xFEu : 0xFFFFFFFFu;
    constexpr uint32_t kDirect = 14;  // staged rows: levels from here on are stored as found
    typedef typename DeepEntry<B>::T Entry;
    constexpr uint32_t VPW = 32 / B;  // vertices per LDS word
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_task;
    __shared__ uint32_t s_cnt[4];
    __shared__ Entry s_sink[64];  // target of the appends a lane does not make

    uint32_t *const vis = lds;
    const uint32_t tid = threadIdx.x, nth = blockDim.x;
    const uint32_t lane = lane_id();
    Entry *const lq0 = (Entry *)(lds + p.lds_vis_words);
    Entry *const lq1 = lq0 + p.lds_qcap;
    const uint64_t gcap = p.cap - p.lds_qcap;  // spill entries per queue
    Entry *const gq0 = (Entry *)p.queues + (uint64_t)blockIdx.x * 2 * gcap;
    Entry *const gq1 = gq0 + gcap;

    for (;;) {
        if (tid == 0) s_task = atomicAdd(p.work, 1u);
        __syncthreads();
        const uint32_t t = s_task;
        if (t >= p.ntasks) break;
        const DeepTask task = p.tasks[t];
        const DeepGraph g = p.graphs[task.graph];
        const uint32_t V = g.V;
        const uint32_t nwords = (V + VPW - 1) / VPW;
        // staged rows: 8 codes per word after this task's visited bits
        __attribute__((address_space(3))) uint32_t *const nib =
            (__attribute__((address_space(3))) uint32_t *)(vis + nwords);
        const uint32_t nclear = nwords + (STG ? (V + 7) / 8 : 0);
        for (uint32_t i = tid; i < nclear; i += nth) vis[i] = 0;
        if (tid == 0) {
            s_cnt[0] = task.nsrc;
            s_cnt[1] = 0;
            s_cnt[2] = 0;
            s_cnt[3] = 0;
        }
        __syncthreads();

        DeepQueue<Entry> qa{lq0, gq0, p.lds_qcap, gcap};
        DeepQueue<Entry> qb{lq1, gq1, p.lds_qcap, gcap};
        if (tid < task.nsrc) {
            const uint32_t s = p.sources[task.src_off + tid];
            const uint32_t bit = 1u << tid;
            atomicOr(&vis[s / VPW], bit << ((s % VPW) * B));
            qa.put(tid, pack_entry(s, bit, (Entry *)nullptr));
        }
        __syncthreads();

        void *const out_task =
            OUT ? (void *)((char *)p.out + task.out_off * ESZ) : nullptr;
        // Distances are written when an entry is expanded: the entries of
        // level L are the (vertex, source bits) pairs first reached at
        // distance L.  The store of an entry's lowest bit is issued right
        // after its neighbour-row load, so it has a whole level to complete
        // before the next level's load waits on the memory counter.
        uint32_t level = 0;
        for (;;) {
            // Rotating counters: level L reads s_cnt[L&3], appends to
            // s_cnt[(L+1)&3] and clears s_cnt[(L+2)&3], last read at L-2, so
            // one barrier per level orders everything.
            // The first pass's entry is read alongside the count (both LDS,
            // one wait); the queue's LDS part always holds index tid.
            Entry x0 = 0;
            if constexpr (ELL > 0)
                if (tid < qa.lcap) x0 = ((__attribute__((address_space(3))) Entry *)qa.lds)[tid];
            const uint32_t n = s_cnt[level & 3];
            if (n == 0) break;
            uint32_t *const cnt_next = &s_cnt[(level + 1) & 3];
            if (tid == 0) {
                s_cnt[(level + 2) & 3] = 0;
                if (OUT && level > DMAX) atomicOr(&p.flags[FlagU16Overflow], 1u);
            }
            const uint32_t dval = min(level, DMAX);
            const uint32_t code = min(level + 1, 15u);  // staged rows
            for (uint32_t base = 0; base < n; base += nth) {
                const uint32_t i = base + tid;
                uint32_t u = 0, f = 0;
                if constexpr (ELL > 0) {
                    if (base + (tid & ~63u) >= n) continue;  // wave-uniform: no entries
                    // A pass within the LDS part of the queue reads it with
                    // ds_reads only, so the row load is not held behind a
                    // wait on outstanding stores.
                    auto pass = [&](auto spill) {
                        const bool act = i < n;
                        if (act) {
                            Entry x;
                            if constexpr (decltype(spill)::value)
                                x = qa.get(i);
                            else
                                x = base == 0 ? x0 : ((__attribute__((address_space(3))) Entry *)qa.lds)[i];
                            unpack_entry(x, u, f);
                        }
                        // All neighbour slots in one or two 16-byte global loads.
                        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                        uint32_t nb[ELL];
                        if (act) {
                            const __attribute__((address_space(1))) u32x4 *row =
                                (const __attribute__((address_space(1))) u32x4 *)(g.ell +
                                                                                  (uint64_t)u * ELL);
#pragma unroll
                            for (int k = 0; k < ELL / 4; ++k) {
                                const u32x4 x = row[k];
                                nb[4 * k] = x.x;
                                nb[4 * k + 1] = x.y;
                                nb[4 * k + 2] = x.z;
                                nb[4 * k + 3] = x.w;
                            }
                        }
                        // This entry's lowest source bit, one straight-line
                        // store per lane; a lane without an entry repeats lane
                        // 0's store (entries are contiguous, so lane 0 has one).
                        if constexpr (STG)
                            if (act)
                                __hip_atomic_fetch_or(nib + u / 8, code << ((u % 8) * 4),
                                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (OUT && (!STG || level >= kDirect)) {
                            const uint32_t b0 = __builtin_ctz(f | 0x80000000u);
                            const uint32_t u_l0 = (uint32_t)__builtin_amdgcn_readlane((int)u, 0);
                            const uint32_t b_l0 = (uint32_t)__builtin_amdgcn_readlane((int)b0, 0);
                            const uint32_t off = (act ? b0 : b_l0) * V + (act ? u : u_l0);
                            if constexpr (KIND == 1)
                                ((__attribute__((address_space(1))) uint16_t *)out_task)[off] = (uint16_t)dval;
                            else if constexpr (KIND == 2)
                                ((__attribute__((address_space(1))) uint8_t *)out_task)[off] = (uint8_t)dval;
                            else
                                ((__attribute__((address_space(1))) uint32_t *)out_task)[off] = dval;
                        }
                        // Claims: all slots' LDS atomics issued at once, then
                        // one wait; the returned words give the new bits.
                        // (Empty slots name the vertex itself: nothing new.)
                        uint32_t newly[ELL];
                        uint32_t c = 0;
                        if (act) {
                            __attribute__((address_space(3))) uint32_t *const lv =
                                (__attribute__((address_space(3))) uint32_t *)vis;
                            uint32_t old[ELL];
#pragma unroll
                            for (int k = 0; k < ELL; ++k)
                                old[k] = __hip_atomic_fetch_or(lv + nb[k] / VPW,
                                                               f << ((nb[k] % VPW) * B),
                                                               __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
                            for (int k = 0; k < ELL; ++k) {
                                newly[k] = f & ~(old[k] >> ((nb[k] % VPW) * B));
                                c += newly[k] != 0;
                            }
                        } else {
#pragma unroll
                            for (int k = 0; k < ELL; ++k) newly[k] = 0;
                        }
                        // Exclusive prefix of c over the wave from ELL ballots:
                        // lanes below with c > j, summed over j.
                        uint32_t pre = 0, tot = 0;
#pragma unroll
                        for (int j = 0; j < ELL; ++j) {
                            const u64 bj = __ballot(c > (uint32_t)j);
                            pre = mbcnt64(bj, pre);
                            tot += (uint32_t)__popcll(bj);
                        }
                        uint32_t wbase = 0;
                        if (tot) {
                            if (lane == 0) wbase = atomicAdd(cnt_next, tot);
                            wbase = (uint32_t)__builtin_amdgcn_readfirstlane((int)wbase);
                        }
                        uint32_t pos = wbase + pre;
                        if (wbase + tot <= qb.lcap) {
                            // The wave's appends all land in LDS: every slot
                            // writes, the ones without new bits into this
                            // lane's sink word, so there are no branches.
                            __attribute__((address_space(3))) Entry *const q =
                                (__attribute__((address_space(3))) Entry *)qb.lds;
                            __attribute__((address_space(3))) Entry *const sk =
                                (__attribute__((address_space(3))) Entry *)s_sink + lane;
#pragma unroll
                            for (int k = 0; k < ELL; ++k) {
                                *(newly[k] ? q + pos : sk) = pack_entry(nb[k], newly[k], (Entry *)nullptr);
                                pos += newly[k] != 0;
                            }
                        } else {
#pragma unroll
                            for (int k = 0; k < ELL; ++k) {
                                if (newly[k]) {
                                    if (!qb.put(pos, pack_entry(nb[k], newly[k], (Entry *)nullptr)))
                                        atomicOr(&p.flags[FlagQueueOverflow], 1u);
                                    ++pos;
                                }
                            }
                        }
                        // The entry's other source bits (rare: sources whose
                        // searches reach the vertex at the same level).
                        if constexpr (OUT && B > 1)
                            if (act) deep_write_bits<B, KIND>(out_task, V, u, f & (f - 1), dval);
                    };
                    if (base + nth <= qa.lcap)
                        pass(std::false_type{});
                    else
                        pass(std::true_type{});
                } else {
                    if (base + (tid & ~63u) >= n) continue;  // wave-uniform: no entries
                    // CSR neighbour lists: four edges per lane per step, the
                    // indices clamped to the list (a repeated neighbour claims
                    // nothing new), all claims issued before one wait, one
                    // queue reservation per step, branch-free LDS appends.
                    auto pass = [&](auto spill) {
                        typedef const __attribute__((address_space(1))) uint32_t *g32;
                        const bool act = i < n;
                        if (act) {
                            Entry x;
                            if constexpr (decltype(spill)::value)
                                x = qa.get(i);
                            else
                                x = ((__attribute__((address_space(3))) Entry *)qa.lds)[i];
                            unpack_entry(x, u, f);
                            if constexpr (STG) {
                                __hip_atomic_fetch_or(nib + u / 8, code << ((u % 8) * 4),
                                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                if (level >= kDirect) deep_write_bits<B, KIND>(out_task, V, u, f, dval);
                            } else if (OUT) {
                                deep_write_bits<B, KIND>(out_task, V, u, f, dval);
                            }
                        }
                        uint32_t e = 0, eend = 0;
                        if (act) {
                            e = ((g32)g.ro)[u];
                            eend = ((g32)g.ro)[u + 1];
                        }
                        __attribute__((address_space(3))) uint32_t *const lv =
                            (__attribute__((address_space(3))) uint32_t *)vis;
                        constexpr int U = 4;
                        while (__any(e < eend)) {
                            uint32_t nb[U], newly[U];
                            uint32_t c = 0;
#pragma unroll
                            for (int k = 0; k < U; ++k) newly[k] = 0;
                            if (e < eend) {
#pragma unroll
                                for (int k = 0; k < U; ++k)
                                    nb[k] = ((g32)g.ci)[min(e + (uint32_t)k, eend - 1)];
                                uint32_t old[U];
#pragma unroll
                                for (int k = 0; k < U; ++k)
                                    old[k] = __hip_atomic_fetch_or(lv + nb[k] / VPW,
                                                                   f << ((nb[k] % VPW) * B),
                                                                   __ATOMIC_RELAXED,
                                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
                                for (int k = 0; k < U; ++k) {
                                    newly[k] = f & ~(old[k] >> ((nb[k] % VPW) * B));
                                    c += newly[k] != 0;
                                }
                            }
                            uint32_t pre = 0, tot = 0;
#pragma unroll
                            for (int j = 0; j < U; ++j) {
                                const u64 bj = __ballot(c > (uint32_t)j);
                                pre = mbcnt64(bj, pre);
                                tot += (uint32_t)__popcll(bj);
                            }
                            if (tot) {
                                uint32_t wbase = 0;
                                if (lane == 0) wbase = atomicAdd(cnt_next, tot);
                                wbase = (uint32_t)__builtin_amdgcn_readfirstlane((int)wbase);
                                uint32_t pos = wbase + pre;
                                if (wbase + tot <= qb.lcap) {
                                    __attribute__((address_space(3))) Entry *const q =
                                        (__attribute__((address_space(3))) Entry *)qb.lds;
                                    __attribute__((address_space(3))) Entry *const sk =
                                        (__attribute__((address_space(3))) Entry *)s_sink + lane;
#pragma unroll
                                    for (int k = 0; k < U; ++k) {
                                        *(newly[k] ? q + pos : sk) =
                                            pack_entry(newly[k] ? nb[k] : 0u, newly[k], (Entry *)nullptr);
                                        pos += newly[k] != 0;
                                    }
                                } else {
#pragma unroll
                                    for (int k = 0; k < U; ++k) {
                                        if (newly[k]) {
                                            if (!qb.put(pos, pack_entry(nb[k], newly[k], (Entry *)nullptr)))
                                                atomicOr(&p.flags[FlagQueueOverflow], 1u);
                                            ++pos;
                                        }
                                    }
                                }
                            }
                            e += U;
                        }
                    };
                    if (base + nth <= qa.lcap)
                        pass(std::false_type{});
                    else
                        pass(std::true_type{});
                }
            }
            DeepQueue<Entry> tq = qa;
            qa = qb;
            qb = tq;
            ++level;
            __syncthreads();
        }
        if (tid == 0) atomicMax(&p.flags[FlagLevels], level);

        // Staged rows: the whole row from its codes, consecutive threads on
        // consecutive vertices (every visited vertex has a code: it was
        // expanded the level after it was claimed).
        if constexpr (STG) {
            for (uint32_t v = tid; v < V; v += nth) {
                const uint32_t c = (nib[v / 8] >> ((v % 8) * 4)) & 15u;
                if (c == 0)
                    store_sentinel(out_task, KIND, v);
                else if (c != 15u)
                    store_elem<false>(out_task, v, (typename std::conditional<U16, uint16_t, uint32_t>::type)(c - 1));
            }
        }
        // Unreached (source, vertex) pairs get the sentinel; rows are written
        // by consecutive threads, so the stores coalesce.  A wave first tests
        // 64 visited words at once and skips their vertices when every bit is
        // set (a connected graph has nothing to write).
        if constexpr (OUT && !STG) {
            const uint32_t lane = lane_id(), wv = tid / 64, nw = max(1u, nth / 64);
            for (uint32_t i0 = wv * 64; i0 < nwords; i0 += nw * 64) {
                const uint32_t i = i0 + lane;
                if (__all(i >= nwords || vis[i] == 0xFFFFFFFFu)) continue;
                const uint32_t v1 = min(V, (i0 + 64) * VPW);
                for (uint32_t b = 0; b < task.nsrc; ++b) {
                    const uint64_t row = task.out_off + (uint64_t)b * V;
                    for (uint32_t v = i0 * VPW + lane; v < v1; v += 64)
                        if (!((vis[v / VPW] >> ((v % VPW) * B + b)) & 1u))
                            store_sentinel(p.out, KIND, row + v);
                }
            }
        }
        if (p.reach_t) {
            // each source's row of the source-major bitmap, 64 vertices per
            // ballot and wave (no atomics: the row belongs to this task)
            const uint32_t lane = lane_id(), wv = tid / 64, nw = max(1u, nth / 64);
            for (uint32_t b = 0; b < task.nsrc; ++b) {
                uint32_t *const row = p.reach_t + (task.reach_bit + b) * (uint64_t)p.reach_vw;
                for (uint32_t v0 = wv * 64; v0 < V; v0 += nw * 64) {
                    const uint32_t v = v0 + lane;
                    const u64 m = __ballot(v < V && ((vis[v / VPW] >> ((v % VPW) * B + b)) & 1u));
                    if (lane == 0) row[v0 / 32] = (uint32_t)m;
                    if (lane == 32 && v0 + 32 < V) row[v0 / 32 + 1] = (uint32_t)(m >> 32);
                }
            }
        }
        __syncthreads();
    }
}

// One task per wave (workgroups of 64 threads) for graphs with a neighbour
// table: the frontier counts live in scalar registers, queue slots are
// reserved with ballot prefix sums instead of an LDS atomic, and the wave
// needs no barrier (its LDS operations execute in order).  LDS layout,
// queues with global spill, outputs and flags as in deep_bfs_kernel.
//
// DT = 4 (u32 rows, one source per task) writes no distance in place while
// the search runs.  Scattered 4-byte stores into 2,048 rows of 458 KB fill
// each 128-byte line over many levels, and the lines leave L2 and the
// Infinity Cache partly written (measured: 6.4x the output written, and a
// third of the kernel time at road-like S=2048).  Instead each expanded
// entry appends (vertex mod 2^log_bw, level) to its vertex bin's log, and
// the bin's log lives in the bin's own span of the task's output row (a
// bin of n vertices receives at most n entries), so the appends are
// sequential streams that L2 merges into whole lines.  When the task ends,
// each bin's log is scattered into LDS (sentinels first) and the span is
// rewritten in order with coalesced stores.
//
// LAT > 0: the lattice kernel.  The graph's lattice adjacency (LAT bits per
// vertex, DeepGraph::lat) is copied into LDS once per workgroup and shared
// by its waves, each running its own tasks in its own LDS region, so a
// level's chain is LDS operations only: entry read, adjacency fields of the
// vertex and of its lower neighbours, claims, appends.  (Road-like graph on
// a 339-column grid, 114,599 vertices: 2 bits per vertex, 28.6 KB.)

// Bit `off % 32` of w sign-extended: 0 or all ones (one v_bfe_i32 with the
// offset in a register).
__device__ __forceinline__ uint32_t bit_mask_at(uint32_t w, uint32_t off) {
    return (uint32_t)__builtin_amdgcn_sbfe((int)w, off, 1u);
}

// The lattice neighbours of vertex u.  The fields follow `pad` zero fields
// (pad >= every offset).  Vertex ids are biased (see the kernel) and `c`
// maps a biased id to its field index in LDS terms: field x lives in the
// word at LDS byte address (x / FPW) * 4, so no base address is added per
// read.
template <int LAT>
__device__ __forceinline__ void lattice_neighbours(const uint32_t (&d)[4], uint32_t c, uint32_t u,
                                                   uint32_t (&nb)[2 * LAT]) {
    constexpr uint32_t FPW = 32 / LAT;  // vertex fields per word
    typedef const __attribute__((address_space(3))) uint32_t *cl32_t;
    // bit k of field x: bit (x % FPW) * LAT + k of its word (the v_bfe_i32
    // offset takes the low 5 bits of x * LAT + k)
    auto word = [](uint32_t x) -> uint32_t { return *(cl32_t)(size_t)((x / FPW) * 4u); };
    // every word is read before any is used (the reads overlap)
    const uint32_t x = u + c;
    const uint32_t wu = word(x);
    uint32_t wm[LAT];
#pragma unroll
    for (int k = 0; k < LAT; ++k)
        // the lower neighbour's field: a padding field below vertex 0, or
        // the vertex's own for an unused offset 0 (its bit k is never set)
        wm[k] = word(x - d[k]);
    // (keeps the scheduler from placing a use, and its wait, between the reads)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < LAT; ++k) {
        // a missing edge leaves the vertex itself, whose claim finds nothing
        nb[k] = u + (bit_mask_at(wu, x * LAT + k) & d[k]);
        nb[LAT + k] = u - (bit_mask_at(wm[k], (x - d[k]) * LAT + k) & d[k]);
    }
}

// Per lane: mask bit of this lane set ? if_set : if_clear (one v_cndmask on a
// wave mask already in scalar registers, e.g. a ballot).
__device__ __forceinline__ uint32_t lane_select(u64 mask, uint32_t if_clear, uint32_t if_set) {
    uint32_t r;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(if_clear), "v"(if_set), "s"(mask));
    return r;
}

template <int B, int ELL, int DT, int LAT = 0>
__global__ void __launch_bounds__(LAT ? 64 * kLatticeMaxWaves : 64) deep_wave_kernel(DeepParams p) {
    static_assert(ELL == 4 || ELL == 8, "the wave kernel reads neighbour tables");
    static_assert(LAT == 0 || (ELL == 2 * LAT && B == 1),
                  "a lattice vertex has 2 x LAT neighbour slots; one source per task");
    static_assert(DT != 4 || B == 1, "logged output holds one source per task");
    constexpr bool LOG = DT == 4;
    constexpr bool OUT = DT != 0;
    constexpr uint32_t KIND = DT == 1 ? 1u : DT == 3 ? 2u : 0u;
    constexpr uint32_t ESZ = DT == 1 ? 2u : DT == 3 ? 1u : 4u;
    constexpr uint32_t DMAX = DT == 1 ? 0xFFFEu : DT == 3 ? 0xFEu : 0xFFFFFFFFu;
    typedef typename DeepEntry<B>::T Entry;
    typedef __attribute__((address_space(3))) Entry *lq_t;
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    constexpr uint32_t VPW = 32 / B;
    // No static LDS: without a lattice the visited bits start at address 0,
    // so claim addresses need no base added.  With one, the adjacency
    // words come first and each wave's region follows.
    extern __shared__ uint32_t lds[];
    // (wave-uniform: scalar registers, so the region's addresses are too)
    const uint32_t wave = LAT ? (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64u)) : 0u;
    uint32_t *const wl = LAT ? lds + p.lat_words + wave * p.lat_region : lds;
    const uint32_t gwave = LAT ? blockIdx.x * p.lat_waves + wave : blockIdx.x;
    Entry *const s_sink = (Entry *)(wl + p.lds_sink_off);  // appends a lane does not make

    uint32_t *const vis = wl;
    __attribute__((address_space(3))) uint32_t *const lvis =
        (__attribute__((address_space(3))) uint32_t *)vis;
    const uint32_t lane = lane_id();
    Entry *const lq0 = (Entry *)(wl + p.lds_vis_words);
    Entry *const lq1 = lq0 + p.lds_qcap;
    const uint64_t gcap = p.cap - p.lds_qcap;
    Entry *const gq0 = (Entry *)p.queues + (uint64_t)gwave * 2 * gcap;
    Entry *const gq1 = gq0 + gcap;
    typedef __attribute__((address_space(1))) uint32_t *g32_t;
    typedef __attribute__((address_space(3))) uint32_t *l32_t;
    const uint32_t lbw = LOG ? p.log_bw : 0u, lmask = (1u << lbw) - 1u;
    const l32_t lbin = (l32_t)(wl + p.lds_bin_off);
    const g32_t sink = LOG ? (g32_t)p.log_sink + gwave * 64 + lane : nullptr;
    // LDS byte address of a generic pointer into LDS (its low 32 bits)
    auto lds_addr = [](const void *ptr) -> uint32_t { return (uint32_t)(uint64_t)ptr; };
    // Lattice kernel: queue entries hold biased vertex ids u + vbias, where
    // vbias = 8 x the LDS byte address of the visited bits (a multiple of
    // 32), so vertex u's visited word is at LDS byte address
    // ((u + vbias) >> 5) * 4 and its bit is (u + vbias) & 31: a claim adds
    // no base address.
    const uint32_t vbias = LAT ? 8u * lds_addr(vis) : 0u;
    // Queue entries: (vertex, source bits) packed; the lattice kernel runs
    // one source per task, so its entries are bare vertex ids.
    auto pk = [](uint32_t v, uint32_t m) -> Entry {
        if constexpr (LAT > 0) return (Entry)v;
        else return pack_entry(v, m, (Entry *)nullptr);
    };
    auto upk = [](Entry x, uint32_t &v, uint32_t &m) {
        if constexpr (LAT > 0) {
            v = (uint32_t)x;
            m = 1u;
        } else {
            unpack_entry(x, v, m);
        }
    };
    if constexpr (LAT > 0) {
        // one graph per lattice launch
        const uint32_t *const ga = p.graphs[0].lat;
        for (uint32_t i = threadIdx.x; i < p.lat_words; i += blockDim.x) lds[i] = ga[i];
        __syncthreads();
    }

    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(p.work, 1u);
        t = (uint32_t)__builtin_amdgcn_readlane((int)t, 0);
        if (t >= p.ntasks) break;
        const DeepTask task = p.tasks[t];
        const DeepGraph g = p.graphs[task.graph];
        const uint32_t V = g.V;
        // the table base is uniform: scalar registers for the whole task
        const uint64_t ell_a = (uint64_t)g.ell;
        const uint32_t *const ell_s = (const uint32_t *)(
            ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(ell_a >> 32)) << 32) |
            (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)ell_a));
        uint32_t ld[4] = {0, 0, 0, 0};  // lattice offsets (scalar registers)
        // biased id -> LDS field index: the vertex's field follows `pad`
        // padding fields in the adjacency words at the start of LDS
        uint32_t lfield = 0;
        if constexpr (LAT > 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) ld[k] = (uint32_t)__builtin_amdgcn_readfirstlane((int)g.lat_d[k]);
            const uint32_t lpad = (uint32_t)__builtin_amdgcn_readfirstlane((int)g.lat_pad);
            lfield = lpad + (32u / LAT) * (lds_addr(lds) / 4u) - vbias;
        }
        const uint32_t nwords = (V + VPW - 1) / VPW;
        for (uint32_t i = lane; i < nwords; i += 64) vis[i] = 0;
        const uint32_t nbins = LOG ? (V + lmask) >> lbw : 0u;
        if constexpr (LOG)
            for (uint32_t i = lane; i < nbins; i += 64) lbin[i] = 0;
        DeepQueue<Entry> qa{lq0, gq0, p.lds_qcap, gcap};
        DeepQueue<Entry> qb{lq1, gq1, p.lds_qcap, gcap};
        if (lane < task.nsrc) {
            const uint32_t s = p.sources[task.src_off + lane];
            const uint32_t bit = 1u << lane;
            atomicOr(&vis[s / VPW], bit << ((s % VPW) * B));
            qa.put(lane, pk(s + vbias, bit));
        }
        void *const out_task = OUT ? (void *)((char *)p.out + task.out_off * ESZ) : nullptr;
        // the lattice kernel's row, indexed by biased ids
        void *const out_lat = OUT ? (void *)((char *)out_task - (uint64_t)vbias * ESZ) : nullptr;
        uint32_t n = task.nsrc;  // entries of this level (wave-uniform)
        uint32_t level = 0;
        // logged output: the previous pass's log entry, stored once the next
        // pass's row load is in flight (its address waits on an LDS atomic,
        // and a store issued late in a pass would hold up the next pass's
        // first wait on the memory counter)
        uint32_t pend_at = 0xFFFFFFFFu, pend_pos = 0, pend_val = 0;
        // levels of at most nfast entries: one pass whose entries are all in
        // the LDS part of the queue, and whose appends (at most ELL each) fit it
        const uint32_t nfast = p.lds_qcap >= 64u ? min(64u, p.lds_qcap / ELL) : 0u;
        while (n) {
            if (OUT && level > DMAX && lane == 0) atomicOr(&p.flags[FlagU16Overflow], 1u);
            uint32_t dval = min(level, DMAX);
            uint32_t nn = 0;  // entries appended for the next level (wave-uniform)
            // One pass over entries base .. base + 63 of this level.  SPILL:
            // they may lie in the queue's global part; FIT: every append is
            // known to fit the queue's LDS part.
            auto pass = [&](uint32_t base, auto spill, auto fit) {
                const uint32_t i = base + lane;
                const bool act = i < n;
                // (vertex 0 is id vbias in the lattice kernel)
                uint32_t u = vbias, f = 0;
                if constexpr (decltype(spill)::value) {
                    if (act) upk(qa.get(i), u, f);
                } else {
                    // all lanes read (i < lcap); a lane without an entry
                    // takes vertex 0 and no source bits, so its row load
                    // is valid and it appends nothing
                    if constexpr (LAT > 0)
                        upk(*(lq_t)(size_t)(lds_addr(qa.lds) + (uint32_t)sizeof(Entry) * i), u, f);
                    else
                        upk(((lq_t)qa.lds)[i], u, f);
                    u = act ? u : vbias;
                    f = act ? f : 0u;
                }
                constexpr bool ALL = !decltype(spill)::value;
                uint32_t nb[ELL];
                if constexpr (LAT > 0) {
                    lattice_neighbours<LAT>(ld, lfield, u, nb);
                } else if (ALL || act) {
                    const __attribute__((address_space(1))) u32x4 *row =
                        (const __attribute__((address_space(1))) u32x4 *)(ell_s + (uint64_t)u * ELL);
#pragma unroll
                    for (int k = 0; k < ELL / 4; ++k) {
                        const u32x4 r = row[k];
                        nb[4 * k] = r.x;
                        nb[4 * k + 1] = r.y;
                        nb[4 * k + 2] = r.z;
                        nb[4 * k + 3] = r.w;
                    }
                }
                if constexpr (LOG) {
                    // the previous pass's log entry: one unconditional
                    // store per lane (lanes without an entry write their
                    // sink word), so the wait for this pass's row load
                    // may leave it in flight; its bin position (an LDS
                    // atomic issued late in that pass) is added only
                    // here, so no wait for it ends that pass
                    *(pend_at != 0xFFFFFFFFu ? (g32_t)out_task + pend_at + pend_pos : sink) = pend_val;
                    pend_at = 0xFFFFFFFFu;
                } else if constexpr (OUT && LAT == 0) {
                    // the entry's lowest source bit, one straight-line
                    // store per lane (a lane without an entry repeats
                    // lane 0's)
                    const uint32_t b0 = B == 1 ? 0u : __builtin_ctz(f | 0x80000000u);
                    const uint32_t u_l0 = (uint32_t)__builtin_amdgcn_readlane((int)u, 0);
                    const uint32_t b_l0 = (uint32_t)__builtin_amdgcn_readlane((int)b0, 0);
                    const uint32_t off = (act ? b0 : b_l0) * V + (act ? u : u_l0);
                    if constexpr (KIND == 1)
                        ((__attribute__((address_space(1))) uint16_t *)out_task)[off] = (uint16_t)dval;
                    else if constexpr (KIND == 2)
                        ((__attribute__((address_space(1))) uint8_t *)out_task)[off] = (uint8_t)dval;
                    else
                        ((__attribute__((address_space(1))) uint32_t *)out_task)[off] = dval;
                }
                // claims: all slots' LDS atomics, then one wait
                uint32_t newly[ELL];
                if (act) {  // (lanes without an entry would all hit vertex 0's words)
                    uint32_t old[ELL];
                    if constexpr (LAT > 0) {
                        f = 1u;  // one source per task
                        // biased ids: the visited word's LDS byte address
                        // and the bit come straight from the id
#pragma unroll
                        for (int k = 0; k < ELL; ++k)
                            old[k] = __hip_atomic_fetch_or((l32_t)(size_t)((nb[k] >> 5) * 4u), 1u << (nb[k] & 31u),
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    } else {
#pragma unroll
                        for (int k = 0; k < ELL; ++k)
                            old[k] = __hip_atomic_fetch_or(lvis + nb[k] / VPW, f << ((nb[k] % VPW) * B),
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                    if constexpr (LAT > 0 && OUT && !LOG) {
                        // the lattice kernel has no global load to keep
                        // ahead of, so its store goes in this masked block
                        if constexpr (KIND == 1)
                            ((__attribute__((address_space(1))) uint16_t *)out_lat)[u] = (uint16_t)dval;
                        else if constexpr (KIND == 2)
                            ((__attribute__((address_space(1))) uint8_t *)out_lat)[u] = (uint8_t)dval;
                        else
                            ((__attribute__((address_space(1))) uint32_t *)out_lat)[u] = dval;
                    }
                    if constexpr (LAT > 0) {
#pragma unroll
                        for (int k = 0; k < ELL; ++k) newly[k] = ~old[k] & (1u << (nb[k] & 31u));
                    } else {
#pragma unroll
                        for (int k = 0; k < ELL; ++k) newly[k] = f & ~(old[k] >> ((nb[k] % VPW) * B));
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < ELL; ++k) newly[k] = 0;
                }
                // slots, slot-major: the new entries of slot k follow those
                // of slots < k, in lane order within the slot (one ballot
                // and one v_mbcnt pair per slot; queue order is free)
                u64 bk[ELL];
                uint32_t base_k[ELL];
                uint32_t tot = 0;
#pragma unroll
                for (int k = 0; k < ELL; ++k) {
                    bk[k] = __ballot(newly[k] != 0);
                    base_k[k] = nn + tot;
                    tot += (uint32_t)__popcll(bk[k]);
                }
                if (decltype(fit)::value || nn + tot <= qb.lcap) {
                    // straight-line: an LDS byte address for every lane
                    // (the slot's run, wave-uniform, plus the lane's rank in
                    // it), then a select on the slot's ballot: the lane's
                    // sink word when it has no new entry (a masked branch
                    // per slot costs more instructions)
                    const uint32_t qaddr = lds_addr(qb.lds);
                    const uint32_t saddr = lds_addr(s_sink) + (uint32_t)sizeof(Entry) * lane;
#pragma unroll
                    for (int k = 0; k < ELL; ++k) {
                        // (the run's address is summed in scalar registers)
                        const uint32_t run = (uint32_t)__builtin_amdgcn_readfirstlane(
                            (int)(qaddr + (uint32_t)sizeof(Entry) * base_k[k]));
                        const uint32_t a = run + (uint32_t)sizeof(Entry) * mbcnt64(bk[k], 0u);
                        *(lq_t)(size_t)lane_select(bk[k], saddr, a) = pk(nb[k], newly[k]);
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < ELL; ++k) {
                        if (newly[k]) {
                            if (!qb.put(mbcnt64(bk[k], base_k[k]), pk(nb[k], newly[k])))
                                atomicOr(&p.flags[FlagQueueOverflow], 1u);
                        }
                    }
                }
                nn += tot;
                if constexpr (LOG) {
                    // append (vertex in bin, level) to the bin's log: the
                    // position is reserved after the claims and appends
                    if (act) {
                        const uint32_t ur = u - vbias;  // (lattice ids are biased)
                        const uint32_t bin = ur >> lbw;
                        pend_pos = __hip_atomic_fetch_add(lbin + bin, 1u, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_WORKGROUP);
                        pend_at = bin << lbw;
                        pend_val = (ur & lmask) | (level << lbw);
                    }
                }
                // the entry's other source bits (sources whose searches
                // reach the vertex at the same level)
                if constexpr (OUT && B > 1)
                    if (act) deep_write_bits<B, KIND>(out_task, V, u, f & (f - 1), dval);
            };
            // the level's end: the queues swap
            auto next_level = [&]() {
                DeepQueue<Entry> tq = qa;
                qa = qb;
                qb = tq;
                n = nn;
                ++level;
            };
            if (__builtin_expect(n <= nfast, 1)) {
                // the usual levels: one pass each (no spill, appends that
                // fit), in a loop of their own whose only taken branch is its
                // back edge; it ends at an empty level or a larger one
                for (;;) {
                    pass(0u, std::false_type{}, std::true_type{});
                    next_level();
                    if (n - 1u >= nfast) break;  // (n == 0 wraps)
                    if (OUT && level > DMAX && lane == 0) atomicOr(&p.flags[FlagU16Overflow], 1u);
                    dval = min(level, DMAX);
                    nn = 0;
                }
                continue;
            }
            for (uint32_t base = 0; base < n; base += 64) {
                if (base + 64 <= qa.lcap)
                    pass(base, std::false_type{}, std::false_type{});
                else
                    pass(base, std::true_type{}, std::false_type{});
            }
            next_level();
        }
        if constexpr (LOG)
            if (pend_at != 0xFFFFFFFFu) ((g32_t)out_task)[pend_at + pend_pos] = pend_val;
        if (lane == 0) atomicMax(&p.flags[FlagLevels], level);
        if (p.reach_t) {
            // as in the workgroup kernel: one ballot per 64 vertices
            for (uint32_t b = 0; b < task.nsrc; ++b) {
                uint32_t *const row = p.reach_t + (task.reach_bit + b) * (uint64_t)p.reach_vw;
                for (uint32_t v0 = 0; v0 < V; v0 += 64) {
                    const uint32_t v = v0 + lane;
                    const u64 m = __ballot(v < V && ((vis[v / VPW] >> ((v % VPW) * B + b)) & 1u));
                    if (lane == 0) row[v0 / 32] = (uint32_t)m;
                    if (lane == 32 && v0 + 32 < V) row[v0 / 32 + 1] = (uint32_t)(m >> 32);
                }
            }
        }
        // The tasks that finish last leave their rows to a separate pass, so
        // that their rewrites run on the whole device rather than one wave
        // each after the traversal's longest levels (road-like |V| = 114,599,
        // 2,048 sources: the in-kernel rewrites were 0.54 ms of 2.72).
        bool defer = false;
        if constexpr (LOG) {
            uint32_t done = 0;
            if (lane == 0) done = atomicAdd(p.work + 1, 1u);
            done = (uint32_t)__builtin_amdgcn_readlane((int)done, 0);
            if (done >= p.log_defer_after) {
                defer = true;
                uint32_t k = 0;
                if (lane == 0) k = atomicAdd(p.work + 2, 1u);
                k = (uint32_t)__builtin_amdgcn_readlane((int)k, 0);
                const uint32_t nlist = p.ntasks - p.log_defer_after;
                if (lane == 0) p.log_defer[k] = t;
                uint32_t *const counts = p.log_defer + nlist + (uint64_t)k * nbins;
                for (uint32_t c = lane; c < nbins; c += 64) counts[c] = lbin[c];
            }
        }
        if constexpr (LOG) if (!defer) {
            // The log entries were stored by this wave; make them visible
            // to its loads (the vector L1 is not kept coherent with them).
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
            const l32_t chunk = (l32_t)wl;  // visited bits and queues are free now
            // A bin holds at most 4,096 entries: a lane loads its 64 at once,
            // and the next bin's loads are issued before this bin is
            // scattered, so the rewrite costs about one memory round trip
            // per bin.
            static_assert(kDeepLogBits == 12, "64 log entries per lane and bin");
            uint32_t cur[64], nxt[64];
            auto load_bin = [&](uint32_t c, uint32_t (&e)[64]) {
                const uint32_t cnt = (uint32_t)__builtin_amdgcn_readfirstlane((int)lbin[c]);
                const g32_t span = (g32_t)out_task + (c << lbw);
#pragma unroll
                for (int q = 0; q < 64; ++q) {
                    const uint32_t j = q * 64 + lane;
                    e[q] = j < cnt ? span[j] : 0xFFFFFFFFu;
                }
            };
            auto rewrite_bin = [&](uint32_t c, const uint32_t (&e)[64]) {
                const uint32_t c0 = c << lbw;
                const uint32_t len = min(lmask + 1u, V - c0);
                for (uint32_t x = lane; x < len; x += 64) chunk[x] = 0xFFFFFFFFu;
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int q = 0; q < 64; ++q)
                    if (e[q] != 0xFFFFFFFFu) chunk[e[q] & lmask] = e[q] >> lbw;
                __builtin_amdgcn_wave_barrier();
                uint32_t *const span = (uint32_t *)out_task + c0;
                for (uint32_t x = lane; x < len; x += 64) __builtin_nontemporal_store(chunk[x], span + x);
                __builtin_amdgcn_wave_barrier();
            };
            if (nbins) load_bin(0, cur);
            for (uint32_t c = 0; c < nbins; c += 2) {
                if (c + 1 < nbins) load_bin(c + 1, nxt);
                rewrite_bin(c, cur);
                if (c + 1 < nbins) {
                    if (c + 2 < nbins) load_bin(c + 2, cur);
                    rewrite_bin(c + 1, nxt);
                }
            }
        }
        if constexpr (!LOG && OUT) {
            // sentinels for unreached pairs, skipping runs of 64 full visited words
            for (uint32_t i0 = 0; i0 < nwords; i0 += 64) {
                const uint32_t i = i0 + lane;
                if (__all(i >= nwords || vis[i] == 0xFFFFFFFFu)) continue;
                const uint32_t v1 = min(V, (i0 + 64) * VPW);
                for (uint32_t b = 0; b < task.nsrc; ++b) {
                    const uint64_t row = task.out_off + (uint64_t)b * V;
                    for (uint32_t v = i0 * VPW + lane; v < v1; v += 64)
                        if (!((vis[v / VPW] >> ((v % VPW) * B + b)) & 1u))
                            store_sentinel(p.out, KIND, row + v);
                }
            }
        }
    }
}

// Rewrite of the logged rows listed by deep_wave_kernel: one workgroup per
// (listed task, bin).  The bin's log (at most 2^log_bw entries of (vertex
// in bin, level), the span's own first elements) is loaded, the span's
// distances are built in LDS over a sentinel fill, and the span is stored
// whole.  Workgroups beyond the listed tasks return at once.
constexpr uint32_t kLogRewriteThreads = 256;
__global__ void __launch_bounds__(kLogRewriteThreads) deep_log_rewrite_kernel(DeepParams p,
                                                                              uint32_t nbins) {
    __shared__ uint32_t chunk[1u << kDeepLogBits];
    constexpr uint32_t PER = (1u << kDeepLogBits) / kLogRewriteThreads;
    const uint32_t k = blockIdx.x / nbins, c = blockIdx.x % nbins;
    if (k >= p.work[2]) return;  // uniform (written by the traversal's launch)
    const uint32_t nlist = p.ntasks - p.log_defer_after;
    const uint32_t t = p.log_defer[k];
    const uint32_t cnt = p.log_defer[nlist + (uint64_t)k * nbins + c];
    const DeepTask task = p.tasks[t];
    const uint32_t V = p.graphs[task.graph].V;
    const uint32_t lbw = p.log_bw, lmask = (1u << lbw) - 1u;
    const uint32_t c0 = c << lbw, len = min(lmask + 1u, V - c0);
    uint32_t *const span = (uint32_t *)p.out + task.out_off + c0;
    uint32_t e[PER];
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t j = q * kLogRewriteThreads + threadIdx.x;
        e[q] = j < cnt ? span[j] : 0xFFFFFFFFu;
    }
    for (uint32_t x = threadIdx.x; x < len; x += kLogRewriteThreads) chunk[x] = 0xFFFFFFFFu;
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q)
        if (e[q] != 0xFFFFFFFFu) chunk[e[q] & lmask] = e[q] >> lbw;
    __syncthreads();
    for (uint32_t x = threadIdx.x; x < len; x += kLogRewriteThreads)
        __builtin_nontemporal_store(chunk[x], span + x);
}

hipError_t launch_log_rewrite(const DeepParams &p, uint32_t nbins, hipStream_t st) {
    const uint64_t blocks = (uint64_t)(p.ntasks - p.log_defer_after) * nbins;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(deep_log_rewrite_kernel, dim3((unsigned)blocks), dim3(kLogRewriteThreads), 0,
                       st, p, nbins);
    return hipGetLastError();
}

// The deep kernel for a launch: the one-wave kernel for 64-thread
// workgroups on a neighbour table, the workgroup kernel otherwise.
// DT = 4 (logged u32 rows) exists only as the one-wave kernel with one
// source per task; it is null otherwise.
template <int B, int ELL, int DT>
static const void *deep_kernel_for(int threads) {
    if constexpr (DT == 5 || DT == 6) {
        if constexpr (B == 1) return (const void *)deep_bfs_kernel<1, ELL, DT>;
        return nullptr;
    } else if constexpr (DT == 4) {
        if constexpr (ELL > 0 && B == 1)
            if (threads == 64) return (const void *)deep_wave_kernel<B, ELL, 4>;
        return nullptr;
    } else {
        if constexpr (ELL > 0)
            if (threads == 64) return (const void *)deep_wave_kernel<B, ELL, DT>;
        return (const void *)deep_bfs_kernel<B, ELL, DT>;
    }
}

template <int B, int ELL>
static hipError_t deep_launch_t(const DeepParams &p, int blocks, int threads, size_t lds,
                                hipStream_t st) {
    const int dt = !p.out               ? 0
                   : p.out_kind == 1    ? (p.staged ? 6 : 1)
                   : p.out_kind == 2    ? 3
                   : p.log_bw           ? 4
                   : p.staged           ? 5
                                        : 2;
    const void *k = dt == 0   ? deep_kernel_for<B, ELL, 0>(threads)
                    : dt == 1 ? deep_kernel_for<B, ELL, 1>(threads)
                    : dt == 3 ? deep_kernel_for<B, ELL, 3>(threads)
                    : dt == 4 ? deep_kernel_for<B, ELL, 4>(threads)
                    : dt == 5 ? deep_kernel_for<B, ELL, 5>(threads)
                    : dt == 6 ? deep_kernel_for<B, ELL, 6>(threads)
                              : deep_kernel_for<B, ELL, 2>(threads);
    if (!k) return hipErrorInvalidConfiguration;
    void *args[] = {(void *)&p};
    return hipLaunchKernel(k, dim3(blocks), dim3(threads), args, lds, st);
}

template <int B, int ELL, int DT>
static hipError_t deep_prepare_one(const void *k, size_t lds) {
    hipFuncAttributes a;
    hipError_t e = hipFuncGetAttributes(&a, k);
    if (e != hipSuccess) return e;
    if (a.sharedSizeBytes > kDeepStaticLds) return hipErrorInvalidConfiguration;
    return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

// Only the kernel shape that will run (one-wave or workgroup) is prepared:
// their static LDS differs, so one dynamic size need not suit both.
template <int B, int ELL, int DT>
static hipError_t deep_prepare_dt(int threads, size_t lds) {
    const void *k = deep_kernel_for<B, ELL, DT>(threads);
    return k ? deep_prepare_one<B, ELL, DT>(k, lds) : hipSuccess;
}

template <int B, int ELL>
static hipError_t deep_prepare_t(int threads, size_t lds) {
    hipError_t e = deep_prepare_dt<B, ELL, 0>(threads, lds);
    if (e == hipSuccess) e = deep_prepare_dt<B, ELL, 1>(threads, lds);
    if (e == hipSuccess) e = deep_prepare_dt<B, ELL, 2>(threads, lds);
    if (e == hipSuccess) e = deep_prepare_dt<B, ELL, 3>(threads, lds);
    if (e == hipSuccess) e = deep_prepare_dt<B, ELL, 4>(threads, lds);
    if (e == hipSuccess) e = deep_prepare_dt<B, ELL, 5>(threads, lds);
    if (e == hipSuccess) e = deep_prepare_dt<B, ELL, 6>(threads, lds);
    return e;
}

// Resident workgroups per CU: the least over the output variants.
template <int B, int ELL>
static hipError_t deep_occ_t(int threads, size_t lds, int *out) {
    constexpr int NK = 7;
    int o[NK];
    for (int i = 0; i < NK; ++i) o[i] = 1 << 30;
    const void *k[NK] = {deep_kernel_for<B, ELL, 0>(threads), deep_kernel_for<B, ELL, 1>(threads),
                         deep_kernel_for<B, ELL, 2>(threads), deep_kernel_for<B, ELL, 3>(threads),
                         deep_kernel_for<B, ELL, 4>(threads), deep_kernel_for<B, ELL, 5>(threads),
                         deep_kernel_for<B, ELL, 6>(threads)};
    hipError_t e = hipSuccess;
    for (int i = 0; i < NK && e == hipSuccess; ++i)
        if (k[i]) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&o[i], k[i], threads, lds);
    *out = *std::min_element(o, o + NK);
    return e;
}

// Dispatches a templated call over bits in {1,2,4,8,16,32} x ell in {0,4,8}.
#define DEEP_SWITCH(FN, ...)                                                 \
    switch (bits * 16 + ell) {                                               \
    case 1 * 16 + 0: return FN<1, 0>(__VA_ARGS__);                           \
    case 1 * 16 + 4: return FN<1, 4>(__VA_ARGS__);                           \
    case 1 * 16 + 8: return FN<1, 8>(__VA_ARGS__);                           \
    case 2 * 16 + 0: return FN<2, 0>(__VA_ARGS__);                           \
    case 2 * 16 + 4: return FN<2, 4>(__VA_ARGS__);                           \
    case 2 * 16 + 8: return FN<2, 8>(__VA_ARGS__);                           \
    case 4 * 16 + 0: return FN<4, 0>(__VA_ARGS__);                           \
    case 4 * 16 + 4: return FN<4, 4>(__VA_ARGS__);                           \
    case 4 * 16 + 8: return FN<4, 8>(__VA_ARGS__);                           \
    case 8 * 16 + 0: return FN<8, 0>(__VA_ARGS__);                           \
    case 8 * 16 + 4: return FN<8, 4>(__VA_ARGS__);                           \
    case 8 * 16 + 8: return FN<8, 8>(__VA_ARGS__);                           \
    case 16 * 16 + 0: return FN<16, 0>(__VA_ARGS__);                         \
    case 16 * 16 + 4: return FN<16, 4>(__VA_ARGS__);                         \
    case 16 * 16 + 8: return FN<16, 8>(__VA_ARGS__);                         \
    case 32 * 16 + 0: return FN<32, 0>(__VA_ARGS__);                         \
    case 32 * 16 + 4: return FN<32, 4>(__VA_ARGS__);                         \
    case 32 * 16 + 8: return FN<32, 8>(__VA_ARGS__);                         \
    default: return hipErrorInvalidValue;                                    \
    }

hipError_t launch_deep(int bits, int ell, const DeepParams &p, int blocks, int threads,
                       size_t lds, hipStream_t st) {
    DEEP_SWITCH(deep_launch_t, p, blocks, threads, lds, st)
}

hipError_t deep_prepare(int bits, int ell, int threads, size_t lds) {
    DEEP_SWITCH(deep_prepare_t, threads, lds)
}

hipError_t deep_max_blocks_per_cu(int bits, int ell, int threads, size_t lds, int *out) {
    DEEP_SWITCH(deep_occ_t, threads, lds, out)
}

// Lattice kernel variants by output: 0 none, 1 u16, 2 u32, 3 u8, 4 logged u32.
template <int LAT>
static const void *deep_lattice_kernel(int dt) {
    switch (dt) {
    case 0: return (const void *)deep_wave_kernel<1, 2 * LAT, 0, LAT>;
    case 1: return (const void *)deep_wave_kernel<1, 2 * LAT, 1, LAT>;
    case 3: return (const void *)deep_wave_kernel<1, 2 * LAT, 3, LAT>;
    case 4: return (const void *)deep_wave_kernel<1, 2 * LAT, 4, LAT>;
    default: return (const void *)deep_wave_kernel<1, 2 * LAT, 2, LAT>;
    }
}

static const void *deep_lattice_kernel_any(int lat, int dt) {
    return lat == 2 ? deep_lattice_kernel<2>(dt) : lat == 4 ? deep_lattice_kernel<4>(dt) : nullptr;
}

hipError_t launch_deep_lattice(int lat, const DeepParams &p, int blocks, size_t lds,
                               hipStream_t st) {
    const int dt = !p.out ? 0 : p.out_kind == 1 ? 1 : p.out_kind == 2 ? 3 : p.log_bw ? 4 : 2;
    const void *k = deep_lattice_kernel_any(lat, dt);
    if (!k || p.lat_waves < 1 || p.lat_waves > (uint32_t)kLatticeMaxWaves)
        return hipErrorInvalidConfiguration;
    void *args[] = {(void *)&p};
    return hipLaunchKernel(k, dim3(blocks), dim3(64 * p.lat_waves), args, lds, st);
}

hipError_t deep_lattice_prepare(int lat, int waves, size_t lds) {
    (void)waves;
    for (int dt = 0; dt <= 4; ++dt) {
        const void *k = deep_lattice_kernel_any(lat, dt);
        if (!k) return hipErrorInvalidValue;
        hipFuncAttributes a;
        hipError_t e = hipFuncGetAttributes(&a, k);
        if (e != hipSuccess) return e;
        if (a.sharedSizeBytes != 0) return hipErrorInvalidConfiguration;
        e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t deep_lattice_blocks_per_cu(int lat, int waves, size_t lds, int *out) {
    int best = 1 << 30;
    for (int dt = 0; dt <= 4; ++dt) {
        const void *k = deep_lattice_kernel_any(lat, dt);
        if (!k) return hipErrorInvalidValue;
        int o = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k, 64 * waves, lds);
        if (e != hipSuccess) return e;
        best = std::min(best, o);
    }
    *out = best;
    return hipSuccess;
}

// ---------------------------------------------------------------- wide family
//
// State is vertex-major: the W source words of vertex v are the contiguous
// run [v*W, v*W + W), so a parent's whole frontier (W*8 bytes, 256 B at
// S = 2048) is one coalesced read by the TW lanes that share a vertex.
// Wave tile: TV = 64/TW vertices x TW lanes (lane = wi*TV + vi), one word
// per lane, or two adjacent words per lane in deferred pull levels when the
// parameters ask for it (WideParams::pair).

__device__ __forceinline__ u64 word_valid(uint64_t S, uint32_t w) {
    const uint64_t lo = (uint64_t)w * 64;
    if (lo + 64 <= S) return ~0ull;
    if (lo >= S) return 0ull;
    return (1ull << (S - lo)) - 1ull;
}

__device__ __forceinline__ uint32_t ld_dev_u32(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 ld_dev_u64(const u64 *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// OR over the TV lanes that share a word (lanes wi*TV .. wi*TV + TV-1).
template <uint32_t TV>
__device__ __forceinline__ u64 group_or64(u64 x) {
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
#pragma unroll
    for (uint32_t off = 1; off < TV; off <<= 1) {
        lo |= (uint32_t)__shfl_xor((int)lo, (int)off);
        hi |= (uint32_t)__shfl_xor((int)hi, (int)off);
    }
    return ((u64)hi << 32) | lo;
}

// A lane's source words in the pull: one (u64) or two adjacent ones (u64x2,
// 16-byte loads; the source word count is even).
typedef u64 u64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bool nz(u64 a) { return a != 0; }
__device__ __forceinline__ bool nz(u64x2 a) { return (a.x | a.y) != 0; }
__device__ __forceinline__ uint32_t popc(u64 a) { return (uint32_t)__popcll(a); }
__device__ __forceinline__ uint32_t popc(u64x2 a) {
    return (uint32_t)__popcll(a.x) + (uint32_t)__popcll(a.y);
}
template <uint32_t TV>
__device__ __forceinline__ u64 group_or(u64 a) { return group_or64<TV>(a); }
template <uint32_t TV>
__device__ __forceinline__ u64x2 group_or(u64x2 a) {
    u64x2 r;
    r.x = group_or64<TV>(a.x);
    r.y = group_or64<TV>(a.y);
    return r;
}
// All-ones and zero words.
__device__ __forceinline__ void wfill(u64 &a, u64 v) { a = v; }
__device__ __forceinline__ void wfill(u64x2 &a, u64 v) {
    a.x = v;
    a.y = v;
}
template <typename WT>
__device__ __forceinline__ WT wconst(u64 v) {
    WT a;
    wfill(a, v);
    return a;
}
// Valid and active source bits of the lane's words (sb: the vertex's source
// count in a batch union traversal, 0 for all S sources).
__device__ __forceinline__ u64 words_live(const WideParams &p, const u64 *act, uint32_t w, u64,
                                          uint32_t sb = 0) {
    return word_valid(sb ? sb : p.S, w) & act[w];
}
__device__ __forceinline__ u64x2 words_live(const WideParams &p, const u64 *act, uint32_t w, u64x2,
                                            uint32_t sb = 0) {
    const uint64_t lim = sb ? sb : p.S;
    u64x2 r;
    r.x = word_valid(lim, w) & act[w];
    r.y = word_valid(lim, w + 1) & act[w + 1];
    return r;
}
// Valid source bits of the lane's words.
__device__ __forceinline__ u64 words_valid(const WideParams &p, uint32_t w, u64) {
    return word_valid(p.S, w);
}
__device__ __forceinline__ u64x2 words_valid(const WideParams &p, uint32_t w, u64x2) {
    u64x2 r;
    r.x = word_valid(p.S, w);
    r.y = word_valid(p.S, w + 1);
    return r;
}
__device__ __forceinline__ void words_or(u64 *a, u64x2 n) {
    if (n.x) atomicOr(&a[0], n.x);
    if (n.y) atomicOr(&a[1], n.y);
}
__device__ __forceinline__ void act_or(u64 *act_lds, uint32_t w, u64 n) {
    if (n) atomicOr(&act_lds[w], n);
}
__device__ __forceinline__ void act_or(u64 *act_lds, uint32_t w, u64x2 n) {
    if (n.x) atomicOr(&act_lds[w], n.x);
    if (n.y) atomicOr(&act_lds[w + 1], n.y);
}

// Per-lane distance stores for scattered vertices (push filter, heavy
// vertices): one store per set bit.
// Output element of (source s, vertex v): row s of the source-major output,
// or, in a batch union traversal, row s of v's own graph.
__device__ __forceinline__ uint64_t dist_index(const WideParams &p, uint64_t s, uint32_t v) {
    if (!p.bmap) return s * p.V + v;
    const BatchGraphMap g = p.bmap[p.vgraph[v]];
    return g.out_off + s * g.V + (v - g.base);
}

__device__ __forceinline__ void wide_write_lane(const WideParams &p, uint32_t w, uint32_t v,
                                                u64 n, uint32_t d) {
    const uint32_t u16 = p.out_kind;
    while (n) {
        const uint32_t b = (uint32_t)__builtin_ctzll(n);
        n &= n - 1;
        store_dist(p.out, u16, dist_index(p, (uint64_t)w * 64 + b, v), d, p.flags);
    }
}

// Seed i: vertex sources[i] with source bit i, or, in a batch union
// traversal, vertex seed_v[i] with bit seed_b[i].
__global__ void wide_seed_kernel(WideParams p) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= (p.seed_v ? (uint64_t)p.nseeds : p.S)) return;
    const uint32_t s = p.seed_v ? p.seed_v[i] : p.sources[i];
    const uint64_t sb = p.seed_v ? (uint64_t)p.seed_b[i] : i;
    const uint32_t w = (uint32_t)(sb / 64);
    const u64 bit = 1ull << (sb % 64);
    const uint64_t idx = (uint64_t)s * p.W + w;
    atomicOr(&p.seen[idx], bit);
    atomicOr(&p.fr[idx], bit);  // slot of distance 0
    atomicOr(&p.fbA[s >> 5], 1u << (s & 31));
    if (p.defer) atomicOr(&p.lv[s], 1u);  // in the frontier of level 0
    if (p.out && p.defer == 0) store_dist(p.out, p.out_kind, dist_index(p, sb, s), 0, p.flags);
    if (i < p.W) {  // active sources: all of them at level 0
        p.act[i] = word_valid(p.S, (uint32_t)i);
        p.act[p.W + i] = 0;
    }
    if (atomicExch(&p.stamp[s], p.tick_base) != p.tick_base) {
        const uint32_t pos = atomicAdd(&p.ctr[3].next_count, 1u);
        p.qf0[pos] = s;
    }
}

// The level-0 slot rows of the sources' vertices, cleared before the seed
// kernel ORs their bits in (frontier slots are not cleared after use).
__global__ void wide_seed_rows_kernel(WideParams p) {
    const uint64_t n = (p.seed_v ? (uint64_t)p.nseeds : p.S) * p.W;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = p.seed_v ? p.seed_v[i / p.W] : p.sources[i / p.W];
        p.fr[(uint64_t)v * p.W + i % p.W] = 0;
    }
}

// Source-major reachability rows (reach_vw words per source) to the
// vertex-major words reach[v * wpv + s / 32]: one wave per (pair of row
// words, source word); lane j holds row word (2 vw2 + j / 32) of source
// 32 w + j % 32, and a ballot per bit position k collects the 32 sources of
// vertices 64 vw2 + k and 64 vw2 + 32 + k.
__global__ void reach_transpose_kernel(const uint32_t *rt, uint32_t vw_n, uint64_t S, uint32_t V,
                                       uint32_t *reach, uint32_t wpv) {
    const uint64_t gw = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / 64;
    const uint32_t lane = lane_id();
    const uint64_t nvw2 = (vw_n + 1) / 2;
    if (gw >= nvw2 * wpv) return;  // wave-uniform
    const uint32_t w = (uint32_t)(gw % wpv), vw2 = (uint32_t)(gw / wpv);
    const uint64_t s = (uint64_t)w * 32 + (lane & 31);
    const uint32_t vw = 2 * vw2 + lane / 32;
    const uint32_t word = (s < S && vw < vw_n) ? rt[s * vw_n + vw] : 0u;
    uint32_t mine = 0;
#pragma unroll
    for (uint32_t k = 0; k < 32; ++k) {
        const u64 m = __ballot((word >> k) & 1u);
        if (lane == k) mine = (uint32_t)m;
        if (lane == 32 + k) mine = (uint32_t)(m >> 32);
    }
    const uint32_t v = vw2 * 64 + lane;
    if (v < V) reach[(uint64_t)v * wpv + w] = mine;
}

hipError_t launch_reach_transpose(const uint32_t *rt, uint32_t vw_n, uint64_t S, uint32_t V,
                                  uint32_t *reach, uint32_t wpv, hipStream_t st) {
    const uint64_t waves = (uint64_t)((vw_n + 1) / 2) * wpv;
    const uint64_t blocks = (waves + 3) / 4;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(reach_transpose_kernel, dim3((unsigned)blocks), dim3(256), 0, st, rt, vw_n,
                       S, V, reach, wpv);
    return hipGetLastError();
}

hipError_t launch_wide_seed(const WideParams &p, hipStream_t st) {
    const uint64_t nseed = p.seed_v ? (uint64_t)p.nseeds : p.S;
    const uint64_t rblocks = std::max<uint64_t>(1, std::min<uint64_t>((nseed * p.W + 255) / 256, 4096));
    hipLaunchKernelGGL(wide_seed_rows_kernel, dim3((unsigned)rblocks), dim3(256), 0, st, p);
    const uint64_t blocks = std::max<uint64_t>(1, (nseed + 255) / 256);
    hipLaunchKernelGGL(wide_seed_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p);
    return hipGetLastError();
}

enum { DirPush = 0, DirPull = 1 };

// Per-wave indices shared by the phases of the persistent kernel.
struct WaveCtx {
    uint32_t lane, vi, wi;
    uint64_t gwave, nwaves;
    u64 *act_lds;        // this workgroup's new-bit sources of the level (LDS, W words)
    const u64 *act_cur;  // sources whose frontier is not empty at this level (W words)
    uint32_t *blk;       // LDS: pull block's new frontier vertices: [0] count, [1..3]
                         // bitmap words from the block's first, [4..68) the vertices
    bool direct;         // this level's new distances are stored now (not deferred)
    uint32_t *wq;        // this wave's 64 LDS slots for WaveAppender
    uint32_t *par;       // this wave's 64 LDS slots: each lane group's frontier parents of a step
};

// A lane group's frontier parents of one scan step, listed in LDS: the lane
// that loaded in-edge parent u (frontier bit set) writes u at its rank
// among the group's such lanes, so a gather round reads eight parents with
// two 16-byte LDS reads instead of finding the ballot's next bit and
// broadcasting it from its lane eight times.  `gl` = the group's TW slots;
// returns the group's count.  (A wave's LDS operations execute in order,
// so the list is read after it is written without a wait.)
__device__ __forceinline__ uint32_t list_parents(uint32_t *gl, bool inf, uint32_t u, u64 gmask) {
    const u64 m = __ballot(inf) & gmask;
    if (inf) gl[mbcnt64(m, 0)] = u;
    return (uint32_t)__popcll(m);
}

// Eight parents of a list from LDS (slots r .. r+7; r a multiple of 8).
__device__ __forceinline__ void read_parents(const uint32_t *gl, uint32_t r, uint32_t u[8]) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const __attribute__((address_space(3))) u32x4 *q =
        (const __attribute__((address_space(3))) u32x4 *)(gl + r);
    const u32x4 a = q[0], b = q[1];
    u[0] = a.x; u[1] = a.y; u[2] = a.z; u[3] = a.w;
    u[4] = b.x; u[5] = b.y; u[6] = b.z; u[7] = b.w;
}

// Appends to a global queue through the wave's LDS slots: one atomic on the
// queue counter per 64 entries instead of one per call (the counters are
// single words that every wave of the grid hits).  Every lane of the wave
// calls push and flush; flush before the queue is read.
struct WaveAppender {
    uint32_t *buf;
    uint32_t n = 0;  // entries held, wave-uniform
    __device__ __forceinline__ void flush(uint32_t *ctr, uint32_t *queue) {
        if (n == 0) return;
        const int leader = __ffsll((long long)__ballot(1)) - 1;
        uint32_t base = 0;
        if ((int)lane_id() == leader) base = atomicAdd(ctr, n);
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
        if (lane_id() < n) queue[base + lane_id()] = buf[lane_id()];
        n = 0;
    }
    __device__ __forceinline__ void push(bool pred, uint32_t v, uint32_t *ctr, uint32_t *queue) {
        const u64 bal = __ballot(pred);
        if (!bal) return;
        const uint32_t c = (uint32_t)__popcll(bal);
        if (n + c > 64) flush(ctr, queue);
        const u64 below = lane_id() ? (bal & ((~0ull) >> (64 - lane_id()))) : 0ull;
        if (pred) buf[n + (uint32_t)__popcll(below)] = v;
        n += c;
    }
};

__device__ __forceinline__ bool in_frontier(const uint32_t *fb, uint32_t u) {
    return (fb[u >> 5] >> (u & 31)) & 1u;
}
__device__ __forceinline__ void mark_frontier(uint32_t *fb, uint32_t v) {
    atomicOr(&fb[v >> 5], 1u << (v & 31));
}
// Clears the bitmap of a consumed frontier (vertices qf[0, n)): the words
// of its vertices, or the whole array when the frontier outnumbers its words.
__device__ __forceinline__ void clear_frontier(uint32_t *fb, const uint32_t *qf, uint32_t n,
                                               uint32_t V, uint64_t gthread, uint64_t nthreads) {
    const uint64_t nfw = (uint64_t)V / 32 + 1;
    if (n > nfw)
        for (uint64_t k = gthread; k < nfw; k += nthreads) fb[k] = 0;
    else
        for (uint64_t k = gthread; k < n; k += nthreads) fb[qf[k] >> 5] = 0;
}

// Push, part 1: OR each frontier vertex's words into its children's
// candidate words (64-bit atomicOr), skipping children whose bits are all
// seen; the first touch of a child appends it to the candidate queue.
// Light vertices: TW lanes per vertex walk its edges together.  Heavy
// vertices: their out-lists come in chunks of kHeavyChunkOut edges, one wave
// per chunk, the vertex's active words one at a time, 64 edges per step.
template <int TW>
__device__ __forceinline__ void wide_push_expand(const WideParams &p, const WaveCtx &x, u64 *cur,
                                                 const uint32_t *qcur, uint32_t n_front,
                                                 const uint32_t *fb_cur, WideCounters *c,
                                                 uint32_t tick) {
    constexpr uint32_t TV = 64 / TW;
    const uint32_t W = p.W, WG = (W + TW - 1) / TW;
    WaveAppender qa{x.wq};
    // The whole wave ORs vertex h's active words into the children along
    // out-edges [e0, e1), 64 edges per step, one word at a time.
    auto walk = [&](uint32_t h, uint32_t e0, uint32_t e1) {
        for (uint32_t w0 = 0; w0 < W; w0 += 64) {
            const uint32_t wl = w0 + x.lane;
            const u64 fl = wl < W ? cur[(uint64_t)h * W + wl] : 0ull;
            u64 act = __ballot(fl != 0);
            while (act) {
                const uint32_t k = (uint32_t)__builtin_ctzll(act);
                act &= act - 1;
                const uint32_t w = w0 + k;
                const u64 f = __shfl(fl, (int)k);
                for (uint32_t base = e0; base < e1; base += 64) {
                    const uint32_t e = base + x.lane;
                    bool app = false;
                    uint32_t v = 0;
                    if (e < e1) {
                        v = p.ci[e];
                        const uint64_t idx = (uint64_t)v * W + w;
                        const u64 m = f & ~p.seen[idx];
                        // Only the first toucher of (v, w) can be the first
                        // toucher of v, so only it consults the level stamp.
                        if (m && atomicOr(&p.nx[idx], m) == 0 &&
                            atomicExch(&p.stamp[v], tick) != tick)
                            app = true;
                    }
                    qa.push(app, v, &c->cand_count, p.qc);
                }
            }
        }
    };
    if (n_front <= x.nwaves) {
        // Small frontier (sources, last levels): a wave per light vertex, so
        // its edges are not walked one by one by a single lane.
        for (uint64_t i = x.gwave; i < n_front; i += x.nwaves) {
            const uint32_t u = qcur[i];
            const uint32_t e0 = p.ro[u], e1 = p.ro[u + 1];
            if (e1 - e0 <= p.heavy_t) walk(u, e0, e1);  // heavy: chunks below
        }
    } else {
    const uint64_t ntiles = (uint64_t)((n_front + TV - 1) / TV) * WG;
    for (uint64_t t = x.gwave; t < ntiles; t += x.nwaves) {
        const uint32_t i = (uint32_t)(t / WG) * TV + x.vi;
        const uint32_t w = (uint32_t)(t % WG) * TW + x.wi;
        uint32_t u = 0, e = 0, eend = 0;
        u64 f = 0;
        if (i < n_front && w < W) {
            u = qcur[i];
            const uint32_t e0 = p.ro[u], e1 = p.ro[u + 1];
            if (e1 - e0 <= p.heavy_t) {
                const uint64_t idx = (uint64_t)u * W + w;
                f = cur[idx];
                if (f) {
                    e = e0;
                    eend = e1;
                }
            }
        }
        while (__any(e < eend)) {
            bool app = false;
            uint32_t v = 0;
            if (e < eend) {
                v = p.ci[e];
                const uint64_t idx = (uint64_t)v * W + w;
                const u64 m = f & ~p.seen[idx];
                // Only the first toucher of (v, w) can be the first toucher
                // of v, so only it consults the level stamp.
                if (m && atomicOr(&p.nx[idx], m) == 0 &&
                    atomicExch(&p.stamp[v], tick) != tick)
                    app = true;
                ++e;
            }
            qa.push(app, v, &c->cand_count, p.qc);
        }
    }
    }
    // Heavy out-lists: chunk descriptors are scanned 64 at a time, lane =
    // chunk, with a hub's consecutive chunks on different waves; the wave
    // then walks each chunk whose vertex is in the frontier.
    const uint64_t nhc = p.n_hchunk_out;
    for (uint64_t it = 0; x.gwave + it * 64 * x.nwaves < nhc; ++it) {
        const uint64_t tl = x.gwave + (it * 64 + x.lane) * x.nwaves;
        uint32_t cx = 0, cy = 0, cz = 0;
        bool cact = false;
        if (tl < nhc) {
            const uint4 ch = p.hchunk_out[tl];
            cx = ch.x;
            cy = ch.y;
            cz = ch.z;
            cact = in_frontier(fb_cur, cx);
        }
        u64 am = __ballot(cact);
        while (am) {
            const int kc = (int)__builtin_ctzll(am);
            am &= am - 1;
            walk((uint32_t)__builtin_amdgcn_readlane((int)cx, kc),
                 (uint32_t)__builtin_amdgcn_readlane((int)cy, kc),
                 (uint32_t)__builtin_amdgcn_readlane((int)cz, kc));
        }
    }
    qa.flush(&c->cand_count, p.qc);
}

// Push, part 2: new bits of each candidate become its next-frontier words
// (all of them, zeros included: the slot is not cleared between uses) and
// distances; the candidate joins the next frontier queue.  The consumed
// frontier's bitmap is cleared here.
template <int TW>
__device__ __forceinline__ void wide_push_filter(const WideParams &p, const WaveCtx &x, u64 *cur,
                                                 const uint32_t *qcur, uint32_t n_front,
                                                 uint32_t *fb_cur, u64 *nxt, uint32_t *qnext,
                                                 uint32_t *fb_next, WideCounters *c, uint32_t nd,
                                                 uint64_t gthread, uint64_t nthreads,
                                                 u64 &my_bits, u64 &my_deg) {
    constexpr uint32_t TV = 64 / TW;
    const uint32_t W = p.W, WG = (W + TW - 1) / TW;
    const uint32_t ncand = ld_dev_u32(&c->cand_count);
    const uint64_t ftiles = (ncand + TV - 1) / TV;
    WaveAppender qa{x.wq};
    for (uint64_t t = x.gwave; t < ftiles; t += x.nwaves) {
        const uint32_t i = (uint32_t)t * TV + x.vi;
        const bool okv = i < ncand;
        const uint32_t v = okv ? p.qc[i] : 0;
        bool any = false;
        for (uint32_t g = 0; g < WG; ++g) {
            const uint32_t w = g * TW + x.wi;
            u64 n = 0;
            if (okv && w < W) {
                const uint64_t idx = (uint64_t)v * W + w;
                const u64 xw = p.nx[idx];
                if (xw) {
                    p.nx[idx] = 0;
                    const u64 s = p.seen[idx];
                    n = xw & ~s;
                    if (n) {
                        atomicOr(&x.act_lds[w], n);
                        p.seen[idx] = s | n;
                        if (x.direct) wide_write_lane(p, w, v, n, nd);
                    }
                }
                nxt[idx] = n;
            }
            any |= n != 0;
            my_bits += __popcll(n);
        }
        const u64 bal = __ballot(any);
        bool vany = false;
#pragma unroll
        for (uint32_t k = 0; k < TW; ++k) vany |= ((bal >> (k * TV + x.vi)) & 1ull) != 0;
        const bool app = x.wi == 0 && okv && vany;
        qa.push(app, v, &c->next_count, qnext);
        if (app) {
            mark_frontier(fb_next, v);
            if (nd < p.defer) p.lv[v] |= 1u << nd;  // one writer per vertex and level
            my_deg += p.ro[v + 1] - p.ro[v];
        }
    }
    qa.flush(&c->next_count, qnext);
    clear_frontier(fb_cur, qcur, n_front, p.V, gthread, nthreads);
}

// Scan-then-gather step for tiles of TW >= 8 words (TV = 64/TW lane
// groups, group vi = lanes wi*TV + vi): each lane of a group loads one
// in-edge of [ebase, eend) (TW edges per group), tests its frontier bit, and
// the group's ballot lists the parents in the frontier; their frontier rows
// are then gathered eight at a time (the gathers are latency-bound: on
// R-MAT20 S=2048 eight per round beat four by 7%, twelve and sixteen spill
// registers and lose), parent index broadcast from the lane that loaded it.
// Returns the OR of this lane's word over those parents.
// need / have: the lane's word's still-needed bits and those already found;
// a group whose words all have their needed bits stops after the round.
template <int TW, typename WT>
__device__ __forceinline__ WT scan_gather(const WideParams &p, const WaveCtx &x, const u64 *cur,
                                          const uint32_t *fb, uint32_t ebase, uint32_t eend,
                                          uint32_t w, bool wok, WT need, WT have) {
    constexpr uint32_t TV = 64 / TW;
    static_assert(TV <= 8, "scan_gather needs tiles of 8 or more words");
    typedef const __attribute__((address_space(1))) uint32_t *g32;
    typedef const __attribute__((address_space(1))) WT *gw;
    const u64 gmask = (~0ull / ((1ull << TV) - 1ull)) << x.vi;
    const uint32_t e = ebase + x.wi;
    uint32_t uu = 0;
    bool inf = false;
    if (e < eend) {
        uu = ((g32)p.ri)[e];
        inf = (((g32)fb)[uu >> 5] >> (uu & 31)) & 1u;
    }
    uint32_t *const gl = x.par + x.vi * TW;
    uint32_t cnt = list_parents(gl, inf, uu, gmask);
    WT acc = wconst<WT>(0);
    const uint64_t W = p.W;
    for (uint32_t r = 0; __any(r < cnt); r += 8) {
        uint32_t u[8];
        read_parents(gl, r, u);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            acc |= *(gw)((r + k < cnt && wok) ? cur + (uint64_t)u[k] * W + w : p.zrow + (w & 63));
        if (!(__ballot(nz(((have | acc) & need) ^ need)) & gmask)) cnt = 0;
    }
    return acc;
}

// One light vertex of a pull level on its TW lanes (all lanes of the wave
// call it; lanes with okv false take part in the shuffles only): every word
// with unseen bits ORs the frontier words of its parents (only those set in
// the frontier bitmap are read) until all still-needed bits are found (early
// exit); new bits go to seen, the next frontier slot and, when lds_row is
// given, to lds_row[w].  A vertex that joins the next frontier has all its
// words written to the slot, zeros included (slots are not cleared between
// uses).  Returns whether the vertex got any new bit (the same on all its
// lanes).
// sb: the vertex's source count in a batch union traversal (bits at or past
// it are never needed), 0 for all S sources.
template <int TW, typename WT = u64>
__device__ __forceinline__ bool pull_vertex(const WideParams &p, const WaveCtx &x, const u64 *cur,
                                            const uint32_t *fb_cur, u64 *nxt, uint32_t v,
                                            bool okv, uint32_t e0, uint32_t e1, u64 *lds_row,
                                            u64 &my_bits, uint32_t sb = 0) {
    constexpr uint32_t TV = 64 / TW;
    constexpr uint32_t K = sizeof(WT) / 8;  // source words per lane
    static_assert(K == 1 || TW >= 8, "two words per lane need tiles of 8 or more lanes");
    const uint32_t W = p.W, WG = (W + TW * K - 1) / (TW * K);
    bool any = false;
    WT n0 = wconst<WT>(0);  // this lane's new bits of word group 0
    for (uint32_t g = 0; g < WG; ++g) {
        const uint32_t w = (g * TW + x.wi) * K;
        WT n = wconst<WT>(0);
        if constexpr (TW >= 8) {
            // The vertex's TW lanes scan its in-list together.
            const bool wok = okv && w < W;
            const uint64_t idx = (uint64_t)v * W + w;
            const WT s = wok ? *(const WT *)(p.seen + idx) : wconst<WT>(~0ull);
            const WT need = wok ? ~s & words_live(p, x.act_cur, w, s, sb) : wconst<WT>(0);
            const u64 gmask = (~0ull / ((1ull << TV) - 1ull)) << x.vi;
            bool gon = (__ballot(nz(need)) & gmask) != 0 && e0 < e1;
            WT acc = wconst<WT>(0);
            uint32_t eb = e0;
            while (__any(gon)) {
                // a lane whose words already have every needed bit
                // reads the zero row instead of parent words
                acc |= scan_gather<TW, WT>(p, x, cur, fb_cur, gon ? eb : e1, e1, w,
                                           wok && nz(need & ~acc), need, acc);
                eb += TW;
                const bool sat = !nz((acc & need) ^ need);
                gon = gon && eb < e1 && (__ballot(!sat) & gmask) != 0;
            }
            n = acc & need;
            if (nz(n)) {
                act_or(x.act_lds, w, n);
                *(WT *)(p.seen + idx) = s | n;
                *(WT *)(nxt + idx) = n;
            } else if (WG > 1 && wok) {
                // several word groups: the row's other words are not known yet
                *(WT *)(nxt + idx) = n;
            }
        } else if (okv && w < W) {
            const uint64_t idx = (uint64_t)v * W + w;
            const u64 s = p.seen[idx];
            const u64 need = ~s & word_valid(sb ? sb : p.S, w) & x.act_cur[w];
            if (need) {
                // Parents in steps of 8 with no branches inside a
                // step: indices past the end repeat the last parent,
                // parents outside the frontier read the zero row, so
                // each step is three rounds of independent loads
                // (indices, bitmap words, frontier words).
                typedef const __attribute__((address_space(1))) uint32_t *g32;
                typedef const __attribute__((address_space(1))) u64 *g64;
                constexpr int U = 8;
                u64 acc = 0;
                for (uint32_t e = e0; e < e1; e += U) {
                    uint32_t uu[U];
#pragma unroll
                    for (int k = 0; k < U; ++k) uu[k] = ((g32)p.ri)[min(e + (uint32_t)k, e1 - 1)];
                    uint32_t fw[U];
#pragma unroll
                    for (int k = 0; k < U; ++k) fw[k] = ((g32)fb_cur)[uu[k] >> 5];
#pragma unroll
                    for (int k = 0; k < U; ++k) {
                        const bool inf = (fw[k] >> (uu[k] & 31)) & 1u;
                        acc |= *(g64)(inf ? cur + (uint64_t)uu[k] * W + w : p.zrow + w);
                    }
                    if ((acc & need) == need) break;
                }
                n = acc & need;
                if (n) {
                    atomicOr(&x.act_lds[w], n);
                    p.seen[idx] = s | n;
                    nxt[idx] = n;
                }
            }
            if (WG > 1 && !n) nxt[idx] = 0;
        }
        if constexpr (K == 1)
            if (lds_row && w < W) lds_row[w] = n;
        any |= nz(n);
        if (g == 0) n0 = n;
        my_bits += popc(n);
    }
    const u64 bal = __ballot(any);
    bool vany = false;
#pragma unroll
    for (uint32_t k = 0; k < TW; ++k) vany |= ((bal >> (k * TV + x.vi)) & 1ull) != 0;
    // one word group: the zero words of a vertex that joins the frontier
    if (WG == 1 && vany && okv && x.wi * K < W && !nz(n0))
        *(WT *)(nxt + (uint64_t)v * W + x.wi * K) = wconst<WT>(0);
    return vany;
}

// Heavy in-list chunk [e0, e1) for tiles of TW >= 8 words: steps of 64
// in-edges, lane group vi taking edges vi*TW .. vi*TW+TW-1 of each step, as
// scan_gather, but software-pipelined: the next step's bitmap word and the
// index after it are loaded before this step's gathers, so only the gathers
// stay on the chain.  Stops once every word's needed bits are found.
// Returns the OR of this lane's word over the chunk's frontier parents.
template <int TW, typename WT>
__device__ __forceinline__ WT heavy_gather(const WideParams &p, const WaveCtx &x, const u64 *cur,
                                           const uint32_t *fb, uint32_t e0, uint32_t e1,
                                           uint32_t w, bool wok, WT need) {
    constexpr uint32_t TV = 64 / TW;
    static_assert(TV <= 8, "heavy_gather needs tiles of 8 or more words");
    typedef const __attribute__((address_space(1))) uint32_t *g32;
    typedef const __attribute__((address_space(1))) WT *gw;
    const u64 gmask = (~0ull / ((1ull << TV) - 1ull)) << x.vi;
    const uint64_t W = p.W;
    uint32_t e = e0 + x.vi * TW + x.wi;
    uint32_t uu = e < e1 ? ((g32)p.ri)[e] : 0u;
    uint32_t fw = e < e1 ? ((g32)fb)[uu >> 5] : 0u;
    uint32_t un = e + 64 < e1 ? ((g32)p.ri)[e + 64] : 0u;
    uint32_t *const gl = x.par + x.vi * TW;
    WT acc = wconst<WT>(0), got = wconst<WT>(0);
    for (uint32_t base = e0; base < e1; base += 64) {
        const bool inf = e < e1 && ((fw >> (uu & 31)) & 1u);
        const uint32_t cnt = list_parents(gl, inf, uu, gmask);
        const uint32_t en = e + 64;
        const uint32_t fwn = en < e1 ? ((g32)fb)[un >> 5] : 0u;
        const uint32_t unn = en + 64 < e1 ? ((g32)p.ri)[en + 64] : 0u;
        // a lane whose word already has every needed bit reads the zero row
        const bool lok = wok && nz(need & ~got);
        for (uint32_t r = 0; __any(r < cnt); r += 8) {
            uint32_t u[8];
            read_parents(gl, r, u);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                acc |= *(gw)((r + k < cnt && lok) ? cur + (uint64_t)u[k] * W + w : p.zrow + (w & 63));
        }
        got = group_or<TV>(acc);
        if (__all(!nz((got & need) ^ need))) break;
        e = en;
        uu = un;
        fw = fwn;
        un = unn;
    }
    return acc;
}

// Pull.  Levels whose distances are deferred to the final pass: each wave
// takes its own tiles of TV consecutive vertices and appends new frontier
// vertices through its LDS buffer, so no wave waits for another's in-lists.
// Levels that store distances now: each workgroup takes blocks of NB
// consecutive vertices, the new bits go to LDS, and the block's distances
// are then stored one source row at a time with lane = vertex, so each
// store covers NB consecutive entries of a row.  Heavy in-lists come in
// chunks, one wave per chunk; a chunk's bits are ORed into nx and resolved
// by wide_pull_resolve after the grid barrier.
template <int TW>
__device__ __forceinline__ void wide_pull(const WideParams &p, const WaveCtx &x, const u64 *cur,
                                          const uint32_t *fb_cur, u64 *nxt, uint32_t *qnext,
                                          uint32_t *fb_next, WideCounters *c, uint32_t nd,
                                          u64 *lds_n, u64 &my_bits, u64 &my_deg) {
    constexpr uint32_t TV = 64 / TW;
    const uint32_t V = p.V, W = p.W, WG = (W + TW - 1) / TW;
    const uint32_t NB = p.pull_nb, LW = W + 1;  // LDS row stride (padded)
    const uint32_t wib = threadIdx.x / 64, nwb = blockDim.x / 64;
    const uint64_t nblk = (V + NB - 1) / NB;
    const uint32_t u16 = p.out_kind;
    if (!x.direct) {
        // Light vertices from their list (vertex and in-list bounds in one
        // load; vertices without in-edges are not in it), TV per tile;
        // heavy ones are chunked below.
        const uint32_t nl = p.n_light_in;
        const uint64_t ntv = (nl + TV - 1) / TV;
        WaveAppender qa{x.wq};
        for (uint64_t t = x.gwave; t < ntv; t += x.nwaves) {
            const uint32_t i = (uint32_t)t * TV + x.vi;
            const bool okv = i < nl;
            uint4 it = make_uint4(0u, 0u, 0u, 0u);
            if (okv) it = p.light_in[i];
            const uint32_t v = it.x, e0 = it.y, e1 = it.z, sb = it.w;
            bool vany;
            if constexpr (TW >= 8) {
                if (p.pair)  // two adjacent words per lane: TV vertices over 2*TW words
                    vany = pull_vertex<TW, u64x2>(p, x, cur, fb_cur, nxt, v, okv, e0, e1,
                                                  nullptr, my_bits, sb);
                else
                    vany = pull_vertex<TW>(p, x, cur, fb_cur, nxt, v, okv, e0, e1, nullptr,
                                           my_bits, sb);
            } else {
                vany = pull_vertex<TW>(p, x, cur, fb_cur, nxt, v, okv, e0, e1, nullptr, my_bits,
                                       sb);
            }
            const bool app = x.wi == 0 && okv && vany;
            qa.push(app, v, &c->next_count, qnext);
            if (app) {
                if (nd < p.defer) p.lv[v] |= 1u << nd;
                mark_frontier(fb_next, v);
                my_deg += p.ro[v + 1] - p.ro[v];
            }
        }
        qa.flush(&c->next_count, qnext);
    } else
    for (uint64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        const uint32_t v0 = (uint32_t)blk * NB;
        for (uint32_t st = wib; st < NB / TV; st += nwb) {
            const uint32_t vl = st * TV + x.vi;
            const uint32_t v = v0 + vl;
            bool okv = v < V;
            uint32_t e0 = 0, e1 = 0;
            if (okv) {
                e0 = p.co[v];
                e1 = p.co[v + 1];
                if (e1 - e0 > p.heavy_t) okv = false;  // chunked below
            }
            const bool vany = pull_vertex<TW>(p, x, cur, fb_cur, nxt, v, okv, e0, e1,
                                              lds_n + vl * LW, my_bits);
            // New frontier vertices are gathered in LDS and published once
            // per block (one global atomic instead of one per wave).
            const bool app = x.wi == 0 && okv && vany;
            const u64 ab = __ballot(app);
            if (ab) {
                uint32_t lb = 0;
                const int leader = __ffsll((long long)ab) - 1;
                if ((int)x.lane == leader) lb = atomicAdd(&x.blk[0], (uint32_t)__popcll(ab));
                lb = (uint32_t)__builtin_amdgcn_readlane((int)lb, leader);
                if (app) {
                    if (nd < p.defer) p.lv[v] |= 1u << nd;
                    x.blk[4 + lb + (uint32_t)__popcll(ab & ((1ull << x.lane) - 1ull))] = v;
                    atomicOr(&x.blk[1 + ((v >> 5) - (v0 >> 5))], 1u << (v & 31));
                    my_deg += p.ro[v + 1] - p.ro[v];
                }
            }
        }
        __syncthreads();
        if (wib == 0) {
            const uint32_t cnt = x.blk[0];
            uint32_t gb = 0;
            if (cnt) {
                if (x.lane == 0) gb = atomicAdd(&c->next_count, cnt);
                gb = (uint32_t)__builtin_amdgcn_readfirstlane((int)gb);
            }
            if (x.lane < cnt) qnext[gb + x.lane] = x.blk[4 + x.lane];
            if (x.lane < 3) {
                const uint32_t m = x.blk[1 + x.lane];
                if (m) atomicOr(&fb_next[(v0 >> 5) + x.lane], m);
            }
        }
        if (x.direct) {
            for (uint32_t w = wib; w < W; w += nwb) {
                const bool okl = x.lane < NB && v0 + x.lane < V;
                const u64 n = okl ? lds_n[x.lane * LW + w] : 0ull;
                u64 un = wave_or64(n);
                while (un) {
                    const uint32_t b = (uint32_t)__builtin_ctzll(un);
                    un &= un - 1;
                    if ((n >> b) & 1ull)
                        store_dist<true>(p.out, u16, dist_index(p, (uint64_t)w * 64 + b, v0 + x.lane),
                                         nd, p.flags);
                }
            }
        }
        if (wib == 0 && x.lane < 4) x.blk[x.lane] = 0;  // wave 0 read them above
        __syncthreads();
    }
    // Heavy in-lists: one wave per (chunk, word group); lanes vi take every
    // TV-th in-edge, U parents per lane per step (R-MAT20 S=64, level 1:
    // U = 1 394 us, U = 8 289 us); a chunk stops once its words' needed
    // bits (less those other chunks already found) are in.
    if constexpr (TW >= 8) {
        if (p.pair) {
            // two adjacent words per lane, as the light vertices above
            const uint32_t WG2 = (W + 2 * TW - 1) / (2 * TW);
            const uint64_t nh2 = (uint64_t)p.n_hchunk_in * WG2;
            for (uint64_t t = x.gwave; t < nh2; t += x.nwaves) {
                const uint4 ch = p.hchunk_in[t / WG2];
                const uint32_t h = ch.x, e0 = ch.y, e1 = ch.z;
                const uint32_t w = ((uint32_t)(t % WG2) * TW + x.wi) * 2;
                const uint64_t idx = (uint64_t)h * W + w;
                const bool wok = w < W;
                u64x2 need = wconst<u64x2>(0);
                if (wok)
                    need = ~*(const u64x2 *)(p.seen + idx) & ~*(const u64x2 *)(p.nx + idx) &
                           words_live(p, x.act_cur, w, need, ch.w);
                if (!__any(nz(need))) continue;  // wave-uniform
                const u64x2 acc = heavy_gather<TW, u64x2>(p, x, cur, fb_cur, e0, e1, w, wok, need);
                const u64x2 g = group_or<TV>(acc) & need;
                if (x.vi == 0) {
                    if (g.x) atomicOr(&p.nx[idx], g.x);
                    if (g.y) atomicOr(&p.nx[idx + 1], g.y);
                }
            }
            return;
        }
    }
    const uint64_t nh = (uint64_t)p.n_hchunk_in * WG;
    constexpr uint32_t U = 8;  // (narrow tiles) independent in-edges per lane per step
    for (uint64_t t = x.gwave; t < nh; t += x.nwaves) {
        const uint4 ch = p.hchunk_in[t / WG];
        const uint32_t h = ch.x, e0 = ch.y, e1 = ch.z;
        const uint32_t w = (uint32_t)(t % WG) * TW + x.wi;
        const uint64_t idx = (uint64_t)h * W + w;
        u64 need = 0;
        if (w < W) need = ~p.seen[idx] & word_valid(ch.w ? ch.w : p.S, w) & ~p.nx[idx] & x.act_cur[w];
        if (!__any(need != 0)) continue;  // wave-uniform
        // Branch-free steps as in the light pull: indices past the chunk
        // repeat its last parent, parents outside the frontier (and lanes
        // past the last word) read the zero row.
        typedef const __attribute__((address_space(1))) uint32_t *g32;
        typedef const __attribute__((address_space(1))) u64 *g64;
        const bool wok = w < W;
        u64 acc = 0;
        if constexpr (TW >= 8) {
            acc = heavy_gather<TW, u64>(p, x, cur, fb_cur, e0, e1, w, wok, need);
        } else
        for (uint32_t base = e0; base < e1; base += TV * U) {
            uint32_t uu[U];
#pragma unroll
            for (uint32_t k = 0; k < U; ++k) uu[k] = ((g32)p.ri)[min(base + k * TV + x.vi, e1 - 1)];
            uint32_t fw[U];
#pragma unroll
            for (uint32_t k = 0; k < U; ++k) fw[k] = ((g32)fb_cur)[uu[k] >> 5];
#pragma unroll
            for (uint32_t k = 0; k < U; ++k) {
                const bool inf = wok && ((fw[k] >> (uu[k] & 31)) & 1u);
                acc |= *(g64)(inf ? cur + (uint64_t)uu[k] * W + w : p.zrow + (w & 63));
            }
            const u64 g = group_or64<TV>(acc);
            if (__all((g & need) == need)) break;
        }
        const u64 g = group_or64<TV>(acc) & need;
        if (x.vi == 0 && g) atomicOr(&p.nx[idx], g);
    }
}

// After a pull level's barrier: the old frontier's words and bitmap are
// cleared, and each heavy vertex's chunk results (ORed into nx) become its
// new bits, distances and next-frontier membership.
template <int TW>
__device__ __forceinline__ void wide_pull_resolve(const WideParams &p, const WaveCtx &x, u64 *cur,
                                                  const uint32_t *qcur, uint32_t n_front,
                                                  uint32_t *fb_cur, u64 *nxt, uint32_t *qnext,
                                                  uint32_t *fb_next, WideCounters *c,
                                                  uint32_t nd, uint32_t tick, uint64_t gthread,
                                                  uint64_t nthreads, u64 &my_bits,
                                                  u64 &my_deg) {
    const uint32_t W = p.W;
    // the frontier words stay (the final pass reads the kept levels'; a
    // slot's rows are all rewritten when it is next used): only the bitmap goes
    clear_frontier(fb_cur, qcur, n_front, p.V, gthread, nthreads);
    const uint32_t nh = p.n_heavy_in;
    if constexpr (TW >= 8) {
        if (!x.direct && p.pair) {
            // as below with two adjacent words per lane (one word group)
            constexpr uint32_t TV = 64 / TW;
            const uint32_t WG2 = (W + 2 * TW - 1) / (2 * TW);
            const uint64_t htiles = (nh + TV - 1) / TV;
            WaveAppender qa{x.wq};
            for (uint64_t t = x.gwave; t < htiles; t += x.nwaves) {
                const uint32_t i = (uint32_t)t * TV + x.vi;
                const bool okv = i < nh;
                const uint32_t h = okv ? p.heavy_in[i] : 0;
                bool any = false;
                for (uint32_t g = 0; g < WG2; ++g) {
                    const uint32_t w = (g * TW + x.wi) * 2;
                    u64x2 n = wconst<u64x2>(0);
                    if (okv && w < W) {
                        const uint64_t idx = (uint64_t)h * W + w;
                        const u64x2 xw = *(const u64x2 *)(p.nx + idx);
                        if (nz(xw)) {
                            *(u64x2 *)(p.nx + idx) = wconst<u64x2>(0);
                            const u64x2 s = *(const u64x2 *)(p.seen + idx);
                            n = xw & ~s & words_valid(p, w, n);
                            if (nz(n)) {
                                words_or(x.act_lds + w, n);
                                *(u64x2 *)(p.seen + idx) = s | n;
                            }
                        }
                        *(u64x2 *)(nxt + idx) = n;  // every word: the slot is not cleared
                    }
                    any |= nz(n);
                    my_bits += popc(n);
                }
                const u64 bal = __ballot(any);
                bool vany = false;
#pragma unroll
                for (uint32_t k = 0; k < TW; ++k) vany |= ((bal >> (k * TV + x.vi)) & 1ull) != 0;
                const bool app = x.wi == 0 && okv && vany;
                qa.push(app, h, &c->next_count, qnext);
                if (app) {
                    mark_frontier(fb_next, h);
                    if (nd < p.defer) p.lv[h] |= 1u << nd;
                    my_deg += p.ro[h + 1] - p.ro[h];
                }
            }
            qa.flush(&c->next_count, qnext);
            return;
        }
    }
    if (!x.direct) {
        // Distances deferred: tiles of TV heavy vertices x TW words as in the
        // push filter, so a vertex's words are read and written as one
        // contiguous run and its lanes own it (no level stamp needed).
        constexpr uint32_t TV = 64 / TW;
        const uint32_t WG = (W + TW - 1) / TW;
        const uint64_t htiles = (nh + TV - 1) / TV;
        WaveAppender qa{x.wq};
        for (uint64_t t = x.gwave; t < htiles; t += x.nwaves) {
            const uint32_t i = (uint32_t)t * TV + x.vi;
            const bool okv = i < nh;
            const uint32_t h = okv ? p.heavy_in[i] : 0;
            bool any = false;
            for (uint32_t g = 0; g < WG; ++g) {
                const uint32_t w = g * TW + x.wi;
                u64 n = 0;
                if (okv && w < W) {
                    const uint64_t idx = (uint64_t)h * W + w;
                    const u64 xw = p.nx[idx];
                    if (xw) {
                        p.nx[idx] = 0;
                        const u64 s = p.seen[idx];
                        n = xw & ~s & word_valid(p.S, w);
                        if (n) {
                            atomicOr(&x.act_lds[w], n);
                            p.seen[idx] = s | n;
                        }
                    }
                    nxt[idx] = n;  // every word: the slot is not cleared
                }
                any |= n != 0;
                my_bits += __popcll(n);
            }
            const u64 bal = __ballot(any);
            bool vany = false;
#pragma unroll
            for (uint32_t k = 0; k < TW; ++k) vany |= ((bal >> (k * TV + x.vi)) & 1ull) != 0;
            const bool app = x.wi == 0 && okv && vany;
            qa.push(app, h, &c->next_count, qnext);
            if (app) {
                mark_frontier(fb_next, h);
                if (nd < p.defer) p.lv[h] |= 1u << nd;
                my_deg += p.ro[h + 1] - p.ro[h];
            }
        }
        qa.flush(&c->next_count, qnext);
        return;
    }
    // Distances stored now: tiles of (word, 64 heavy vertices), lane =
    // vertex, so a source bit's distances for the tile's vertices go out in
    // one store instruction (heavy vertices are listed in increasing order).
    const uint64_t hg = (nh + 63) / 64;
    const uint64_t tiles = hg * W;
    const uint32_t u16 = p.out_kind;
    WaveAppender qa{x.wq};
    for (uint64_t t = x.gwave; t < tiles; t += x.nwaves) {
        const uint32_t w = (uint32_t)(t / hg);
        const uint32_t k = (uint32_t)(t % hg) * 64 + x.lane;
        bool app = false;
        uint32_t h = 0;
        u64 n = 0;
        if (k < nh) {
            h = p.heavy_in[k];
            const uint64_t idx = (uint64_t)h * W + w;
            const u64 xw = p.nx[idx];
            if (xw) {
                p.nx[idx] = 0;
                const u64 s = p.seen[idx];
                n = xw & ~s & word_valid(p.S, w);
                if (n) {
                    atomicOr(&x.act_lds[w], n);
                    p.seen[idx] = s | n;
                    my_bits += __popcll(n);
                    app = atomicExch(&p.stamp[h], tick) != tick;
                }
            }
            nxt[idx] = n;  // every word: the slot is not cleared
        }
        if (x.direct) {
            u64 un = wave_or64(n);
            while (un) {
                const uint32_t b = (uint32_t)__builtin_ctzll(un);
                un &= un - 1;
                if ((n >> b) & 1ull)
                    store_dist(p.out, u16, dist_index(p, (uint64_t)w * 64 + b, h), nd, p.flags);
            }
        }
        qa.push(app, h, &c->next_count, qnext);
        if (app) {
            mark_frontier(fb_next, h);
            if (nd < p.defer) p.lv[h] |= 1u << nd;
            my_deg += p.ro[h + 1] - p.ro[h];
        }
    }
    qa.flush(&c->next_count, qnext);
}

// Bits 0..7 of x to bit positions 0, 4, ..., 28.
__device__ __forceinline__ uint32_t spread8x4(uint32_t x) {
    x &= 0xFFu;
    x = (x | (x << 12)) & 0x000F000Fu;
    x = (x | (x << 6)) & 0x03030303u;
    x = (x | (x << 3)) & 0x11111111u;
    return x;
}

// Final-pass row stores of one tile, T = output element: wave i (of 8)
// writes the rows b = 8k + i of each of the tile's kFinalWords words, so the
// distance it needs is always nibble i of nibble word k; lane l writes the
// tile's vertices 2l and 2l+1 (one pair store when both entries are stored
// and the pair is aligned, single stores otherwise).  nib[w][k][v]: nibble
// i of byte k = distance of source bit 8k+i (15 = unreached); put[w][v]: bits
// to store (a pair reached at a later, directly stored level is left alone).
template <typename T, typename T2, bool BM>
__device__ __forceinline__ void final_tile_rows(T *out, uint64_t V, uint32_t v0, uint32_t w0,
                                                uint32_t W, uint64_t S, const uint32_t *nib,
                                                const u64 *put, uint32_t wave, uint32_t lane,
                                                const BatchGraphMap *bmap, const uint32_t *vgraph) {
    const uint32_t va = 2 * lane;
    const bool oka = v0 + va < V, okb = v0 + va + 1 < V;
    // BM (batch union traversal): each vertex's rows are its own graph's,
    // and only its graph's source bits exist
    uint64_t ra = 0, rb = 0, sta = 0, stb = 0, lima = S, limb = S;
    bool same = true;
    if constexpr (BM) {
        const uint32_t ga = oka ? vgraph[v0 + va] : 0u, gb = okb ? vgraph[v0 + va + 1] : ga;
        const BatchGraphMap A = bmap[ga], B = bmap[gb];
        ra = A.out_off + (v0 + va - A.base);
        sta = A.V;
        lima = A.S;
        rb = B.out_off + (v0 + va + 1 - B.base);
        stb = B.V;
        limb = B.S;
        same = ga == gb;
    }
    for (uint32_t wl = 0; wl < kFinalWords; ++wl) {
        const uint32_t w = w0 + wl;
        if (w >= W) break;  // uniform
        // both vertices' put words, read once per word (adjacent: one ds_read2)
        const u64 *pw = put + wl * kFinalVerts + va;
        const u64 pwa = oka ? pw[0] : 0ull, pwb = okb ? pw[1] : 0ull;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
            const uint32_t b = 8 * k + wave;
            const uint64_t sidx = (uint64_t)w * 64 + b;
            if (sidx >= S) break;  // uniform, b grows with k
            const uint32_t *nw = nib + (wl * 8 + k) * kFinalVerts + va;
            const uint32_t da = __builtin_amdgcn_ubfe(nw[0], 4 * wave, 4);
            const uint32_t db = __builtin_amdgcn_ubfe(nw[1], 4 * wave, 4);
            const bool pa = ((pwa >> b) & 1ull) && sidx < lima, pb = ((pwb >> b) & 1ull) && sidx < limb;
            const T ta = da == 15u ? (T)~(T)0 : (T)da, tb = db == 15u ? (T)~(T)0 : (T)db;
            T *dst, *dstb;
            if constexpr (BM) {
                dst = out + ra + sidx * sta;
                dstb = out + rb + sidx * stb;
            } else {
                dst = out + sidx * V + v0 + va;
                dstb = dst + 1;
            }
            if (pa && pb && same && ((uintptr_t)dst & (sizeof(T2) - 1)) == 0) {
                T2 two = (T2)ta | ((T2)tb << (8 * sizeof(T)));
                __builtin_nontemporal_store(two, (T2 *)dst);
            } else {
                if (pa) __builtin_nontemporal_store(ta, dst);
                if (pb) __builtin_nontemporal_store(tb, dstb);
            }
        }
    }
}

// Grid barrier of the cooperative wide kernel (every workgroup is resident):
// each workgroup arrives at its group's counter (workgroup index mod 8), the
// last of a group at the top counter, and the last there advances the
// generation the others wait on.  An agent-scope release before arriving and
// an acquire after leaving make every workgroup's writes before the barrier
// visible to all after it.  The wait is bounded: on a timeout (or another
// workgroup's) the abort flag is set and the barrier returns false, so the
// kernel always ends.  About 10 us at 512 workgroups on gfx950 against
// 48 us for cooperative_groups' grid sync.
__device__ __forceinline__ bool grid_barrier(uint32_t *bw) {
    __shared__ uint32_t s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t *const top = bw, *const gen = bw + 32 * 9, *const abort_flag = bw + 32 * 10;
        const uint32_t g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t nb = gridDim.x, grp = blockIdx.x & 7u;
        uint32_t *const gc = bw + 32 * (1 + grp);
        bool last = false;
        if (__hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 ==
            (nb - grp + 7u) / 8u) {
            __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            last = __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 ==
                   (nb < 8u ? nb : 8u);
        }
        bool ok = true;
        if (last) {
            __hip_atomic_store(top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0;
            while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                    ++spins > (1u << 28)) {
                    __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

template <int TW>
__global__ void __launch_bounds__(kWideThreads) __attribute__((amdgpu_waves_per_eu(4))) wide_bfs_kernel(WideParams p) {
    extern __shared__ u64 lds_n[];
    constexpr uint32_t TV = 64 / TW;
    WaveCtx x;
    x.lane = lane_id();
    x.vi = x.lane % TV;
    x.wi = x.lane / TV;
    x.gwave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / 64;
    x.nwaves = (uint64_t)gridDim.x * blockDim.x / 64;
    const uint64_t gthread = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
    const uint32_t V = p.V, W = p.W;
    const bool first = blockIdx.x == 0 && threadIdx.x == 0;
    x.act_lds = lds_n + (size_t)p.pull_nb * (W + 1);
    for (uint32_t w = threadIdx.x; w < W; w += blockDim.x) x.act_lds[w] = 0;
    __shared__ uint32_t s_blk[4 + 64];
    __shared__ uint32_t s_wq[kWideThreads];
    __shared__ __align__(16) uint32_t s_par[kWideThreads];
    __shared__ u64 s_sum[2];
    if (threadIdx.x < 2) s_sum[threadIdx.x] = 0;
    x.blk = s_blk;
    x.wq = s_wq + (threadIdx.x / 64) * 64;
    x.par = s_par + (threadIdx.x / 64) * 64;
    if (threadIdx.x < 4) s_blk[threadIdx.x] = 0;
    __syncthreads();

    const uint64_t stride = (uint64_t)V * W;
    const uint32_t D = p.defer;
    auto slot = [&](uint32_t d) { return p.fr + (d < D ? d : D + ((d - D) & 1u)) * stride; };
    u64 *cur = slot(0), *nxt = slot(1);
    uint32_t *fb_cur = p.fbA, *fb_next = p.fbB;
    uint32_t *qcur = p.qf0, *qnext = p.qf1;
    uint32_t n_front = ld_dev_u32(&p.ctr[3].next_count);
    u64 prev_bits = p.S, prev_deg = 0;
    uint32_t dir = (p.policy == 2) ? DirPull : DirPush;
    uint32_t push_levels = 0, pull_levels = 0;
    uint32_t L = 0;
    bool aborted = false;

    for (;; ++L) {
        if (n_front == 0 || L > V) break;
        WideCounters *c = &p.ctr[L & 3];
        if (p.policy == 0 && L > 0) {
            // Direction controller (crates/gpu_bfs/src/dispatch/push_pull.rs:483-498):
            // Push->Pull when m_f > |E|/alpha, Pull->Push when n_f < |V|/beta.
            u64 n_f, m_f;
            if (p.rule == 1) {
                // the reference's estimate from the new-bit count
                n_f = (prev_bits + p.S - 1) / p.S;
                m_f = V ? (n_f * p.E) / V : 0;
            } else {
                // measured: union frontier size and its out-degree sum
                n_f = n_front;
                m_f = prev_deg;
            }
            if (dir == DirPush) {
                if (m_f > p.E / (p.alpha ? p.alpha : 1)) dir = DirPull;
            } else {
                if (n_f < (u64)V / (p.beta ? p.beta : 1)) dir = DirPush;
            }
        }
        const uint32_t nd = L + 1;
        const uint32_t tick = p.tick_base + L + 1;
        x.act_cur = p.act + (size_t)(L % 3) * W;
        x.direct = p.out != nullptr && nd >= D;
        u64 my_bits = 0, my_deg = 0;
        if (p.trace && first && L < p.trace_cap) {
            p.trace[4 * L] = wall_clock64();
            p.trace[4 * L + 1] = ((u64)dir << 32) | n_front;
        }

        if (dir == DirPush) {
            ++push_levels;
            wide_push_expand<TW>(p, x, cur, qcur, n_front, fb_cur, c, tick);
            if (!grid_barrier(p.gbar)) {
                aborted = true;
                break;
            }
            if (p.trace && first && L < p.trace_cap) p.trace[4 * L + 2] = wall_clock64();
            wide_push_filter<TW>(p, x, cur, qcur, n_front, fb_cur, nxt, qnext, fb_next, c, nd,
                                 gthread, nthreads, my_bits, my_deg);
        } else {
            ++pull_levels;
            wide_pull<TW>(p, x, cur, fb_cur, nxt, qnext, fb_next, c, nd, lds_n, my_bits, my_deg);
            if (!grid_barrier(p.gbar)) {
                aborted = true;
                break;
            }
            if (p.trace && first && L < p.trace_cap) p.trace[4 * L + 2] = wall_clock64();
            wide_pull_resolve<TW>(p, x, cur, qcur, n_front, fb_cur, nxt, qnext, fb_next, c, nd,
                                  tick, gthread, nthreads, my_bits, my_deg);
        }
        // Level counters for the direction controller: summed per workgroup
        // in LDS, one global atomic per workgroup (after the barrier below).
        {
            const u64 sb = wave_sum64(my_bits), sd = wave_sum64(my_deg);
            if (x.lane == 0) {
                if (sb) atomicAdd(&s_sum[0], sb);
                if (sd) atomicAdd(&s_sum[1], sd);
            }
        }
        if (first) {
            WideCounters *z = &p.ctr[(L + 2) & 3];
            z->next_count = 0;
            z->cand_count = 0;
            z->active_bits = 0;
            z->deg_sum = 0;
        }
        // Sources that reached new vertices at this level are the ones with a
        // frontier at the next: one atomic per word per workgroup.  The ring
        // slot of level L+2 was last read at level L-1 and is cleared here.
        __syncthreads();
        if (threadIdx.x == 0) {
            if (s_sum[0]) atomicAdd(&c->active_bits, s_sum[0]);
            if (s_sum[1]) atomicAdd(&c->deg_sum, s_sum[1]);
            s_sum[0] = 0;
            s_sum[1] = 0;
        }
        for (uint32_t w = threadIdx.x; w < W; w += blockDim.x) {
            const u64 a = x.act_lds[w];
            if (a) {
                atomicOr(&p.act[(size_t)((L + 1) % 3) * W + w], a);
                x.act_lds[w] = 0;
            }
            if (blockIdx.x == 0) p.act[(size_t)((L + 2) % 3) * W + w] = 0;
        }
        if (!grid_barrier(p.gbar)) {
            aborted = true;
            break;
        }
        n_front = ld_dev_u32(&c->next_count);
        prev_bits = ld_dev_u64(&c->active_bits);
        prev_deg = ld_dev_u64(&c->deg_sum);
        if (p.trace && first && L < p.trace_cap) p.trace[4 * L + 3] = prev_bits;
        cur = nxt;
        nxt = slot(L + 2);
        uint32_t *tb = fb_cur;
        fb_cur = fb_next;
        fb_next = tb;
        uint32_t *tq = qcur;
        qcur = qnext;
        qnext = tq;
    }

    if (aborted) {
        if (first) p.flags[FlagAbort] = 1;
        return;
    }
    if (first) {
        p.flags[FlagLevels] = L;
        p.flags[FlagPushLevels] = push_levels;
        p.flags[FlagPullLevels] = pull_levels;
        if (p.trace && L < p.trace_cap) p.trace[4 * L] = wall_clock64();
    }
}

// ---------------------------------------------------------------- stepped wide
//
// The same levels as wide_bfs_kernel, one launch per phase: the kernel
// boundaries take the place of the grid barriers, so no launch needs its
// whole grid resident and other work (another lane's traversal) may share
// the device.  wide_ctl_kernel (one thread) runs before every level: it
// reads the previous level's counters, applies the direction rule, clears
// the ring slots the level after next will use, and records the level in
// WideCtl; once the frontier is empty it marks the traversal done and the
// remaining enqueued launches return at once.

// `probe` (the launch after the last enqueued level) only marks the
// traversal done when no frontier is left; it decides no level, so launches
// enqueued later continue from the next level.
__global__ void wide_ctl_kernel(WideParams p, uint32_t probe) {
    WideCtl *ct = p.ctl;
    if (ld_dev_u32(&ct->done)) return;
    const uint32_t L = ld_dev_u32(&ct->next), V = p.V, W = p.W;
    uint32_t n_front, dir = ld_dev_u32(&ct->dir);
    u64 prev_bits, prev_deg;
    if (L == 0) {
        n_front = ld_dev_u32(&p.ctr[3].next_count);
        prev_bits = p.S;
        prev_deg = 0;
        dir = (p.policy == 2) ? DirPull : DirPush;
    } else {
        WideCounters *c = &p.ctr[(L - 1) & 3];
        n_front = ld_dev_u32(&c->next_count);
        prev_bits = ld_dev_u64(&c->active_bits);
        prev_deg = ld_dev_u64(&c->deg_sum);
    }
    // per-level trace, as wide_bfs_kernel records it: the new bits of the
    // level just finished, and the start of the next (or the end)
    if (p.trace && L > 0 && L - 1 < p.trace_cap) p.trace[4 * (L - 1) + 3] = prev_bits;
    if (n_front == 0 || L > V) {
        ct->done = 1;
        p.flags[FlagLevels] = L;
        p.flags[FlagPushLevels] = ct->push_levels;
        p.flags[FlagPullLevels] = ct->pull_levels;
        if (p.trace && L < p.trace_cap) p.trace[4 * L] = wall_clock64();
        __threadfence();
        return;
    }
    if (probe) return;
    if (p.policy == 0 && L > 0) {
        // the rule of wide_bfs_kernel (crates/gpu_bfs/src/dispatch/push_pull.rs:483-498)
        u64 n_f, m_f;
        if (p.rule == 1) {
            n_f = (prev_bits + p.S - 1) / p.S;
            m_f = V ? (n_f * p.E) / V : 0;
        } else {
            n_f = n_front;
            m_f = prev_deg;
        }
        if (dir == DirPush) {
            if (m_f > p.E / (p.alpha ? p.alpha : 1)) dir = DirPull;
        } else {
            if (n_f < (u64)V / (p.beta ? p.beta : 1)) dir = DirPush;
        }
    }
    if (dir == DirPush)
        ++ct->push_levels;
    else
        ++ct->pull_levels;
    // slots of level L + 2, last used at level L - 1 (act) and L - 2 (ctr)
    WideCounters *z = &p.ctr[(L + 2) & 3];
    z->next_count = 0;
    z->cand_count = 0;
    z->active_bits = 0;
    z->deg_sum = 0;
    for (uint32_t w = 0; w < W; ++w) p.act[(size_t)((L + 2) % 3) * W + w] = 0;
    ct->L = L;
    ct->dir = dir;
    ct->n_front = n_front;
    ct->next = L + 1;
    if (p.trace && L < p.trace_cap) {
        p.trace[4 * L] = wall_clock64();
        p.trace[4 * L + 1] = ((u64)dir << 32) | n_front;
    }
    __threadfence();
}

// Phase A (push expand / pull) or phase B (push filter / pull resolve) of
// level WideCtl::L; per-workgroup sums and active-source words are flushed
// at the end of each phase.
template <int TW, int PHASE>
__global__ void __launch_bounds__(kWideThreads) __attribute__((amdgpu_waves_per_eu(4))) wide_step_kernel(WideParams p) {
    extern __shared__ u64 lds_n[];
    const WideCtl *ct = p.ctl;
    if (ld_dev_u32(&ct->done)) return;
    const uint32_t L = ld_dev_u32(&ct->L), dir = ld_dev_u32(&ct->dir);
    const uint32_t n_front = ld_dev_u32(&ct->n_front);
    constexpr uint32_t TV = 64 / TW;
    WaveCtx x;
    x.lane = lane_id();
    x.vi = x.lane % TV;
    x.wi = x.lane / TV;
    x.gwave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / 64;
    x.nwaves = (uint64_t)gridDim.x * blockDim.x / 64;
    const uint64_t gthread = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
    const uint32_t V = p.V, W = p.W;
    x.act_lds = lds_n + (size_t)p.pull_nb * (W + 1);
    for (uint32_t w = threadIdx.x; w < W; w += blockDim.x) x.act_lds[w] = 0;
    __shared__ uint32_t s_blk[4 + 64];
    __shared__ uint32_t s_wq[kWideThreads];
    __shared__ __align__(16) uint32_t s_par[kWideThreads];
    __shared__ u64 s_sum[2];
    if (threadIdx.x < 2) s_sum[threadIdx.x] = 0;
    x.blk = s_blk;
    x.wq = s_wq + (threadIdx.x / 64) * 64;
    x.par = s_par + (threadIdx.x / 64) * 64;
    if (threadIdx.x < 4) s_blk[threadIdx.x] = 0;
    __syncthreads();

    const uint64_t stride = (uint64_t)V * W;
    const uint32_t D = p.defer;
    auto slot = [&](uint32_t d) { return p.fr + (d < D ? d : D + ((d - D) & 1u)) * stride; };
    u64 *const cur = slot(L), *const nxt = slot(L + 1);
    uint32_t *const fb_cur = (L & 1) ? p.fbB : p.fbA, *const fb_next = (L & 1) ? p.fbA : p.fbB;
    uint32_t *const qcur = (L & 1) ? p.qf1 : p.qf0, *const qnext = (L & 1) ? p.qf0 : p.qf1;
    WideCounters *const c = &p.ctr[L & 3];
    const uint32_t nd = L + 1;
    const uint32_t tick = p.tick_base + L + 1;
    x.act_cur = p.act + (size_t)(L % 3) * W;
    x.direct = p.out != nullptr && nd >= D;
    u64 my_bits = 0, my_deg = 0;
    if (PHASE == 1 && p.trace && L < p.trace_cap && blockIdx.x == 0 && threadIdx.x == 0)
        p.trace[4 * L + 2] = wall_clock64();
    if (PHASE == 0) {
        if (dir == DirPush)
            wide_push_expand<TW>(p, x, cur, qcur, n_front, fb_cur, c, tick);
        else
            wide_pull<TW>(p, x, cur, fb_cur, nxt, qnext, fb_next, c, nd, lds_n, my_bits, my_deg);
    } else {
        if (dir == DirPush)
            wide_push_filter<TW>(p, x, cur, qcur, n_front, fb_cur, nxt, qnext, fb_next, c, nd,
                                 gthread, nthreads, my_bits, my_deg);
        else
            wide_pull_resolve<TW>(p, x, cur, qcur, n_front, fb_cur, nxt, qnext, fb_next, c, nd,
                                  tick, gthread, nthreads, my_bits, my_deg);
    }
    {
        const u64 sb = wave_sum64(my_bits), sd = wave_sum64(my_deg);
        if (x.lane == 0) {
            if (sb) atomicAdd(&s_sum[0], sb);
            if (sd) atomicAdd(&s_sum[1], sd);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_sum[0]) atomicAdd(&c->active_bits, s_sum[0]);
        if (s_sum[1]) atomicAdd(&c->deg_sum, s_sum[1]);
    }
    for (uint32_t w = threadIdx.x; w < W; w += blockDim.x) {
        const u64 a = x.act_lds[w];
        if (a) atomicOr(&p.act[(size_t)((L + 1) % 3) * W + w], a);
    }
}

// Final pass of a wide traversal, launched after wide_bfs_kernel on the same
// stream (its own launch, so its registers and LDS are sized for it alone).
template <typename T, typename T2>
__global__ void __launch_bounds__(kWideThreads) wide_final_kernel(WideParams p) {
    __shared__ uint32_t L_nib[kFinalWords * 8 * kFinalVerts];  // [kFinalWords][8][kFinalVerts]
    __shared__ u64 L_put[kFinalWords * kFinalVerts];            // [kFinalWords][kFinalVerts]
    if (p.flags[FlagAbort]) return;  // the traversal did not finish
    if (p.ctl && !ld_dev_u32(&p.ctl->done)) {
        // a stepped traversal that needs more levels than were enqueued:
        // the host enqueues more and this pass again
        if (blockIdx.x == 0 && threadIdx.x == 0) p.flags[FlagIncomplete] = 1;
        return;
    }
    const uint32_t V = p.V, W = p.W;
    const uint64_t stride = (uint64_t)V * W;
    const uint32_t kept = min(p.flags[FlagLevels], p.defer);
    // With every level kept (no distance stored at once) and no
    // reachability words wanted, the kept slots hold all of a vertex's bits:
    // the visited words are cleared without being read.
    const bool use_seen = p.reach != nullptr || p.flags[FlagLevels] > kept;
    // Lane = vertex in the row stores, so a source row's stores coalesce: the
    // distances of the kept levels are read from their slots as bit planes
    // (plane j = OR of the slots whose distance has bit j) and every source
    // row of the tile is written once, with the sentinel for unreached pairs;
    // pairs reached at a later, directly stored level are left alone.  The
    // visited words are cleared for the next run (the kept slots are not:
    // their rows are rewritten whole when next used, and only rows of
    // vertices in a level's frontier are ever read), and the reachability
    // words are written when asked.
    // Work items (tile, word group) in runs of whole tiles per workgroup, so
    // the word groups of one tile (sharing the vertices' cache lines) run on
    // one CU, in order: the last one clears the tile's level masks.
    const uint64_t ntile = (V + kFinalVerts - 1) / kFinalVerts;
    const uint64_t ngrp = (W + kFinalWords - 1) / kFinalWords;
    const uint64_t items = ntile * ngrp;
    const uint64_t per = ((ntile + gridDim.x - 1) / gridDim.x) * ngrp;
    const uint64_t i0 = min(items, blockIdx.x * per), i1 = min(items, i0 + per);
    const uint32_t wave = threadIdx.x / 64;
    // thread -> (vertex, word) of an item: kFinalWords neighbouring threads
    // read one vertex's words of the group (contiguous)
    const uint32_t vl = threadIdx.x / kFinalWords, wl = threadIdx.x % kFinalWords;
    static_assert(kWideThreads == kFinalVerts * kFinalWords, "one (vertex, word) per thread");
    // Software pipeline: an item's loads are issued before the previous
    // item's row stores.
    // Only the slots of levels whose frontier held the vertex are read.
    // The level mask is loaded one item further ahead than the slots it
    // selects, so a slot load never waits on a mask load issued behind the
    // previous item's row stores (a wave's memory counter retires in order).
    u64 sw = 0, f[16];
    auto mask_of = [&](uint64_t item) -> uint32_t {
        const uint32_t v = (uint32_t)(item / ngrp) * kFinalVerts + vl;
        return (item < i1 && v < V && kept) ? p.lv[v] : 0u;
    };
    auto load = [&](uint64_t item, uint32_t lvm) {
        const uint32_t v = (uint32_t)(item / ngrp) * kFinalVerts + vl;
        const uint32_t w = (uint32_t)(item % ngrp) * kFinalWords + wl;
        const bool ok = item < i1 && v < V && w < W;
        const uint64_t idx = (uint64_t)v * W + w;
        sw = (ok && use_seen) ? p.seen[idx] : 0ull;
        lvm = ok ? lvm : 0u;
#pragma unroll
        for (uint32_t d = 0; d < 16; ++d)
            f[d] = (d < kept && ((lvm >> d) & 1u)) ? p.fr[(uint64_t)d * stride + idx] : 0ull;
    };
    uint32_t lv_next = mask_of(i0 + 1);
    load(i0, mask_of(i0));
    for (uint64_t item = i0; item < i1; ++item) {
        const uint32_t v0 = (uint32_t)(item / ngrp) * kFinalVerts;
        const uint32_t w0 = (uint32_t)(item % ngrp) * kFinalWords;
        const uint32_t v = v0 + vl, w = w0 + wl;
        const bool ok = v < V && w < W;
        const uint64_t idx = (uint64_t)v * W + w;
        u64 found = 0, pl0 = 0, pl1 = 0, pl2 = 0, pl3 = 0;
#pragma unroll
        for (uint32_t d = 0; d < 16; ++d) {
            found |= f[d];
            if (d & 1u) pl0 |= f[d];
            if (d & 2u) pl1 |= f[d];
            if (d & 4u) pl2 |= f[d];
            if (d & 8u) pl3 |= f[d];
        }
        if (ok) {
            if (use_seen ? sw != 0 : found != 0) p.seen[idx] = 0;
            // the tile's last word group clears the level mask (this thread's
            // load of it is done; the tile's other items ran before on this block)
            if (wl == 0 && kept && item % ngrp == ngrp - 1) p.lv[v] = 0;
            if (p.reach) {
                const u64 r = sw & word_valid(p.S, w);
                const uint64_t base = (uint64_t)v * p.reach_wpv + 2 * w;
                p.reach[base] = (uint32_t)r;
                if (2 * w + 1 < p.reach_wpv) p.reach[base + 1] = (uint32_t)(r >> 32);
            }
        }
        const u64 miss = ok ? ~(use_seen ? sw : found) & word_valid(p.S, w) : 0ull;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t sh = 8 * k;
            L_nib[(wl * 8 + k) * kFinalVerts + vl] =
                spread8x4((uint32_t)(pl0 >> sh)) | spread8x4((uint32_t)(pl1 >> sh)) << 1 |
                spread8x4((uint32_t)(pl2 >> sh)) << 2 | spread8x4((uint32_t)(pl3 >> sh)) << 3 |
                spread8x4((uint32_t)(miss >> sh)) * 15u;
        }
        L_put[wl * kFinalVerts + vl] = found | miss;
        // (the mask of item + 2 is read before this item clears any mask:
        // a tile's masks are cleared by its last word group, and item + 2
        // then belongs to the next tile)
        const uint32_t lv_after = mask_of(item + 2);
        load(item + 1, lv_next);
        lv_next = lv_after;
        __syncthreads();
        if (p.out) {
            if (p.bmap)
                final_tile_rows<T, T2, true>((T *)p.out, V, v0, w0, W, p.S, L_nib, L_put, wave,
                                             lane_id(), p.bmap, p.vgraph);
            else
                final_tile_rows<T, T2, false>((T *)p.out, V, v0, w0, W, p.S, L_nib, L_put, wave,
                                              lane_id(), nullptr, nullptr);
        }
        __syncthreads();
    }
}

hipError_t launch_wide_final(const WideParams &p, int blocks, hipStream_t st) {
    if (p.out_kind == 1)
        hipLaunchKernelGGL((wide_final_kernel<uint16_t, uint32_t>), dim3(blocks), dim3(kWideThreads),
                           0, st, p);
    else if (p.out_kind == 2)
        hipLaunchKernelGGL((wide_final_kernel<uint8_t, uint16_t>), dim3(blocks), dim3(kWideThreads),
                           0, st, p);
    else
        hipLaunchKernelGGL((wide_final_kernel<uint32_t, u64>), dim3(blocks), dim3(kWideThreads), 0,
                           st, p);
    return hipGetLastError();
}

template <int TW>
static hipError_t wide_launch_t(const WideParams &p, int blocks, int threads, size_t lds,
                                hipStream_t st) {
    void *args[] = {(void *)&p};
    return hipLaunchCooperativeKernel((const void *)wide_bfs_kernel<TW>, dim3(blocks),
                                      dim3(threads), args, lds, st);
}

template <int TW>
static hipError_t wide_steps_t(const WideParams &p, int blocks, int threads, size_t lds,
                               uint32_t levels, hipStream_t st) {
    hipError_t e = hipFuncSetAttribute((const void *)wide_step_kernel<TW, 0>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)wide_step_kernel<TW, 1>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    for (uint32_t l = 0; l < levels && e == hipSuccess; ++l) {
        hipLaunchKernelGGL(wide_ctl_kernel, dim3(1), dim3(1), 0, st, p, 0u);
        hipLaunchKernelGGL((wide_step_kernel<TW, 0>), dim3(blocks), dim3(threads), lds, st, p);
        hipLaunchKernelGGL((wide_step_kernel<TW, 1>), dim3(blocks), dim3(threads), lds, st, p);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(wide_ctl_kernel, dim3(1), dim3(1), 0, st, p, 1u);
        e = hipGetLastError();
    }
    return e;
}

template <int TW>
static hipError_t wide_occ_t(int threads, size_t lds, int *out) {
    hipError_t e = hipFuncSetAttribute((const void *)wide_bfs_kernel<TW>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, wide_bfs_kernel<TW>, threads, lds);
}

#define WIDE_SWITCH(FN, ...)                  \
    switch (tw) {                             \
    case 1: return FN<1>(__VA_ARGS__);        \
    case 2: return FN<2>(__VA_ARGS__);        \
    case 4: return FN<4>(__VA_ARGS__);        \
    case 8: return FN<8>(__VA_ARGS__);        \
    case 16: return FN<16>(__VA_ARGS__);      \
    case 32: return FN<32>(__VA_ARGS__);      \
    case 64: return FN<64>(__VA_ARGS__);      \
    default: return hipErrorInvalidValue;     \
    }

hipError_t launch_wide(int tw, const WideParams &p, int blocks, int threads, size_t lds,
                       hipStream_t st) {
    WIDE_SWITCH(wide_launch_t, p, blocks, threads, lds, st)
}

hipError_t launch_wide_steps(int tw, const WideParams &p, int blocks, int threads, size_t lds,
                             uint32_t levels, hipStream_t st) {
    WIDE_SWITCH(wide_steps_t, p, blocks, threads, lds, levels, st)
}

hipError_t wide_max_blocks_per_cu(int tw, int threads, size_t lds, int *out) {
    WIDE_SWITCH(wide_occ_t, threads, lds, out)
}

}  // namespace msbfs
