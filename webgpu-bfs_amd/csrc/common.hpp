// Shared host-side helpers of libmsbfs: status/error plumbing and the
// seeded random streams used by the graph generators.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "msbfs.h"

namespace msbfs {

// Records the message of the last failure on the calling thread and returns
// the status code so call sites can write `return fail(code, ...)`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// splitmix64: seeds xoshiro256** from one 64-bit value (Vigna, 2015).
inline uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// xoshiro256** 1.0 (Blackman & Vigna, "Scrambled linear pseudorandom number
// generators", 2018).  Each generator stream in this library is one of these
// seeded through splitmix64 from (seed, stream id).
class Rng {
public:
    explicit Rng(uint64_t seed, uint64_t stream = 0) {
        uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull);
        for (auto &w : s_) w = splitmix64(x);
    }
    uint64_t next() {
        const uint64_t r = rotl(s_[1] * 5, 7) * 9;
        const uint64_t t = s_[1] << 17;
        s_[2] ^= s_[0];
        s_[3] ^= s_[1];
        s_[1] ^= s_[2];
        s_[0] ^= s_[3];
        s_[2] ^= t;
        s_[3] = rotl(s_[3], 45);
        return r;
    }
    // Uniform double in [0, 1) with 53 random bits.
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // Uniform integer in [0, n) by 128-bit multiply (Lemire's method without
    // the rejection step; bias < n / 2^64).
    uint64_t below(uint64_t n) {
        return (uint64_t)(((unsigned __int128)next() * (unsigned __int128)n) >> 64);
    }

private:
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t s_[4];
};

// Host CSR owned by the library (returned through msbfs_host_graph*).
struct HostCsr {
    uint64_t V = 0;
    std::vector<uint32_t> ro;  // V + 1
    std::vector<uint32_t> ci;  // E
    uint64_t E() const { return ci.size(); }
};

// Canonical CSR from directed pairs: drops self loops and out-of-range
// endpoints, sorts every row and removes duplicates (graph.rs:51-80 and
// generators.rs:559-588 give the same result).
void csr_from_pairs(uint64_t V, const std::vector<uint64_t> &pairs, HostCsr &out, int threads);
// Counting-sort transpose (graph.rs:90-126).
void csr_transpose(uint64_t V, const uint32_t *ro, const uint32_t *ci, uint32_t *co,
                   uint32_t *ri);
bool csr_is_symmetric(uint64_t V, uint64_t E, const uint32_t *ro, const uint32_t *ci);
// Lattice offsets: when every edge (u, v) with v > u has v - u among at most
// four distinct values, writes them ascending to d[0..k) (the rest 0) and
// returns k (0 for a graph with no such edge); returns -1 otherwise.
int csr_lattice_offsets(uint64_t V, const uint32_t *ro, const uint32_t *ci, uint32_t d[4]);
// Levels of a breadth-first search from s over the CSR (the eccentricity of
// s in its component, plus one); 0 when V is 0.
uint32_t host_levels_from(uint64_t V, const uint32_t *ro, const uint32_t *ci, uint32_t s);
int host_threads();
// Hands a built CSR to the C API (msbfs_host_graph_get / _free).
msbfs_host_graph *wrap_host_graph(HostCsr &&c);

}  // namespace msbfs

struct msbfs_host_graph {
    msbfs::HostCsr csr;
};
