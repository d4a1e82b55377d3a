// Shared device/host helpers for the DEER-over-MALA kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PMCMC_NT 256  // threads per workgroup of the tile kernels (4 wave64)

namespace pmcmc {

// ---------------------------------------------------------------------------
// Counter-based Rademacher probes.  A probe entry is a pure function of
// (seed, Newton iteration, probe index, t*D + d), so the device generates the
// same matrix the CPU oracle does, with no host RNG and no host->device copy
// (the reference draws z from the sampler's host RNG and uploads it every
// iteration, src/DEER/DEER.jl:298-325, 407).
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

__host__ __device__ inline uint32_t probe_key(uint64_t seed, uint32_t it, uint32_t k) {
    uint32_t lo = (uint32_t)(seed & 0xFFFFFFFFull), hi = (uint32_t)(seed >> 32);
    uint32_t a = fmix32(lo ^ 0x3C6EF372u);
    uint32_t b = a + hi * 0x9E3779B1u + it * 0x85EBCA77u + k * 0xC2B2AE3Du;
    return fmix32(b);
}

// +1 or -1 for global element index idx = t*D + d of one probe matrix.
__host__ __device__ inline float rademacher(uint32_t key, uint64_t idx) {
    uint32_t h = fmix32(((uint32_t)idx * 0x9E3779B1u) ^ key);
    return (h >> 31) ? 1.0f : -1.0f;
}

// ---------------------------------------------------------------------------
// Cross-workgroup state of one single-pass scan launch (decoupled look-back)
// and of the Newton loop's device-side convergence test.  Counters are reset
// by the last workgroup of each launch, so launches need no memset.
// ---------------------------------------------------------------------------
struct SyncState {
    unsigned int ticket;  // dynamic tile index
    unsigned int done;    // finished workgroups of the current launch
    unsigned int epoch;   // launch counter; look-back flags carry it
    unsigned int err;     // nonzero: a bounded wait expired
    unsigned long long dmax_bits;  // max |S_new - S| of this iteration (double bits)
    unsigned long long smax_bits;  // max |S_new| (double bits)
    int converged;        // set by the iteration that met metric <= 1
    int conv_iter;        // that iteration (0-based)
    double last_metric;
    int last_iter;
    int pad;
};

// flag word: (epoch << 2) | status
#define PMCMC_ST_AGG 1u
#define PMCMC_ST_PREFIX 2u

__device__ inline unsigned ld_acquire(const unsigned* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_release(unsigned* p, unsigned v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
struct DevTraits;
template <>
struct DevTraits<float> {
    static constexpr int tile_elems = 4096;
};
template <>
struct DevTraits<double> {
    static constexpr int tile_elems = 2048;
};

}  // namespace pmcmc
