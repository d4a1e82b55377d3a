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
// Device-side state of the Newton loop: convergence maxima of the current
// iteration, the arrival counter that elects the workgroup which closes it,
// and the convergence flag that turns later queued iterations into no-ops.
// ---------------------------------------------------------------------------
struct SyncState {
    unsigned int done;             // arrived workgroups of the current iteration
    unsigned int err;
    unsigned long long dmax_bits;  // max |S_new - S| of this iteration (double bits)
    unsigned long long smax_bits;  // max |S_new| (double bits)
    int converged;                 // set by the iteration that met metric <= 1
    int conv_iter;                 // that iteration (0-based)
    double last_metric;
    int last_iter;
    int pad;
};

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
