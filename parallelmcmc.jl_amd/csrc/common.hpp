// Shared device/host helpers for the DEER-over-MALA kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

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
    // Final prefix of a structured (full-Jacobian) solve: columns below
    // skip_cols had residual FT - S exactly 0, so they sit on the fixed point in
    // both iterate buffers and later iterations leave them out.  live_from
    // collects this iteration's lowest tile start with a nonzero residual.
    int live_from;
    int skip_cols;
    int pad;
    double smax_prefix;  // max |S| over final columns that the current tiling no longer covers
};

// A workgroup's arrival at the end of an iteration's output pass (one thread
// per workgroup).  It hands its maxima of |S_new - S| and |S_new| (and, with
// live_col >= 0, the start of its tile, which still has a nonzero residual)
// to the last workgroup to arrive; only that one returns true, with the
// iteration's maxima read and reset.  Everything handed over is an
// agent-scope atomic, coherent across the XCDs' L2s, so the arrival needs no
// L2 write-back or invalidation (an acq_rel arrival costs both, in every
// workgroup): the atomics return their old values, so once the wave has
// waited for them (vmcnt(0)) they have been performed, and only then does the
// arrival count.  Anything else the last workgroup reads must have been
// written by an earlier launch or by its own thread.
__device__ inline bool sync_arrive(SyncState* sy, unsigned nblocks, double dmax, double smax, int live_col,
                                   double* dm, double* sm) {
    int o2 = 0;
    if (live_col >= 0) o2 = __hip_atomic_fetch_min(&sy->live_from, live_col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // non-negative doubles order like their bit patterns; NaN sorts above +Inf
    const unsigned long long o0 = __hip_atomic_fetch_max(&sy->dmax_bits, (unsigned long long)__double_as_longlong(dmax),
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long o1 = __hip_atomic_fetch_max(&sy->smax_bits, (unsigned long long)__double_as_longlong(smax),
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    asm volatile("" ::"v"(o0), "v"(o1), "v"(o2));  // keeps the returning forms
    __builtin_amdgcn_s_waitcnt(0x0F70);            // vmcnt(0); expcnt and lgkmcnt not waited on
    const unsigned prev = __hip_atomic_fetch_add(&sy->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (prev != nblocks - 1u) return false;
    // read and reset at once: the next launch's first atomics see zero
    *dm = __longlong_as_double(
        (long long)__hip_atomic_exchange(&sy->dmax_bits, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    *sm = __longlong_as_double(
        (long long)__hip_atomic_exchange(&sy->smax_bits, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    __hip_atomic_store(&sy->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// The last workgroup's read and reset of live_from (0x7fffffff: no tile of the
// iteration had a nonzero residual).
__device__ inline int sync_take_live_from(SyncState* sy) {
    return __hip_atomic_exchange(&sy->live_from, 0x7fffffff, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Sliding window of one chain of a batched solve.  The Newton iteration runs
// over the columns [front, hi) of the chain's whole trajectory, hi =
// min(front + span, T).  After each iteration the window's first columns
// whose residual FT - S was exactly 0 (its final prefix, at tile granularity)
// are final, and the whole window is accepted once the convergence metric is
// <= 1 (the reference's block test, src/DEER/DEER.jl:567-577); either way the
// front moves past them and the window slides on.  The device moves `front`
// (last workgroup of the apply pass) and `hi` (window kernel, from the span
// the host last chose); accepted columns [pend_lo, pend_hi) are copied into
// the other iterate buffer by the next window kernel, and columns entering
// the window for the first time (at or beyond `hw`) start from the state just
// before them (the reference's block start, S = broadcast(s0)).
// ---------------------------------------------------------------------------
struct SlideState {
    long long front, hi, T;
    long long hw;                // columns below hw hold an iterate or an initial guess
    long long pend_lo, pend_hi;  // accepted columns to copy into the other buffer
    int seg_start;               // first iteration of the current window segment
    int accepts;                 // windows accepted by the metric test
    int last_adv_iter;           // last iteration that moved the front
    int pad;
};

// ---------------------------------------------------------------------------
// Optional per-kernel timing with HIP events on the launching stream.  Pairs
// are tagged with the Newton iteration so that launches queued after
// convergence (which exit at once) are not counted.
// ---------------------------------------------------------------------------
enum { KT_GLM_PASS = 0, KT_TILE_REDUCE = 1, KT_TILE_APPLY = 2, KT_GLM_REDUCE = 3, KT_SCAN_SPINE = 4,
       KT_ITERATION = 5, KT_LAMBDA_PASS = 6, KT_COUNT = 7 };

struct Prof {
    bool on = false;
    unsigned kinds = ~0u;  // bit k: record pairs of kind k
    struct Pair {
        hipEvent_t a, b;
        int kind, iter;
    };
    std::vector<Pair> pending;
    std::vector<hipEvent_t> pool;
    double ms[KT_COUNT] = {0};
    long long n[KT_COUNT] = {0};

    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    // returns the index of the open pair, or -1
    int begin(int kind, int iter, hipStream_t s) {
        if (!on || !((kinds >> kind) & 1u)) return -1;
        Pair p{get(), get(), kind, iter};
        (void)hipEventRecord(p.a, s);
        pending.push_back(p);
        return (int)pending.size() - 1;
    }
    void end(int idx, hipStream_t s) {
        if (idx >= 0) (void)hipEventRecord(pending[idx].b, s);
    }
    // after the stream has been synchronised: count pairs of iterations <= last_iter
    void flush(int last_iter) {
        for (auto& p : pending) {
            if (p.iter <= last_iter) {
                float t = 0.f;
                if (hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
                    ms[p.kind] += t;
                    n[p.kind] += 1;
                }
            }
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    void reset() {
        for (int k = 0; k < KT_COUNT; ++k) {
            ms[k] = 0;
            n[k] = 0;
        }
    }
    ~Prof() {
        for (auto& p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Pointers that kernels read from descriptor tables in device memory are
// generic (flat) pointers to the compiler: flat accesses also count against
// the LDS counter, and guarded flat loads get a full wait in front of each
// one, which serialises a workgroup's loads.  Hot loops access memory through
// these global-address-space views instead.
#define PMCMC_GLOBAL __attribute__((address_space(1)))
typedef float pmcmc_f4v __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ const PMCMC_GLOBAL T* gp(const T* p) {
    return (const PMCMC_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ PMCMC_GLOBAL T* gp(T* p) {
    return (PMCMC_GLOBAL T*)p;
}
typedef double pmcmc_d2v __attribute__((ext_vector_type(2)));
// 16-byte load of four floats (p 16-byte aligned)
__device__ __forceinline__ pmcmc_f4v gld4(const float* p) { return *(const PMCMC_GLOBAL pmcmc_f4v*)p; }
// 16-byte vector types and loads by element type
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef pmcmc_f4v type;
    static constexpr int n = 4;
};
template <>
struct Vec16<double> {
    typedef pmcmc_d2v type;
    static constexpr int n = 2;
};
template <typename T>
__device__ __forceinline__ typename Vec16<T>::type gld16(const T* p) {
    return *(const PMCMC_GLOBAL typename Vec16<T>::type*)p;
}

// max that propagates NaN from either operand (fmax drops it): a NaN state
// must fail the convergence test, as maximum() does in src/DEER/DEER.jl:567-571
__host__ __device__ inline double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }

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

// One contiguous span of ncol columns (column-major, D floats each, from
// src) into LDS columns dst[col * SV + k] (col < ncol, k < D): 16-byte
// loads, every one of a round (up to 8 x 256 vectors) in flight before the
// first LDS write.  span_issue / span_store split a round in two, so that a
// kernel can put other loads (and VALU work) between them; span_to_lds is
// both, round by round.
struct SpanRegs {
    pmcmc_f4v r[8];
    float hv, tv;
    int head, nv, n;
};

__device__ __forceinline__ void span_issue(SpanRegs& s, const float* src, int D, int ncol, int tid, int base = 0) {
    s.n = D * ncol;
    if (s.n <= 0) {
        s.head = s.nv = 0;
        return;
    }
    const uintptr_t addr = (uintptr_t)src;
    int head = (int)(((16 - (addr & 15)) & 15) / 4);
    if (head > s.n) head = s.n;
    s.head = head;
    s.nv = (s.n - head) / 4;
    const float* sv = src + head;
    // only the span's own vectors are read (a wave past the span issues
    // nothing): clamped duplicates would still pass through the CU's address
    // and data path, which a short kernel's loads saturate
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int idx = base + tid + 256 * q;
        s.r[q] = pmcmc_f4v{0.f, 0.f, 0.f, 0.f};
        if (idx < s.nv) s.r[q] = gld4(sv + 4 * idx);
    }
    s.hv = s.tv = 0.f;
    if (base == 0) {
        const int ti = head + s.nv * 4 + tid;
        if (tid < head) s.hv = gp(src)[tid];
        if (ti < s.n) s.tv = gp(src)[ti];
    }
}

template <int SV>
__device__ __forceinline__ void span_store(float* dst, const SpanRegs& s, int D, int tid, int base = 0) {
    if (s.n <= 0) return;
    // element i of the span -> column i / D, row i % D (32-bit, once per vector)
    auto put4 = [&](int i0, const float* v, int cnt) {
        int col = (int)((unsigned)i0 / (unsigned)D);
        int k = i0 - col * D;
        for (int e = 0; e < cnt; ++e) {
            dst[col * SV + k] = v[e];
            if (++k == D) {
                k = 0;
                ++col;
            }
        }
    };
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int idx = base + tid + 256 * q;
        if (idx < s.nv) {
            const float v[4] = {s.r[q].x, s.r[q].y, s.r[q].z, s.r[q].w};
            put4(s.head + 4 * idx, v, 4);
        }
    }
    if (base == 0) {
        const int ti = s.head + s.nv * 4 + tid;
        if (tid < s.head) put4(tid, &s.hv, 1);
        if (ti < s.n) put4(ti, &s.tv, 1);
    }
}

template <int SV>
__device__ __forceinline__ void span_to_lds(float* dst, const float* src, int D, int ncol, int tid) {
    SpanRegs s;
    span_issue(s, src, D, ncol, tid);
    span_store<SV>(dst, s, D, tid);
    for (int base = 8 * 256; base < s.nv; base += 8 * 256) {
        span_issue(s, src, D, ncol, tid, base);
        span_store<SV>(dst, s, D, tid, base);
    }
}


}  // namespace pmcmc
