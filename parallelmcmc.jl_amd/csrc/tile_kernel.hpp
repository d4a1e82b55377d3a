// Tile kernels of the diagonal affine scan, optionally fused with the DEER
// Newton update of a taped MALA recursion.
//
//   s_t = a_t .* s_{t-1} + b_t   (rows independent; D x T column-major)
//
// One workgroup owns TC consecutive columns (all D rows).  Those columns are
// contiguous in HBM, so they are staged into LDS with fully coalesced loads
// and scanned in LDS (rows x P parts, then a Hillis-Steele combine of the
// parts).  The scan across tiles is a reduce-then-scan:
//   PASS_REDUCE: per tile, the composed map (prod a, combined b) of each row;
//   the D x nTiles aggregates are scanned recursively by the same kernels
//   (they form a D x nTiles diagonal scan); then
//   PASS_APPLY: per tile, the incoming state is the previous tile's end state
//   and the tile is re-scanned from it and written out.
//   PASS_SINGLE: a problem that fits one tile, scanned from s0 directly.
// No workgroup ever waits on another (no spin loops).  PASS_APPLY re-reads
// what PASS_REDUCE read just before; at these working-set sizes the re-read
// is served by the 256 MB last-level cache, so HBM sees close to one read of
// every input and one write of the output.
//
// Modes
//   MODE_SCAN: A, B read from HBM.                       (DEERScan.jl:102-173)
//   MODE_ELEM: A, B built on chip from the iterate S_in, the tape (Xi, U) and
//              on-chip Rademacher probes, for targets with a diagonal Hessian
//              (MALA.jl:556-628 + DEER.jl:398-419 fused with the scan, the
//              damping DEER.jl:470-473 and the convergence maxima
//              DEER.jl:567-571).
//   MODE_GEMM: as MODE_ELEM but the gradients / HVPs / log-densities of the
//              current and proposed states come precomputed by the GEMM
//              target kernels.
// The DEER modes scan the Newton correction, not the state:
//   d_t = A_t d_{t-1} + (FT_t - S_t),  d_{-1} = 0,  S_new_t = S_t + damping d_t
// (S_t replaced by 0 where it is not finite)
// (the same update as s_t = A_t s_{t-1} + FT_t - A_t Xbar_t with
// s_t = S_t + d_t).  Its fixed point is exact in floating point: once the
// prefix up to t-1 has stopped moving, d_t = FT_t - S_t is exact (the two are
// close, Sterbenz) and S_t + d_t lands on FT_t and stays there, whatever the
// rounding of A_t.  The state form lands within an ulp of FT_t that moves
// with every iteration's probes, and that jitter alone can flip an accept
// decision sitting within ~1e-6 of its threshold, for ever.
// MODE_SCAN with s0 == nullptr starts from zero.
#pragma once

#include <climits>

#include <type_traits>

#include "common.hpp"
#include "refine.hpp"

namespace pmcmc {

enum { MODE_SCAN = 0, MODE_ELEM = 1, MODE_GEMM = 2 };
enum { PASS_REDUCE = 0, PASS_APPLY = 1, PASS_SINGLE = 2, PASS_CHAINED = 3 };
// spin cap of one look-back poll: far beyond any legitimate wait (which is
// microseconds); hitting it sets lb_err and the scan reports an error
#define PMCMC_LB_SPIN_CAP (1u << 21)
enum { JAC_STOCH = 0, JAC_DIAG = 1 };
#define PMCMC_MAXP 4
#ifndef PMCMC_FAST_RM
#define PMCMC_FAST_RM 16
#endif
#ifndef PMCMC_BATCHED_WAVES
#define PMCMC_BATCHED_WAVES 2
#endif
// Base of the Newton correction of one entry: the iterate, or 0 where the
// iterate is not finite, so that a non-finite region of S_in shrinks by (at
// least) one column per iteration, as it does in the state form.
template <typename T>
__device__ inline T corr_base(T v) {
    return isfinite(v) ? v : (T)0;
}

// most log-density partials per column (64-row slabs of a D <= 256 target)
#define PMCMC_LP_MAXSLAB 4
// most tiles whose incoming states an apply pass folds in place (agg_inline)
#define PMCMC_AGG_INLINE 16

// Operands of one column in the fast GEMM build: RM rows per lane.
template <typename T, int RM>
struct FastOps {
    T gx[RM], y[RM], gy[RM], w[RM], hw[RM];
    double lpx, lpy;
    T u;  // the tape's uniform, in the iterate's precision
    int8_t o, pv;
};

template <typename T>
struct TileArgs {
    // scan operands (MODE_SCAN)
    const T* A;
    const T* B;
    // common
    const T* s0;
    T* S;           // PASS_APPLY / PASS_SINGLE output
    T* agg_a;       // PASS_REDUCE output: D x numTiles
    T* agg_b;
    const T* prev;  // PASS_APPLY: D x numTiles end states of the tiles (column j = after tile j)
    int64_t D;
    int64_t T_len;
    int TC;        // columns per tile
    int numTiles;
    int P;         // parts per row in the in-tile scan
    const int* skip;  // nonzero => the kernel is a no-op (converged Newton loop)
    // DEER (MODE_ELEM / MODE_GEMM)
    SyncState* sync;
    const T* S_in;
    const T* Xi;
    const T* U;
    uint8_t* acc_out;
    double* la_out;
    double eps, damping, tol_abs, tol_rel;
    int iter;
    uint32_t probe_iter;  // iteration field of the probe keys (fixed for fixed probes)
    int probes;
    int jacobian;
    uint64_t seed;
    // diagonal-Hessian target (nullable: standard normal)
    const T* prec;
    const T* mean;
    // PASS_CHAINED (single pass with decoupled look-back): per-tile status
    // (0 none, 1 aggregate, 2 inclusive state), aggregates and inclusive
    // states (D x numTiles), the tile ticket and an error flag; zeroed before
    // each launch
    int* lb_status;
    T* lb_agg_a;
    T* lb_agg_b;
    T* lb_incl;
    unsigned* lb_ticket;
    unsigned* lb_err;
    // two-level look-back of the register-tile scan: per super-tile (64
    // consecutive tiles) status (1 aggregate, 2 end state), aggregate maps;
    // its end states go in lb_incl (D x super-tiles)
    int* lb_sstatus;
    T* lb_sagg_a;
    T* lb_sagg_b;
    // GEMM target products (MODE_GEMM): D x T, probe k of W/hw at k*D*T
    const T* Y;
    const T* gx;
    const T* gy;
    const T* W;   // w_k = z_k + eps * H(x) z_k
    const T* hw;  // H(y) w_k
    const double* lpx;
    const double* lpy;
    // log-densities in lp_nslab partials (lp[s * lp_stride + t], summed in
    // slab order; at most PMCMC_LP_MAXSLAB); 0 or 1: a single value per column
    int lp_nslab;
    int64_t lp_stride;
    // MODE_GEMM: FP64 decision override per column (-1 none; else bit 0 the
    // decision, bit 1 set when |log alpha - log u| < hyst), or null
    const int8_t* dec;
    // Hysteresis of FP32 iterates (null / 0 for FP64): where |log alpha - log u|
    // < hyst the previous iteration's decision (prev_dec, -1 none) is kept.
    // Near the fixed point the last ulps of the iterate settle one column per
    // iteration from the front; until they do, log alpha of a later column
    // moves by ~1e-7 per ulp, and a decision that close to its threshold would
    // flip back and forth, each flip moving the whole tail.
    int8_t* prev_dec;
    double hyst;
    // with dec: nonzero once the FP64 override is active; an iteration that
    // meets the tolerance before then does not count as converged
    const int* refine_on;
    // MODE_GEMM multi-tile solves (null otherwise): PASS_REDUCE stores the
    // tile's A and B (D x T each) here and writes the per-column outputs
    // (decisions, log alpha, prev_dec); PASS_APPLY then stages A and B instead
    // of building them again.  They may alias W and hw: a tile reads and
    // writes only its own columns, and writes them after it has read them.
    T* a_keep;
    T* b_keep;
    // Final prefix (DEER modes, two-pass tiles; null otherwise): columns below
    // *final_cols had a residual FT - S of exactly 0, so both iterate buffers
    // hold the fixed point there and no later iteration changes them.  Their
    // tiles are left out: PASS_REDUCE writes the identity aggregate (Delta
    // stays 0 through it), PASS_APPLY only adds the tile's cached max |S|
    // (tile_smax, numTiles entries, written whenever the tile runs) to the
    // metric.  The last workgroup moves the floor to the first tile with a
    // nonzero residual, once the FP64 decision override (refine_on) is active.
    const int* final_cols;
    double* tile_smax;
    // Sliding window of a batched chain (null otherwise): the last workgroup
    // moves the window's front past its final prefix, or past the whole
    // window when the metric test passes, and flags the chain done at T.
    SlideState* win;
    // element offset of column 0 in the probe stream (window of a longer
    // trajectory: probes stay attached to absolute columns)
    int64_t zoff;
    // Inline FP64 decisions (dense-Gaussian chains of a batched solve; null
    // otherwise): the pass that forms the decisions re-decides, in FP64, every
    // column whose log alpha is within `band` of log u, once the gate is open
    // (the previous update within `gate` tolerance units, or iteration
    // `gate_iters`; then sticky in *rflag), and rebuilds the column's A and B
    // if the decision changed.  The same FP64 quantities as the separate
    // refinement kernels (the state is the FP32 iterate, exact in double).
    const RefineTarget* rtg;
    int* rflag;
    double band, gate;
    int gate_iters;
    // PASS_APPLY with at most PMCMC_AGG_INLINE tiles and D <= PMCMC_NT
    // (nonzero): a tile's incoming state is folded from the aggregates of the
    // tiles before it (agg_a, agg_b) in the tile itself, in the order the
    // aggregate scan would fold them, instead of being read from `prev`.
    int agg_inline;
};

// FP64 log alpha of one column for a dense Gaussian target, by the whole
// workgroup: the state x (FP32 in LDS, exact in double) and tape column xi;
// v: D doubles of LDS scratch, red: 4 doubles.  One row per thread in the two
// D x D mat-vecs of Lambda (column-major, coalesced over the rows).
template <typename T>
__device__ double gauss_logalpha_fp64(const RefineTarget& tg, const T* x, const T* xi, double eps, double* v,
                                      double* red) {
    const int tid = threadIdx.x;
    const int64_t D = tg.D;
    auto bsum = [&](double q) {
        for (int m = 32; m >= 1; m >>= 1) q += __shfl_xor(q, m, 64);
        if ((tid & 63) == 0) red[tid >> 6] = q;
        __syncthreads();
        const double r = (red[0] + red[1]) + (red[2] + red[3]);
        __syncthreads();
        return r;
    };
    const double sq2e = sqrt(2.0 * eps);
    for (int64_t d = tid; d < D; d += PMCMC_NT) v[d] = (double)x[d] - (tg.mu64 ? tg.mu64[d] : 0.0);
    __syncthreads();
    double g0 = 0.0, yv = 0.0, lpp = 0.0;  // thread tid owns row tid (D <= 256)
    if (tid < D) {
        double p = 0.0;
        for (int64_t k = 0; k < D; ++k) p += tg.L64[tid + k * D] * v[k];
        g0 = -p;
        lpp = 0.5 * v[tid] * g0;
        yv = (double)x[tid] + eps * g0 + sq2e * (double)xi[tid];
    }
    const double lp0 = bsum(lpp);  // its barriers order the reads of v before the rewrite
    if (tid < D) v[tid] = yv - (tg.mu64 ? tg.mu64[tid] : 0.0);
    __syncthreads();
    double lq = 0.0, qyx = 0.0, qxy = 0.0;
    if (tid < D) {
        double p = 0.0;
        for (int64_t k = 0; k < D; ++k) p += tg.L64[tid + k * D] * v[k];
        const double g = -p;
        lq = 0.5 * v[tid] * g;
        const double xx = (double)x[tid];
        const double e1 = yv - xx - eps * g0;
        const double e2 = xx - yv - eps * g;
        qyx = e1 * e1;
        qxy = e2 * e2;
    }
    const double lp1 = bsum(lq);
    qyx = bsum(qyx);
    qxy = bsum(qxy);
    const double c_norm = ((double)D / 2.0) * log(4.0 * 3.14159265358979323846 * eps);
    return (lp1 + (-0.5 * qxy / (2.0 * eps) - c_norm)) - (lp0 + (-0.5 * qyx / (2.0 * eps) - c_norm));
}

// LDS layout (elements of T):
//   xs  : (TC+1)*D  iterate columns c0-1 .. c0+TC-1 (Xbar and S_old); DEER modes only
//   bl  : TC*D      B (DEER modes: the tape Xi first, overwritten by B)
//   al  : TC*D      A, then the output tile
//   pa,pb: max(NT, D) part aggregates
//   sin : D         incoming state per row
//   acc : TC bytes  accept flags
//   cmp : 2*D       composed predecessor aggregates (PASS_CHAINED only; at a
//                   16-byte aligned offset after acc)
template <typename T>
__host__ __device__ inline size_t tile_lds_core(int64_t D, int TC, bool with_iterate) {
    int64_t PA = D > PMCMC_NT ? D : PMCMC_NT;
    int64_t xs = with_iterate ? ((int64_t)TC + 1) * D : 0;
    size_t n = (size_t)(xs + 2 * (int64_t)TC * D + 2 * PA + D);
    return (n * sizeof(T) + (size_t)TC + 15) / 16 * 16;
}
template <typename T>
__host__ __device__ inline size_t tile_lds_bytes(int64_t D, int TC, bool with_iterate, bool chained = false) {
    return tile_lds_core<T>(D, TC, with_iterate) + (chained ? 2 * (size_t)D * sizeof(T) : 0) + 64;
}

// Level-0 tile shape of a D x Tn problem: TC columns per tile (as many as fit
// DevTraits<T>::tile_elems elements and the LDS budget), P parts per row.
template <typename T>
inline void tile_shape(int64_t D, int64_t Tn, bool with_iterate, int* TC, int* P, int* numTiles) {
    int64_t E = DevTraits<T>::tile_elems;
    int64_t tc = E / D;
    if (tc < 1) tc = 1;
    while (tc > 1 && tile_lds_bytes<T>(D, (int)tc, with_iterate) > 64512) tc >>= 1;
    if (tc > Tn && Tn > 0) tc = Tn;
    *TC = (int)tc;
    *P = D >= PMCMC_NT ? 1 : (int)(PMCMC_NT / D);
    if (*P > tc) *P = (int)tc;
    if (*P < 1) *P = 1;
    *numTiles = (int)((Tn + tc - 1) / tc);
}

template <typename T>
__device__ inline T wave_max(T v) {
    for (int m = 32; m >= 1; m >>= 1) v = nan_max(v, __shfl_xor(v, m, 64));
    return v;
}

// Sum over a group of G lanes (G a power of two <= 64, groups aligned).
__device__ inline double group_sum(double v, int G) {
    for (int m = G >> 1; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Copy n elements from global memory into LDS: 16-byte vector loads, eight
// per thread in flight before the LDS writes (the staging loop is otherwise
// bound by one global-load latency per element).
template <typename T>
__device__ inline void stage_lds(T* dst, const T* src, int64_t n, int tid) {
    typedef typename std::conditional<sizeof(T) == 4, float4, double2>::type V;
    constexpr int VEC = 16 / sizeof(T);
    const uintptr_t addr = (uintptr_t)src;
    int64_t head = (int64_t)(((16 - (addr & 15)) & 15) / sizeof(T));
    if (head > n) head = n;
    for (int64_t i = tid; i < head; i += PMCMC_NT) dst[i] = src[i];
    const int64_t nv = (n - head) / VEC;
    const V* sv = reinterpret_cast<const V*>(src + head);
    T* dh = dst + head;
    for (int64_t base = 0; base < nv; base += 8 * PMCMC_NT) {
        V r[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t idx = base + tid + (int64_t)k * PMCMC_NT;
            if (idx < nv) r[k] = sv[idx];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t idx = base + tid + (int64_t)k * PMCMC_NT;
            if (idx < nv) {
                const T* rv = reinterpret_cast<const T*>(&r[k]);
#pragma unroll
                for (int j = 0; j < VEC; ++j) dh[idx * VEC + j] = rv[j];
            }
        }
    }
    for (int64_t i = head + nv * VEC + tid; i < n; i += PMCMC_NT) dst[i] = src[i];
}

// Copy up to NR regions from global memory into LDS in one round trip: every
// region's 16-byte loads (up to eight per thread per region and round) are in
// flight before the first LDS write.  Regions need no common alignment.
template <typename T, int NR>
__device__ inline void stage_lds_multi(T* const (&dst)[NR], const T* const (&src)[NR], const int64_t (&n)[NR],
                                       int tid) {
    typedef typename Vec16<T>::type V;
    constexpr int VEC = Vec16<T>::n;
    int64_t head[NR], nv[NR];
    const T* s0[NR];  // 16-byte aligned start of each region's vectors
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const uintptr_t addr = (uintptr_t)src[r];
        head[r] = (int64_t)(((16 - (addr & 15)) & 15) / sizeof(T));
        if (head[r] > n[r]) head[r] = n[r];
        nv[r] = (n[r] - head[r]) / VEC;
        // with no full vector, loads go to the aligned block holding src[0]
        // (same page), and their values are not used
        s0[r] = nv[r] > 0 ? src[r] + head[r] : (const T*)(addr & ~(uintptr_t)15);
    }
    int64_t nmax = 0;
#pragma unroll
    for (int r = 0; r < NR; ++r) nmax = nv[r] > nmax ? nv[r] : nmax;
    // one round: every region's loads issued (global address space, indices
    // clamped to valid vectors, no branches), then the LDS writes
    auto round = [&](int64_t base) {
        V rv[NR][8];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int64_t last = nv[r] > 0 ? nv[r] - 1 : 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int64_t idx = base + tid + (int64_t)k * PMCMC_NT;
                rv[r][k] = gld16<T>(s0[r] + (idx < last ? idx : last) * VEC);
            }
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            T* dh = dst[r] + head[r];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int64_t idx = base + tid + (int64_t)k * PMCMC_NT;
                if (idx < nv[r]) {
#pragma unroll
                    for (int j = 0; j < VEC; ++j) dh[idx * VEC + j] = rv[r][k][j];
                }
            }
        }
    };
    // the unaligned heads and tails: at most one element per thread each
    T hs[NR], ts[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int64_t ti = head[r] + nv[r] * VEC + tid;
        hs[r] = gp(src[r])[tid < head[r] ? tid : 0];
        ts[r] = gp(src[r])[ti < n[r] ? ti : 0];
    }
    round(0);  // straight-line: the common case
    for (int64_t base = 8 * PMCMC_NT; base < nmax; base += 8 * PMCMC_NT) round(base);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (tid < head[r]) dst[r][tid] = hs[r];
        const int64_t ti = head[r] + nv[r] * VEC + tid;
        if (ti < n[r]) dst[r][ti] = ts[r];
    }
}

// Publication between workgroups of one launch: values are written and read
// with agent-scope relaxed atomics (coherent across the XCDs' L2s without an
// L2 write-back per publication).  Before the status word is stored, each
// publishing thread waits until its stores have completed (vmcnt(0)): a
// workgroup-scope fence does not wait for global stores, and without the wait
// a reader can see the status before the values.  The barrier that follows
// then orders the status store after every publishing thread's wait.
template <typename T>
__device__ inline void lb_store(T* p, T v) {
    if constexpr (sizeof(T) == 8)
        __hip_atomic_store((unsigned long long*)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    else
        __hip_atomic_store((unsigned*)p, __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void lb_publish_fence() {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt and lgkmcnt not waited on
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
}

// Coherent load of a value another workgroup published.
template <typename T>
__device__ inline T lb_load(const T* p) {
    if constexpr (sizeof(T) == 8) {
        const unsigned long long v =
            __hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return __longlong_as_double((long long)v);
    } else {
        const unsigned v = __hip_atomic_load((const unsigned*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return __uint_as_float(v);
    }
}

// Decoupled look-back of tile `tile` (tiles numbered in start order by a
// ticket, so every predecessor is resident and never waits on a successor).
// ga/gb: this tile's aggregate per row.  Publishes the aggregate (status 1),
// composes predecessor aggregates over windows of 256 tiles (one status word
// per thread) back to the nearest published inclusive state, sets sin_ to the
// incoming state, then publishes this tile's inclusive state (status 2).
// Inclusive states advance one window per look-back round trip, so the
// window is as wide as the workgroup.
template <typename T>
__device__ void chained_lookback(const TileArgs<T>& a, const int tile, const int64_t D, const T* ga, const T* gb,
                                 T* sin_, T* cmp_a, T* cmp_b) {
    __shared__ int lb_k, lb_lo;
    __shared__ int lb_w[PMCMC_NT / 64];
    const int tid = threadIdx.x;
    int* st = a.lb_status;
    if (tile == 0) {
        for (int64_t r = tid; r < D; r += PMCMC_NT) {
            const T s = a.s0 ? a.s0[r] : (T)0;
            sin_[r] = s;
            lb_store<T>(&a.lb_incl[r], ga[r] * s + gb[r]);
        }
        lb_publish_fence();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(&st[0], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    for (int64_t r = tid; r < D; r += PMCMC_NT) {
        lb_store<T>(&a.lb_agg_a[(size_t)tile * D + r], ga[r]);
        lb_store<T>(&a.lb_agg_b[(size_t)tile * D + r], gb[r]);
        cmp_a[r] = (T)1;
        cmp_b[r] = (T)0;
    }
    lb_publish_fence();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&st[tile], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int j = tile - 1;
    for (;;) {
        {
            const int idx = j - tid;
            int f = 0;
            if (idx >= 0) {
                unsigned spins = 0;
                f = __hip_atomic_load(&st[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                while (f == 0) {
                    if (++spins > PMCMC_LB_SPIN_CAP) {
                        atomicOr(a.lb_err, 1u);
                        f = 2;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                    f = __hip_atomic_load(&st[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            const unsigned long long m = __ballot(f == 2);
            if ((tid & 63) == 0) lb_w[tid >> 6] = m ? (tid + (int)__builtin_ctzll(m)) : -1;
        }
        __syncthreads();
        if (tid == 0) {
            int l = -1;  // nearest predecessor with an inclusive state
            for (int w = 0; w < PMCMC_NT / 64 && l < 0; ++w) l = lb_w[w];
            lb_k = l >= 0 ? j - l : -1;
            lb_lo = j - (PMCMC_NT - 1) < 0 ? 0 : j - (PMCMC_NT - 1);
        }
        __syncthreads();
        const int k = lb_k, lo = lb_lo;
        const int stop = k >= 0 ? k + 1 : lo;
        for (int64_t r = tid; r < D; r += PMCMC_NT) {
            T ca = cmp_a[r], cb = cmp_b[r];
#pragma unroll 8
            for (int jj = j; jj >= stop; --jj) {
                const T xa = lb_load<T>(&a.lb_agg_a[(size_t)jj * D + r]);
                const T xb = lb_load<T>(&a.lb_agg_b[(size_t)jj * D + r]);
                cb = ca * xb + cb;
                ca = ca * xa;
            }
            if (k >= 0) sin_[r] = ca * lb_load<T>(&a.lb_incl[(size_t)k * D + r]) + cb;
            cmp_a[r] = ca;
            cmp_b[r] = cb;
        }
        if (k >= 0) break;
        j = lo - 1;
        __syncthreads();  // lb_w / lb_k / lb_lo are rewritten
    }
    for (int64_t r = tid; r < D; r += PMCMC_NT) lb_store<T>(&a.lb_incl[(size_t)tile * D + r], ga[r] * sin_[r] + gb[r]);
    lb_publish_fence();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&st[tile], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The kernel body, for tile `tile` of the problem `a`.  a.iter < 0 marks a
// batched solve: the iteration number is then kept on the device
// (sync->last_iter + 1) and the probes do not change between iterations.
template <typename T, int MODE, int PASS>
__device__ __forceinline__ void tile_body(const TileArgs<T>& a, const int tile) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double s_red[2][PMCMC_NT / 64];

    const int tid = threadIdx.x;
    // Newton loop already converged: later queued iterations are no-ops.  The
    // flag is loaded here and tested once this tile's staging loads are in
    // flight too (one memory round trip for both).  It and the final-prefix
    // count below are written by earlier launches only (within a launch, the
    // last workgroup to arrive writes them, after every other one has read
    // them), so plain loads see their current values.
    const int skip_now = a.skip ? *gp(a.skip) : 0;
    // the gate of the inline FP64 decisions (below), loaded here with the
    // skip flag so that it arrives with the staging loads instead of one
    // memory round trip after them; every input is written by earlier
    // launches, except the sticky flag that tile 0 of this launch may set,
    // whose other inputs then open the gate for every tile anyway
    int g_rf = 0, g_li = 0;
    double g_lm = 0.0;
    if (MODE == MODE_GEMM && a.rtg != nullptr) {
        g_rf = *gp(a.rflag);
        g_lm = *gp(&a.sync->last_metric);
        g_li = *gp(&a.sync->last_iter);
    }
    const int64_t D = a.D;
    const int TC = a.TC;
    const int64_t c0 = (int64_t)tile * TC;
    const int tc = (int)((a.T_len - c0) < TC ? (a.T_len - c0) : TC);
    const int P = a.P;
    const int64_t PA = D > PMCMC_NT ? D : PMCMC_NT;
    constexpr bool DEERM = MODE != MODE_SCAN;
    constexpr bool OUT = PASS != PASS_REDUCE;

    // End of a DEER output pass: the tile's maxima, its cached max |S| and
    // residual report (live tiles), the iteration's maxima; the last
    // workgroup forms the metric, the convergence flag and the next floor.
    auto finish = [&](double dmax, double smax, bool nz, bool final_tile) {
        SyncState* sy = a.sync;
        dmax = wave_max(dmax);
        smax = wave_max(smax);
        const int any_nz = __syncthreads_or(nz ? 1 : 0);
        if ((tid & 63) == 0) {
            s_red[0][tid >> 6] = dmax;
            s_red[1][tid >> 6] = smax;
        }
        __syncthreads();
        if (tid == 0) {
            double m0 = 0, m1 = 0;
            for (int w = 0; w < PMCMC_NT / 64; ++w) {
                m0 = nan_max(m0, s_red[0][w]);
                m1 = nan_max(m1, s_red[1][w]);
            }
            int live_col = -1;
            if (a.final_cols && !final_tile) {
                a.tile_smax[tile] = m1;
                if (any_nz) live_col = (int)c0;
            }
            // the last workgroup closes the iteration: metric and convergence flag
            double dm, sm;
            if (sync_arrive(sy, (unsigned)a.numTiles, m0, m1, live_col, &dm, &sm)) {
                const double metric = dm / (a.tol_abs + a.tol_rel * sm);
                const int it = a.iter >= 0 ? a.iter : sy->last_iter + 1;
                sy->last_metric = metric;
                sy->last_iter = it;
                const bool dec_final =
                    !a.refine_on || __hip_atomic_load(a.refine_on, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (!a.win && metric <= 1.0 && dec_final) {
                    sy->conv_iter = it;
                    sy->converged = 1;
                }
                if (a.win) {
                    // sliding window: the front moves past the final prefix (tiles
                    // before the first one with a nonzero residual), or past the
                    // whole window when it meets the tolerance; the window stays
                    // where it is while decisions may still change with the gate
                    SlideState* w = a.win;
                    const int lf = sync_take_live_from(sy);
                    // this iteration's window kernel initialised the window's new
                    // columns and copied the previous acceptance: both are done
                    if (w->front + a.T_len > w->hw) w->hw = w->front + a.T_len;
                    w->pend_lo = w->pend_hi = 0;
                    int64_t adv = 0;
                    if (dec_final) {
                        if (metric <= 1.0 || lf == INT_MAX)
                            adv = a.T_len;
                        else
                            adv = lf;
                    }
                    if (adv > 0) {
                        w->pend_lo = w->front;
                        w->pend_hi = w->front + adv;
                        w->front += adv;
                        w->last_adv_iter = it;
                        if (metric <= 1.0) {
                            w->accepts += 1;
                            w->seg_start = it + 1;
                        }
                    }
                    if (w->front >= w->T) {
                        sy->conv_iter = it;
                        sy->converged = 1;
                    }
                } else if (a.final_cols) {
                    // tiles before the first one with a nonzero residual are final,
                    // once decisions no longer change with the override's gate
                    const int lf = sync_take_live_from(sy);
                    if (dec_final) sy->skip_cols = lf == INT_MAX ? (int)((int64_t)a.numTiles * TC) : lf;
                }
            }
        }
    };
    const int final_now = (DEERM && a.final_cols) ? *gp(a.final_cols) : 0;

    T* xs = reinterpret_cast<T*>(smem);
    T* bl = xs + (DEERM ? (size_t)(TC + 1) * D : (size_t)0);
    T* al = bl + (size_t)TC * D;
    T* pa = al + (size_t)TC * D;
    T* pb = pa + PA;
    T* sin_ = pb + PA;
    uint8_t* accl = reinterpret_cast<uint8_t*>(sin_ + D);
    T* cmp_a = reinterpret_cast<T*>(smem + tile_lds_core<T>(D, TC, DEERM));
    T* cmp_b = cmp_a + D;

    const int64_t nel = (int64_t)tc * D;
    const int64_t gbase = c0 * D;

    // incoming state of an output pass's later tile: its first PMCMC_NT
    // entries are loaded ahead of the staging loads
    constexpr bool SIN = OUT && PASS != PASS_CHAINED && PASS != PASS_SINGLE;
    T sinv = (T)0;
    T fa[PMCMC_AGG_INLINE], fb[PMCMC_AGG_INLINE];
    const int rr = tid < D ? tid : 0;
    if (SIN && tile > 0) {
        if (PASS == PASS_APPLY && a.agg_inline) {
#pragma unroll
            for (int q = 0; q < PMCMC_AGG_INLINE; ++q) {
                const size_t k = q < tile ? (size_t)q : 0;
                fa[q] = gp(a.agg_a)[k * D + rr];
                fb[q] = gp(a.agg_b)[k * D + rr];
            }
        } else {
            sinv = gp(a.prev)[(size_t)(tile - 1) * D + rr];
        }
    }

    // ---- phase 0: stage inputs (coalesced, every region in one round) -------
    // A and B kept by the reduce pass of this iteration (MODE_GEMM multi-tile)
    const bool keep = DEERM && (PASS == PASS_REDUCE || PASS == PASS_APPLY) && a.a_keep != nullptr;
    const bool build = DEERM && !(PASS == PASS_APPLY && keep);
    if (!DEERM) {
        T* const dst[2] = {al, bl};
        const T* const src[2] = {a.A + gbase, a.B + gbase};
        const int64_t n[2] = {nel, nel};
        stage_lds_multi<T, 2>(dst, src, n, tid);
    } else {
        // the iterate's columns c0-1 .. c0+tc-1 (column c0-1 is s0 for the first tile)
        T* xd = c0 == 0 ? xs + D : xs;
        const T* xsrc = c0 == 0 ? a.S_in : a.S_in + gbase - D;
        const int64_t xn = c0 == 0 ? nel : nel + D;
        // s0 (first tile): its first PMCMC_NT entries are loaded ahead of
        // the staging loads and written after them
        T s0v = (T)0;
        if (c0 == 0) s0v = gp(a.s0)[tid < D ? tid : 0];
        if (build && MODE == MODE_GEMM) {
            // GEMM targets read their proposals Y from the pass: the tape's
            // noise is not needed (bl is written before it is read)
            T* const dst[1] = {xd};
            const T* const src[1] = {xsrc};
            const int64_t n[1] = {xn};
            stage_lds_multi<T, 1>(dst, src, n, tid);
        } else if (build) {
            T* const dst[2] = {xd, bl};
            const T* const src[2] = {xsrc, a.Xi + gbase};
            const int64_t n[2] = {xn, nel};
            stage_lds_multi<T, 2>(dst, src, n, tid);
        } else {
            T* const dst[3] = {xd, al, bl};
            const T* const src[3] = {xsrc, a.a_keep + gbase, a.b_keep + gbase};
            const int64_t n[3] = {xn, nel, nel};
            stage_lds_multi<T, 3>(dst, src, n, tid);
        }
        if (c0 == 0) {
            if (tid < D) xs[tid] = s0v;
            for (int64_t i = tid + PMCMC_NT; i < D; i += PMCMC_NT) xs[i] = gp(a.s0)[i];
        }
    }
    if (skip_now) return;
    if constexpr (DEERM && (PASS == PASS_REDUCE || PASS == PASS_APPLY)) {
        if (a.final_cols && c0 + TC <= (int64_t)final_now) {
            if constexpr (PASS == PASS_REDUCE) {
                for (int64_t r = tid; r < D; r += PMCMC_NT) {
                    a.agg_a[(size_t)tile * D + r] = (T)1;
                    a.agg_b[(size_t)tile * D + r] = (T)0;
                }
            } else {
                finish(0.0, tid == 0 ? a.tile_smax[tile] : 0.0, false, true);
            }
            return;
        }
    }

    if (SIN && PASS == PASS_APPLY && tile > 0 && a.agg_inline) {
        T f = (T)0;
#pragma unroll
        for (int q = 0; q < PMCMC_AGG_INLINE; ++q)
            if (q < tile) f = fa[q] * f + fb[q];
        sinv = f;
    }
    // incoming state of the tile (needed only by the output passes)
    if (OUT && PASS != PASS_CHAINED) {
        for (int64_t r = tid; r < D; r += PMCMC_NT)
            sin_[r] = (PASS == PASS_SINGLE || tile == 0) ? ((DEERM || !a.s0) ? (T)0 : gp(a.s0)[r])
                      : (SIN && r == tid)                ? sinv
                                                         : gp(a.prev)[(size_t)(tile - 1) * D + r];
    }
    // per-column outputs: by the pass that writes the tile, or by the reduce
    // pass when it keeps A and B
    const bool side = keep ? PASS == PASS_REDUCE : OUT;
    // inline FP64 decisions (GEMM targets with a refinement target): the gate
    // is read once, fixed for this launch, and the list of near-threshold
    // columns emptied; the barrier below publishes both
    const bool inl = MODE == MODE_GEMM && build && a.rtg != nullptr && side;
    bool inl_on = false;
    __shared__ int fl_n;
    if (inl) {
        inl_on = (g_rf != 0) | (g_lm <= a.gate) | (g_li + 1 >= a.gate_iters);
        if (tid == 0) {
            fl_n = 0;
            if (inl_on && tile == 0) __hip_atomic_store(a.rflag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();

    // ---- phase 1: build A, B per column (MALA step + surrogate JVP) -------
    if (build) {
        const int K = a.probes;
        // GEMM targets with one probe and D <= 1024: a group of G lanes per
        // column with up to 16 rows per lane (G = D / 16 rounded up to a power
        // of two), so each group has one or two columns of the tile and the
        // column's reductions take log2(G) <= 6 shuffle steps.  Otherwise one
        // row per lane (G = D rounded up, at most 64).
        constexpr int RM = PMCMC_FAST_RM;  // rows per lane of the fast path
        const bool fast = MODE == MODE_GEMM && K == 1 && D <= RM * 64;
        int G = 1;
        if (fast) {
            while (G * RM < D) G <<= 1;
            // wider groups while the groups still cover the tile's columns in
            // one round: fewer rows per lane on the chain of dependent FP64 work
            while (G < 64 && PMCMC_NT / (2 * G) >= tc) G <<= 1;
        } else {
            while (G < D && G < 64) G <<= 1;
        }
        const int lane = tid & (G - 1);
        const int grp = tid / G;
        const int ngrp = PMCMC_NT / G;
        const double eps = a.eps;
        const double sq2e = sqrt(2.0 * eps);
        const double c_norm = ((double)D / 2.0) * log(4.0 * 3.14159265358979323846 * eps);
        uint32_t keys[PMCMC_MAXP];
#pragma unroll
        for (int k = 0; k < PMCMC_MAXP; ++k) keys[k] = probe_key(a.seed, a.probe_iter, (uint32_t)k);
        const size_t DT = (size_t)D * (size_t)a.T_len;
        // inline FP64 decisions: the gate (fixed during this launch), and the
        // columns within the band (their FP32-built A and B are fixed up below)
        constexpr int FLMAX = 64;
        __shared__ int fl_c[FLMAX];
        __shared__ int8_t fl_pv[FLMAX];
        __shared__ double fl_dg[FLMAX];
        __shared__ double fl_v[256];
        // a column's log-density: the slab partials summed in slab order
        // (lp_nslab <= PMCMC_LP_MAXSLAB: the partials are loaded together,
        // clamped to slab 0; a loop with a runtime trip count would wait for
        // each load in turn)
        auto lp_load = [&](const double* lp, int64_t t, double* part) {
            const PMCMC_GLOBAL double* p = gp(lp) + t;
            const int ns = a.lp_nslab;
            const size_t st = (size_t)a.lp_stride;
#pragma unroll
            for (int sl = 0; sl < PMCMC_LP_MAXSLAB; ++sl) part[sl] = p[(sl < ns ? (size_t)sl : 0) * st];
        };
        auto lp_sum = [&](const double* part) -> double {
            double v = part[0];
#pragma unroll
            for (int sl = 1; sl < PMCMC_LP_MAXSLAB; ++sl) v = sl < a.lp_nslab ? v + part[sl] : v;
            return v;
        };
        auto lp_of = [&](const double* lp, int64_t t) -> double {
            double part[PMCMC_LP_MAXSLAB];
            lp_load(lp, t, part);
            return lp_sum(part);
        };
        auto load_ops = [&](int c, FastOps<T, RM>& f) {
            const int64_t t = c0 + (c < tc ? c : 0);
            // every operand load in flight at once: global address space,
            // rows beyond D clamped to a valid row (their values unused)
#pragma unroll
            for (int i = 0; i < RM; ++i) {
                if ((int64_t)i * G >= D) break;  // the same for every lane: rows past D unused
                const int64_t d = lane + (int64_t)i * G;
                const size_t gi = (size_t)t * D + (d < D ? d : D - 1);
                f.gx[i] = gp(a.gx)[gi];
                f.y[i] = gp(a.Y)[gi];
                f.gy[i] = gp(a.gy)[gi];
                f.w[i] = gp(a.W)[gi];
                f.hw[i] = gp(a.hw)[gi];
            }
            double px[PMCMC_LP_MAXSLAB], py[PMCMC_LP_MAXSLAB];
            lp_load(a.lpx, t, px);
            lp_load(a.lpy, t, py);
            f.u = gp(a.U)[t];
            // byte loads without a branch: a missing array reads the tape's U
            // (4 T bytes) instead and the value is replaced
            const int8_t o = gp(a.dec ? a.dec : (const int8_t*)a.U)[t];
            const int8_t pv = gp(a.prev_dec ? (const int8_t*)a.prev_dec : (const int8_t*)a.U)[t];
            f.o = a.dec ? o : (int8_t)-1;
            f.pv = a.prev_dec ? pv : (int8_t)-1;
            f.lpx = lp_sum(px);
            f.lpy = lp_sum(py);
        };

        // Accept decision of column c from its (group-summed) sums; returns it
        // and sets gg = g (1 - g) of the acceptance probability g and dg[k],
        // its derivative along probe k.  u: the tape's uniform; o: the FP64 override (-1 none);
        // pv: the previous decision (-1 none).
        auto decide = [&](int c, int64_t t, double s_lpx, double s_lpy, double s_qyx, double s_qxy,
                          const double* d1, const double* d2, const double* d3, double u, int8_t o, int8_t pv,
                          double* dg, double& gg) -> bool {
            // log acceptance ratio (MALA.jl:589-595), in double
            const double lq_yx = -0.5 * s_qyx / (2.0 * eps) - c_norm;
            const double lq_xy = -0.5 * s_qxy / (2.0 * eps) - c_norm;
            const double la = (s_lpy + lq_xy) - (s_lpx + lq_yx);
            const double lu = log(u);
            const double g = 1.0 / (1.0 + exp(-(la - lu)));
            bool acc = lu < la;
            bool tie = fabs(la - lu) < a.hyst;
            if (MODE == MODE_GEMM && o >= 0) {  // FP64 decision (bit 0) and its tie flag (bit 1)
                acc = (o & 1) != 0;
                tie = (o & 2) != 0;
            }
            // decisions an FP32 iterate cannot resolve keep their previous value
            // (every lane has read pv before lane 0 writes: one column per group)
            if (a.prev_dec && tie && pv >= 0) acc = pv != 0;
            gg = g * (1.0 - g);
#pragma unroll
            for (int k = 0; k < PMCMC_MAXP; ++k) dg[k] = gg * (d1[k] - d2[k] - d3[k] / (2.0 * eps));
            if (side && lane == 0) {
                accl[c] = acc ? 1 : 0;
                bool listed = false;
                if (inl_on && fabs(la - lu) < a.band) {
                    const int q = atomicAdd(&fl_n, 1);
                    if (q < FLMAX) {  // beyond the list: the FP32 decision stands
                        fl_c[q] = c;
                        fl_pv[q] = pv;
                        fl_dg[q] = dg[0];
                        listed = true;
                    }
                }
                if (!listed) {
                    if (a.prev_dec) gp(a.prev_dec)[t] = acc ? 1 : 0;
                    if (a.acc_out) gp(a.acc_out)[t] = acc ? 1 : 0;
                }
                if (a.la_out) gp(a.la_out)[t] = la;
            }
            return acc;
        };

        if (fast) {
            // the column's operands are loaded once, all in flight together,
            // and kept in registers for both sweeps
            auto column = [&](int c, const FastOps<T, RM>& f) {
                const int64_t t = c0 + c;
                const T* xcol = xs + (size_t)c * D;
                T* bcol = bl + (size_t)c * D;
                T* acol = al + (size_t)c * D;
                double s_qyx = 0, s_qxy = 0;
                double d1[PMCMC_MAXP] = {0, 0, 0, 0}, d2[PMCMC_MAXP] = {0, 0, 0, 0}, d3[PMCMC_MAXP] = {0, 0, 0, 0};
                static_assert(RM <= 32, "one sign bit per row of the lane");
                uint32_t zneg = 0;  // probe signs of the lane's rows (bit i: row i is -1), hashed once
#pragma unroll
                for (int i = 0; i < RM; ++i) {
                    const int64_t d = lane + (int64_t)i * G;
                    if (d < D) {
                        const double x = (double)xcol[d];
                        const size_t gi = (size_t)t * D + d;
                        const double gx = (double)f.gx[i], y = (double)f.y[i], gy = (double)f.gy[i];
                        const float zf = rademacher(keys[0], gi + (size_t)a.zoff);
                        zneg |= (zf < 0.f ? 1u : 0u) << i;
                        const double z = (double)zf;
                        const double w = (double)f.w[i], hy = (double)f.hw[i];
                        const double r = x - y - eps * gy;
                        const double dr = z - w - eps * hy;
                        d1[0] += gy * w;
                        d2[0] += gx * z;
                        d3[0] += r * dr;
                        const double e1 = y - x - eps * gx;
                        s_qyx += e1 * e1;
                        s_qxy += r * r;
                    }
                }
                s_qyx = group_sum(s_qyx, G);
                s_qxy = group_sum(s_qxy, G);
                d1[0] = group_sum(d1[0], G);
                d2[0] = group_sum(d2[0], G);
                d3[0] = group_sum(d3[0], G);
                double dg[PMCMC_MAXP], gg;
                const bool acc = decide(c, t, f.lpx, f.lpy, s_qyx, s_qxy, d1, d2, d3, (double)f.u, f.o, f.pv,
                                        dg, gg);
                // A and B in the iterate's precision (they are stored in it)
                const T dgt = (T)dg[0];
#pragma unroll
                for (int i = 0; i < RM; ++i) {
                    const int64_t d = lane + (int64_t)i * G;
                    if (d < D) {
                        const T x = xcol[d];
                        const T y = f.y[i];
                        if (MODE == MODE_GEMM && a.jacobian == JAC_DIAG) {
                            // exact diagonal: w holds c (the surrogate's log-alpha
                            // gradient), hw holds diag H_x
                            acol[d] = (T)((acc ? 1.0 + eps * (double)f.hw[i] : 1.0) +
                                          gg * ((double)y - (double)x) * (double)f.w[i]);
                        } else {
                            const T z = (zneg >> i) & 1u ? (T)-1 : (T)1;
                            const T jt = (acc ? f.w[i] : z) + (y - x) * dgt;
                            acol[d] = z * jt;
                        }
                        bcol[d] = (acc ? y : x) - corr_base(xcol[D + d]);  // FT_t - S_t
                    }
                }
            };
            for (int c = grp; c < tc; c += ngrp) {
                FastOps<T, RM> f;
                load_ops(c, f);
                column(c, f);
            }
        } else {
            for (int c = grp; c < tc; c += ngrp) {
                const int64_t t = c0 + c;
                const T* xcol = xs + (size_t)c * D;
                T* bcol = bl + (size_t)c * D;
                T* acol = al + (size_t)c * D;
                double s_lpx = 0, s_lpy = 0, s_qyx = 0, s_qxy = 0;
                double d1[PMCMC_MAXP], d2[PMCMC_MAXP], d3[PMCMC_MAXP];
#pragma unroll
                for (int k = 0; k < PMCMC_MAXP; ++k) d1[k] = d2[k] = d3[k] = 0;
                for (int64_t d = lane; d < D; d += G) {
                    const double x = (double)xcol[d];
                    const size_t gi = (size_t)t * D + d;
                    double gx, y, gy;
                    if (MODE == MODE_ELEM) {
                        const double p = a.prec ? (double)a.prec[d] : 1.0;
                        const double m = a.mean ? (double)a.mean[d] : 0.0;
                        gx = -p * (x - m);
                        y = x + eps * gx + sq2e * (double)bcol[d];
                        gy = -p * (y - m);
                        s_lpx += -0.5 * p * (x - m) * (x - m);
                        s_lpy += -0.5 * p * (y - m) * (y - m);
#pragma unroll
                        for (int k = 0; k < PMCMC_MAXP; ++k) {
                            if (k >= K) break;
                            const double z = (double)rademacher(keys[k], gi + (size_t)a.zoff);
                            const double w = z + eps * (-p * z);
                            const double hy = -p * w;
                            const double r = x - y - eps * gy;
                            const double dr = z - w - eps * hy;
                            d1[k] += gy * w;
                            d2[k] += gx * z;
                            d3[k] += r * dr;
                        }
                    } else {
                        gx = (double)a.gx[gi];
                        y = (double)a.Y[gi];
                        gy = (double)a.gy[gi];
#pragma unroll
                        for (int k = 0; k < PMCMC_MAXP; ++k) {
                            if (k >= K) break;
                            const double z = (double)rademacher(keys[k], gi + (size_t)a.zoff);
                            const double w = (double)a.W[k * DT + gi];
                            const double hy = (double)a.hw[k * DT + gi];
                            const double r = x - y - eps * gy;
                            const double dr = z - w - eps * hy;
                            d1[k] += gy * w;
                            d2[k] += gx * z;
                            d3[k] += r * dr;
                        }
                    }
                    const double e1 = y - x - eps * gx;
                    const double e2 = x - y - eps * gy;
                    s_qyx += e1 * e1;
                    s_qxy += e2 * e2;
                }
                s_qyx = group_sum(s_qyx, G);
                s_qxy = group_sum(s_qxy, G);
                if (MODE == MODE_ELEM) {
                    s_lpx = group_sum(s_lpx, G);
                    s_lpy = group_sum(s_lpy, G);
                } else {
                    s_lpx = lp_of(a.lpx, t);
                    s_lpy = lp_of(a.lpy, t);
                }
#pragma unroll
                for (int k = 0; k < PMCMC_MAXP; ++k) {
                    if (k >= K) break;
                    d1[k] = group_sum(d1[k], G);
                    d2[k] = group_sum(d2[k], G);
                    d3[k] = group_sum(d3[k], G);
                }
                const int8_t o = (MODE == MODE_GEMM && a.dec) ? a.dec[t] : (int8_t)-1;
                const int8_t pv = a.prev_dec ? a.prev_dec[t] : (int8_t)-1;
                double dg[PMCMC_MAXP], gg;
                const bool acc =
                    decide(c, t, s_lpx, s_lpy, s_qyx, s_qxy, d1, d2, d3, (double)a.U[t], o, pv, dg, gg);
                // second pass: A = mean_k z_k * Jt_k (or the exact diagonal), B = FT - A * x
                for (int64_t d = lane; d < D; d += G) {
                    const double x = (double)xcol[d];
                    const size_t gi = (size_t)t * D + d;
                    double gx, y, gy, p = 1.0;
                    if (MODE == MODE_ELEM) {
                        p = a.prec ? (double)a.prec[d] : 1.0;
                        const double m = a.mean ? (double)a.mean[d] : 0.0;
                        gx = -p * (x - m);
                        y = x + eps * gx + sq2e * (double)bcol[d];
                        gy = -p * (y - m);
                    } else {
                        gx = (double)a.gx[gi];
                        y = (double)a.Y[gi];
                        gy = (double)a.gy[gi];
                    }
                    const double ft = acc ? y : x;
                    double Av = 0;
                    if (a.jacobian == JAC_DIAG && MODE == MODE_ELEM) {
                        // exact diag of J = a(I + eps Hx) + (1-a)I + g(1-g)(y-x)c^T
                        const double hx = -p, hy = -p;
                        const double r = x - y - eps * gy;
                        const double cc = (1.0 + eps * hx) * gy - gx -
                                          (r - (1.0 + eps * hx) * (1.0 + eps * hy) * r) / (2.0 * eps);
                        Av = (acc ? (1.0 + eps * hx) : 1.0) + gg * (y - x) * cc;
                    } else if (a.jacobian == JAC_DIAG) {
                        // GEMM target: c and diag H_x precomputed (W, hw)
                        Av = (acc ? 1.0 + eps * (double)a.hw[gi] : 1.0) + gg * (y - x) * (double)a.W[gi];
                    } else {
#pragma unroll
                        for (int k = 0; k < PMCMC_MAXP; ++k) {
                            if (k >= K) break;
                            const double z = (double)rademacher(keys[k], gi + (size_t)a.zoff);
                            double w;
                            if (MODE == MODE_ELEM)
                                w = z + eps * (-p * z);
                            else
                                w = (double)a.W[k * DT + gi];
                            const double jt = (acc ? w : z) + (y - x) * dg[k];
                            Av += z * jt;
                        }
                        if (K > 1) Av *= 1.0 / (double)K;
                    }
                    acol[d] = (T)Av;
                    bcol[d] = (T)(ft - (double)corr_base(xcol[D + d]));  // FT_t - S_t
                }
            }
        }
        __syncthreads();
        if (inl) {
            // FP64 decisions of the listed columns (whole workgroup, one column
            // at a time); a changed decision rebuilds the column's A and B
            const int nfl = fl_n < FLMAX ? fl_n : FLMAX;
            for (int q = 0; q < nfl; ++q) {
                const int c = fl_c[q];
                const int64_t t = c0 + c;
                const T* xcol = xs + (size_t)c * D;
                const double la64 = gauss_logalpha_fp64<T>(*a.rtg, xcol, a.Xi + (size_t)t * D, eps, fl_v, s_red[0]);
                const double lu = log((double)a.U[t]);
                bool acc = lu < la64;
                const int8_t pv = fl_pv[q];
                if (a.prev_dec && fabs(la64 - lu) < a.hyst && pv >= 0) acc = pv != 0;
                if (acc != (accl[c] != 0)) {
                    const T dgt = (T)fl_dg[q];
                    for (int64_t d = tid; d < D; d += PMCMC_NT) {
                        const size_t gi = (size_t)t * D + d;
                        const T x = xcol[d];
                        const T y = a.Y[gi];
                        const T z = (T)rademacher(keys[0], gi + (size_t)a.zoff);
                        const T jt = (acc ? a.W[gi] : z) + (y - x) * dgt;
                        al[(size_t)c * D + d] = z * jt;
                        bl[(size_t)c * D + d] = (acc ? y : x) - corr_base(xcol[D + d]);
                    }
                }
                if (tid == 0) {
                    accl[c] = acc ? 1 : 0;
                    if (a.prev_dec) a.prev_dec[t] = acc ? 1 : 0;
                    if (a.acc_out) a.acc_out[t] = acc ? 1 : 0;
                }
                __syncthreads();
            }
        }
        if (PASS == PASS_REDUCE && keep) {
            for (int64_t i = tid; i < nel; i += PMCMC_NT) {
                gp(a.a_keep)[gbase + i] = al[i];
                gp(a.b_keep)[gbase + i] = bl[i];
            }
        }
    }

    // ---- phase 2: in-tile scan (rows x P parts, then Hillis-Steele over P)
    const int chunk = (tc + P - 1) / P;
    const int64_t nitems = D * P;
    for (int64_t i = tid; i < nitems; i += PMCMC_NT) {
        const int64_t r = i % D;
        const int p = (int)(i / D);
        const int cb = p * chunk;
        const int ce = (cb + chunk) < tc ? (cb + chunk) : tc;
        T al_ = (T)1, be_ = (T)0;
        for (int c = cb; c < ce; ++c) {
            const T av = al[(size_t)c * D + r];
            be_ = av * be_ + bl[(size_t)c * D + r];
            al_ = av * al_;
        }
        pa[p * D + r] = al_;
        pb[p * D + r] = be_;
    }
    __syncthreads();
    for (int off = 1; off < P; off <<= 1) {
        T na = 0, nb = 0;
        const bool act = tid < nitems;
        int64_t r = 0;
        int p = 0;
        if (act) {
            r = tid % D;
            p = (int)(tid / D);
            if (p >= off) {
                const T a1 = pa[p * D + r], b1 = pb[p * D + r];
                const T a0 = pa[(p - off) * D + r], b0 = pb[(p - off) * D + r];
                na = a1 * a0;
                nb = a1 * b0 + b1;
            }
        }
        __syncthreads();
        if (act && p >= off) {
            pa[p * D + r] = na;
            pb[p * D + r] = nb;
        }
        __syncthreads();
    }

    if (PASS == PASS_REDUCE) {
        // tile aggregate of row r: (pa[(P-1)D+r], pb[(P-1)D+r]), column `tile`
        for (int64_t r = tid; r < D; r += PMCMC_NT) {
            gp(a.agg_a)[(size_t)tile * D + r] = pa[(P - 1) * D + r];
            gp(a.agg_b)[(size_t)tile * D + r] = pb[(P - 1) * D + r];
        }
        return;
    }
    if (PASS == PASS_CHAINED) {
        chained_lookback<T>(a, tile, D, pa + (size_t)(P - 1) * D, pb + (size_t)(P - 1) * D, sin_, cmp_a, cmp_b);
        __syncthreads();
    }

    // ---- phase 4: apply the incoming state, damping, convergence maxima ---
    double dmax = 0, smax = 0;
    bool nz = false;  // a nonzero residual (B) in this tile
    const T damp = (T)a.damping;
    for (int64_t i = tid; i < nitems; i += PMCMC_NT) {
        const int64_t r = i % D;
        const int p = (int)(i / D);
        const int cb = p * chunk;
        const int ce = (cb + chunk) < tc ? (cb + chunk) : tc;
        T s = sin_[r];
        if (p > 0) s = pa[(p - 1) * D + r] * s + pb[(p - 1) * D + r];
        for (int c = cb; c < ce; ++c) {
            const size_t li = (size_t)c * D + r;
            s = al[li] * s + bl[li];
            if (!DEERM) {
                al[li] = s;
            } else {
                const T old = xs[li + D];
                const T base = corr_base(old);
                const T v = (a.damping == 1.0) ? base + s : base + damp * s;
                nz = nz || bl[li] != (T)0;  // NaN counts as nonzero
                al[li] = v;
                dmax = nan_max(dmax, fabs((double)v - (double)old));
                smax = nan_max(smax, fabs((double)v));
            }
        }
    }
    __syncthreads();

    // ---- phase 5: store the tile (coalesced) ------------------------------
    for (int64_t i = tid; i < nel; i += PMCMC_NT) gp(a.S)[gbase + i] = al[i];

    if (DEERM) finish(dmax, smax, nz, false);
}

template <typename T, int MODE, int PASS>
__global__ __launch_bounds__(PMCMC_NT, sizeof(T) == 4 ? 3 : 2) void tile_kernel(TileArgs<T> a) {
    tile_body<T, MODE, PASS>(a, blockIdx.x);
}

// Batched launch over many problems: work item blockIdx.x is (problem, tile) =
// items[blockIdx.x]; args[problem] describes the problem.
template <typename T, int MODE, int PASS>
__global__ __launch_bounds__(PMCMC_NT, PMCMC_BATCHED_WAVES) void tile_kernel_batched(const TileArgs<T>* __restrict__ args, const int2* __restrict__ items) {
    const int2 w = items[blockIdx.x];
    if (w.y >= args[w.x].numTiles) return;  // beyond the chain's current window
    tile_body<T, MODE, PASS>(args[w.x], w.y);
}

// Single pass with decoupled look-back: tiles are numbered by a ticket in the
// order workgroups start.
template <typename T, int MODE>
__global__ __launch_bounds__(PMCMC_NT) void tile_kernel_chained(TileArgs<T> a) {
    __shared__ int ticket;
    if (threadIdx.x == 0) ticket = (int)atomicAdd(a.lb_ticket, 1u);
    __syncthreads();
    tile_body<T, MODE, PASS_CHAINED>(a, ticket);
}

}  // namespace pmcmc
