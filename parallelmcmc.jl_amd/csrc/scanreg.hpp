// Single-pass diagonal scan for D = 64 with register-resident tiles.
//
// The two-pass scan reads A and B twice (reduce, then apply), so it moves
// 20 D T bytes for the 12 D T the scan needs; the LDS-tiled look-back scan
// reads them once but its tiles are small (LDS), so the look-back front has
// thousands of tiles to cross.  Here a workgroup's tile is 4 CPW columns
// held in registers: wave w owns columns [w CPW, (w+1) CPW) of the tile and
// lane = row, so every load and store is one coalesced 256-byte column.
//   1. load the tile (A and B, once), each lane composing its row's map over
//      its CPW columns; the four waves' maps compose into the tile's map
//   2. two-level decoupled look-back on tickets (tiles run in ticket order).
//      Tiles are grouped into super-tiles of SUPER consecutive tiles.  A tile
//      publishes its map, composes the maps of the tiles before it in its
//      super-tile (at most SUPER - 1), and takes the state at the end of the
//      previous super-tile from a look-back over super-tiles (their maps,
//      back to the nearest published end state, or s0).  The last tile of a
//      super-tile publishes the super-tile's map, then its end state.  With
//      one level, a tile walks back to the nearest published state over about
//      half the resident tiles' maps; with two, over fewer than SUPER tile
//      maps plus resident / SUPER super-tile maps.  Walks are split into four
//      contiguous chunks (one per wave) composed in order.
//   3. apply from registers: S written once
// Publication and loads are the ones of the LDS-tiled form (lb_store /
// lb_load / lb_publish_fence); every wait is a bounded spin reported in lb_err.
#pragma once

#include "tile_kernel.hpp"

namespace pmcmc {

static constexpr int SCAN_SUPER = 64;  // tiles per super-tile

// Wait until st[idx] >= lvl (bounded); returns the value seen.
__device__ __forceinline__ int lb_wait(int* st, int idx, int lvl, unsigned* err) {
    unsigned spins = 0;
    int f = __hip_atomic_load(&st[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (f < lvl) {
        if (++spins > PMCMC_LB_SPIN_CAP) {
            atomicOr(err, 1u);
            return 2;
        }
        __builtin_amdgcn_s_sleep(2);
        f = __hip_atomic_load(&st[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return f;
}

// Map of entries lo..hi of (ma, mb) (D x n, row `lane`), composed so that
// entry hi is applied last, split over the four waves in contiguous chunks
// (wave 0 nearest hi); the result is in (ra, rb) of row `lane`, all waves.
// Needs a barrier before (entries published) and uses pa / pb.
template <typename T>
__device__ __forceinline__ void lb_compose_range(const T* ma, const T* mb, int lo, int hi, int w, int lane, T (*pa)[64],
                                        T (*pb)[64], T& ra, T& rb) {
    constexpr int D = 64;
    const int n = hi - lo + 1;
    const int chunk = n > 0 ? (n + 3) / 4 : 0;
    const int hq = hi - w * chunk;
    const int lq = (hq - chunk + 1) > lo ? (hq - chunk + 1) : lo;
    T xa = 1, xb = 0;
    // 8 loads in flight: deeper unrolling costs registers, hence resident
    // tiles, which is what keeps HBM busy
#pragma unroll 8
    for (int jj = hq; jj >= lq; --jj) {
        xb = xa * lb_load<T>(&mb[(size_t)jj * D + lane]) + xb;
        xa = xa * lb_load<T>(&ma[(size_t)jj * D + lane]);
    }
    pa[w][lane] = xa;
    pb[w][lane] = xb;
    __syncthreads();
    T ca = 1, cb = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        cb = ca * pb[q][lane] + cb;
        ca = ca * pa[q][lane];
    }
    ra = ca;
    rb = cb;
}

// Workgroup state of the register-tile scans (D = 64).
template <typename T>
struct RegScanLds {
    T wa[2][4][64], wb[2][4][64];  // per-wave maps of the tile (two tiles in flight: parity)
    T pa[4][64], pb[4][64];        // look-back chunk maps
    T sin_[2][64];                 // incoming state of the tile (parity)
    int tk, lb_k, lb_lo;
    int lb_w[4];
};

// Issue the loads of tile `tile` (wave w: columns w CPW .. w CPW + CPW - 1 of
// the tile, lane = row); columns past T read as the identity map.
template <typename T, int CPW>
__device__ inline void reg_load(const TileArgs<T>& a, int tile, int w, int lane, T (&av)[CPW], T (&bv)[CPW]) {
    constexpr int D = 64;
    const int64_t t0 = (int64_t)tile * (4 * CPW) + (int64_t)w * CPW;
#pragma unroll
    for (int i = 0; i < CPW; ++i) {
        const int64_t t = t0 + i;
        const bool in = t < a.T_len;
        av[i] = in ? a.A[t * D + lane] : (T)1;
        bv[i] = in ? a.B[t * D + lane] : (T)0;
    }
}

// The tile's map of row `lane` (ga, gb): each wave composes its columns, the
// four wave maps compose in order.  Leaves the wave maps in L.wa / L.wb.
template <typename T, int CPW>
__device__ inline void reg_map(RegScanLds<T>& L, int par, int w, int lane, const T (&av)[CPW], const T (&bv)[CPW],
                               T& ga, T& gb) {
    T al = 1, be = 0;
#pragma unroll
    for (int i = 0; i < CPW; ++i) {
        be = av[i] * be + bv[i];
        al = av[i] * al;
    }
    L.wa[par][w][lane] = al;
    L.wb[par][w][lane] = be;
    __syncthreads();
    ga = 1;
    gb = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        gb = L.wa[par][q][lane] * gb + L.wb[par][q][lane];
        ga = L.wa[par][q][lane] * ga;
    }
}

// Publish tile `tile`'s map (ga, gb) of row `lane` (status 1).
template <typename T>
__device__ inline void reg_publish(const TileArgs<T>& a, int tile, T ga, T gb) {
    constexpr int D = 64;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (w == 0) {
        lb_store<T>(&a.lb_agg_a[(size_t)tile * D + lane], ga);
        lb_store<T>(&a.lb_agg_b[(size_t)tile * D + lane], gb);
    }
    lb_publish_fence();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&a.lb_status[tile], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Two-level look-back of tile `tile` (map (ga, gb), already published); on
// return (after a barrier) L.sin_[par] holds the tile's incoming state.
template <typename T>
__device__ __forceinline__ void reg_lookback(const TileArgs<T>& a, RegScanLds<T>& L, int par, int tile, T ga,
                                             T gb) {
    constexpr int D = 64;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int sup = tile / SCAN_SUPER, pos = tile % SCAN_SUPER;
    const int s_first = sup * SCAN_SUPER;
    int* st = a.lb_status;
    int* sst = a.lb_sstatus;
    // maps of the tiles before this one in its super-tile (earlier tickets)
    if (tid < pos) lb_wait(st, tile - 1 - tid, 1, a.lb_err);
    __syncthreads();
    T ma = 1, mb = 0;
    if (pos > 0) lb_compose_range<T>(a.lb_agg_a, a.lb_agg_b, s_first, tile - 1, w, lane, L.pa, L.pb, ma, mb);
    const bool last = pos == SCAN_SUPER - 1;
    if (last) {  // the super-tile's map
        if (w == 0) {
            lb_store<T>(&a.lb_sagg_a[(size_t)sup * D + lane], ga * ma);
            lb_store<T>(&a.lb_sagg_b[(size_t)sup * D + lane], ga * mb + gb);
        }
        lb_publish_fence();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(&sst[sup], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // state at the end of super-tile sup - 1: look back over super-tiles in
    // windows of 256 to the nearest published end state (index -1: s0)
    T ca = 1, cb = 0;  // map of the super-tiles walked so far
    T base = 0;
    int j = sup - 1;
    for (;;) {
        {
            const int idx = j - tid;
            int f = 0;
            if (idx == -1)
                f = 2;
            else if (idx >= 0)
                f = lb_wait(sst, idx, 1, a.lb_err);
            const unsigned long long m = __ballot(f == 2);
            if (lane == 0) L.lb_w[w] = m ? (tid + (int)__builtin_ctzll(m)) : -1;
        }
        __syncthreads();
        if (tid == 0) {
            int l = -1;  // nearest super-tile with an end state
            for (int q = 0; q < 4 && l < 0; ++q) l = L.lb_w[q];
            L.lb_k = l >= 0 ? j - l : -2;
            L.lb_lo = j - 255;
        }
        __syncthreads();
        const int k = L.lb_k, lo = L.lb_lo;
        const int stop = k >= -1 ? k + 1 : lo;
        T xa, xb;
        lb_compose_range<T>(a.lb_sagg_a, a.lb_sagg_b, stop, j, w, lane, L.pa, L.pb, xa, xb);
        cb = ca * xb + cb;
        ca = ca * xa;
        if (k >= -1) {
            base = k >= 0 ? lb_load<T>(&a.lb_incl[(size_t)k * D + lane]) : (a.s0 ? a.s0[lane] : (T)0);
            break;
        }
        j = lo - 1;
        __syncthreads();  // lb_w, lb_k, lb_lo, pa, pb are rewritten
    }
    const T sprev = ca * base + cb;   // state at the end of super-tile sup - 1
    const T sin_t = ma * sprev + mb;  // incoming state of this tile
    if (w == 0) L.sin_[par][lane] = sin_t;
    if (last) {
        if (w == 0) lb_store<T>(&a.lb_incl[(size_t)sup * D + lane], ga * sin_t + gb);
        lb_publish_fence();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(&sst[sup], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();  // L.sin_[par] is complete
}

// Apply the tile from registers: S written once.
template <typename T, int CPW>
__device__ inline void reg_apply(const TileArgs<T>& a, const RegScanLds<T>& L, int par, int tile, int w, int lane,
                                 const T (&av)[CPW], const T (&bv)[CPW]) {
    constexpr int D = 64;
    T s = L.sin_[par][lane];
    for (int q = 0; q < w; ++q) s = L.wa[par][q][lane] * s + L.wb[par][q][lane];
    const int64_t t0 = (int64_t)tile * (4 * CPW) + (int64_t)w * CPW;
#pragma unroll
    for (int i = 0; i < CPW; ++i) {
        s = av[i] * s + bv[i];
        const int64_t t = t0 + i;
        if (t < a.T_len) a.S[t * D + lane] = s;
    }
}

// One tile per workgroup.
template <typename T, int CPW>
__global__ __launch_bounds__(256) void scan_reg64_kernel(TileArgs<T> a) {
    __shared__ RegScanLds<T> L;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) L.tk = (int)atomicAdd(a.lb_ticket, 1u);
    __syncthreads();
    const int tile = L.tk;
    T av[CPW], bv[CPW];
    reg_load<T, CPW>(a, tile, w, lane, av, bv);
    T ga, gb;
    reg_map<T, CPW>(L, 0, w, lane, av, bv, ga, gb);
    reg_publish<T>(a, tile, ga, gb);
    reg_lookback<T>(a, L, 0, tile, ga, gb);
    reg_apply<T, CPW>(a, L, 0, tile, w, lane, av, bv);
}

// Persistent form: a workgroup keeps taking tickets, two tiles in flight.
// Step i (parity p): map and publish tile i (registers X); apply and store
// tile i - 1 (registers Y, its look-back done in the previous step); load
// the next ticket's tile into Y; look back for tile i.  So HBM streams the
// next tiles while the resident tiles wait on their predecessors (with one
// tile per workgroup, the resident tiles load together, then all wait, then
// all store), and publishing tile i (which waits for the wave's outstanding
// memory operations) does not wait for stores just issued.  Progress: the
// lowest unfinished tile is always being worked on by its holder (a
// workgroup holds a next ticket only while working on a lower tile), and
// everything it waits on belongs to lower tiles that have published.
template <typename T, int CPW>
__device__ inline void reg_persist_step(const TileArgs<T>& a, RegScanLds<T>& L, int par, int tile, int prev, int w,
                                        int lane, const T (&xa)[CPW], const T (&xb)[CPW], T (&ya)[CPW],
                                        T (&yb)[CPW], int& next) {
    __syncthreads();  // L.tk has been read; the previous step is done with L.wa[par]
    if (threadIdx.x == 0) L.tk = (int)atomicAdd(a.lb_ticket, 1u);
    T ga, gb;
    reg_map<T, CPW>(L, par, w, lane, xa, xb, ga, gb);  // its barrier also publishes L.tk
    next = L.tk;
    reg_publish<T>(a, tile, ga, gb);
    if (prev >= 0) reg_apply<T, CPW>(a, L, par ^ 1, prev, w, lane, ya, yb);
    if (next < a.numTiles) reg_load<T, CPW>(a, next, w, lane, ya, yb);
    reg_lookback<T>(a, L, par, tile, ga, gb);
}

template <typename T, int CPW>
__global__ __launch_bounds__(256) void scan_reg64_persist_kernel(TileArgs<T> a) {
    __shared__ RegScanLds<T> L;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) L.tk = (int)atomicAdd(a.lb_ticket, 1u);
    __syncthreads();
    int tile = L.tk;
    if (tile >= a.numTiles) return;
    T a0[CPW], b0[CPW], a1[CPW], b1[CPW];
    reg_load<T, CPW>(a, tile, w, lane, a0, b0);
    int prev = -1;
    for (;;) {
        int next;
        reg_persist_step<T, CPW>(a, L, 0, tile, prev, w, lane, a0, b0, a1, b1, next);
        if (next >= a.numTiles) {
            reg_apply<T, CPW>(a, L, 0, tile, w, lane, a0, b0);
            break;
        }
        prev = tile;
        tile = next;
        reg_persist_step<T, CPW>(a, L, 1, tile, prev, w, lane, a1, b1, a0, b0, next);
        if (next >= a.numTiles) {
            reg_apply<T, CPW>(a, L, 1, tile, w, lane, a1, b1);
            break;
        }
        prev = tile;
        tile = next;
    }
}

}  // namespace pmcmc
