// Single-pass diagonal scan with register-resident tiles, D up to 256.
//
// The two-pass scan reads A and B twice (reduce, then apply), so it moves
// 20 D T bytes for the 12 D T the scan needs; the LDS-tiled look-back scan
// reads them once but its tiles are small (LDS), so the look-back front has
// thousands of tiles to cross.  Here a workgroup's tile is 4 CPW columns of
// one row block held in registers: wave w owns columns [w CPW, (w+1) CPW) of
// the tile and lane = row of the block, so every load and store is one
// coalesced run of a column.  The rows of a diagonal scan never interact, so
// D > 64 is split into nb = ceil(D / 64) row blocks of ceil(D / nb) rows, each
// scanned as its own chain of tiles; ticket k is tile k / nb of block k mod nb,
// so the blocks advance together and a tile's look-back reads maps of one
// block (64 rows) whatever D is.
//   1. load the tile (A and B, once), each lane composing its row's map over
//      its CPW columns; the four waves' maps compose into the tile's map
//   2. two-level decoupled look-back on tickets (tiles run in ticket order).
//      Tiles are grouped into super-tiles of SUPER consecutive tiles of a
//      block.  A tile publishes its map, composes the maps of the tiles
//      before it in its super-tile (at most SUPER - 1), and takes the state
//      at the end of the previous super-tile from a look-back over
//      super-tiles (their maps, back to the nearest published end state, or
//      s0).  The last tile of a super-tile publishes the super-tile's map,
//      then its end state.  With one level, a tile walks back to the nearest
//      published state over about half the resident tiles' maps; with two,
//      over fewer than SUPER tile maps plus resident / SUPER super-tile
//      maps.  Walks are split into four contiguous chunks (one per wave)
//      composed in order.
//   3. apply from registers: S written once
// Publication and loads are the ones of the LDS-tiled form (lb_store /
// lb_load / lb_publish_fence); every wait is a bounded spin reported in lb_err.
// Look-back storage per block: tile maps and status of numTiles tiles,
// super-tile maps, end states and status of ceil(numTiles / SUPER)
// super-tiles, maps 64 rows wide.
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

// Map of entries lo..hi of (ma, mb) (64-row maps, row `lane`), composed so
// that entry hi is applied last, split over the four waves in contiguous
// chunks (wave 0 nearest hi); the result is in (ra, rb), all waves.  Needs a
// barrier before (entries published) and uses pa / pb.
template <typename T>
__device__ __forceinline__ void lb_compose_range(const T* ma, const T* mb, int lo, int hi, int w, int lane,
                                                 T (*pa)[64], T (*pb)[64], T& ra, T& rb) {
    const int n = hi - lo + 1;
    const int chunk = n > 0 ? (n + 3) / 4 : 0;
    const int hq = hi - w * chunk;
    const int lq = (hq - chunk + 1) > lo ? (hq - chunk + 1) : lo;
    T xa = 1, xb = 0;
    // 8 loads in flight: deeper unrolling costs registers, hence resident
    // tiles, which is what keeps HBM busy
#pragma unroll 8
    for (int jj = hq; jj >= lq; --jj) {
        xb = xa * lb_load<T>(&mb[(size_t)jj * 64 + lane]) + xb;
        xa = xa * lb_load<T>(&ma[(size_t)jj * 64 + lane]);
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

// Workgroup state of the register-tile scans.
template <typename T>
struct RegScanLds {
    T wa[2][4][64], wb[2][4][64];  // per-wave maps of the tile (two tiles in flight: parity)
    T pa[4][64], pb[4][64];        // look-back chunk maps
    T sin_[2][64];                 // incoming state of the tile (parity)
    int tk, lb_k, lb_lo;
    int lb_w[4];
};

// Tile of ticket k: block k mod nb, tile k / nb of that block; this lane's row
// d (valid if row_ok) and the block's look-back storage.
template <typename T>
struct RegTile {
    int i;       // tile index within the block
    int d;       // row of this lane
    bool row_ok;
    int* st;     // block's tile status, tile maps, super-tile status, maps, end states
    T *agg_a, *agg_b;
    int* sst;
    T *sagg_a, *sagg_b, *incl;
};

template <typename T>
__device__ inline RegTile<T> reg_tile(const TileArgs<T>& a, int k, int lane) {
    const int D = (int)a.D;
    const int nb = (D + 63) >> 6;
    const int rpb = (D + nb - 1) / nb;  // rows per block
    const int blk = k % nb;
    const int nS = (a.numTiles + SCAN_SUPER - 1) / SCAN_SUPER;
    RegTile<T> g;
    g.i = k / nb;
    g.d = blk * rpb + lane;
    g.row_ok = lane < rpb && g.d < D;
    g.st = a.lb_status + (size_t)blk * a.numTiles;
    g.agg_a = a.lb_agg_a + (size_t)blk * a.numTiles * 64;
    g.agg_b = a.lb_agg_b + (size_t)blk * a.numTiles * 64;
    g.sst = a.lb_sstatus + (size_t)blk * nS;
    g.sagg_a = a.lb_sagg_a + (size_t)blk * nS * 64;
    g.sagg_b = a.lb_sagg_b + (size_t)blk * nS * 64;
    g.incl = a.lb_incl + (size_t)blk * nS * 64;
    return g;
}

// Issue the loads of a tile (wave w: columns w CPW .. w CPW + CPW - 1 of the
// tile, lane: row g.d); columns past T and lanes past the block read as the
// identity map.
template <typename T, int CPW>
__device__ inline void reg_load(const TileArgs<T>& a, const RegTile<T>& g, int w, T (&av)[CPW], T (&bv)[CPW]) {
    const int64_t D = a.D;
    const int64_t t0 = (int64_t)g.i * (4 * CPW) + (int64_t)w * CPW;
#pragma unroll
    for (int i = 0; i < CPW; ++i) {
        const int64_t t = t0 + i;
        const bool in = g.row_ok && t < a.T_len;
        av[i] = in ? a.A[t * D + g.d] : (T)1;
        bv[i] = in ? a.B[t * D + g.d] : (T)0;
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

// Publish the tile's map (ga, gb) of row `lane` (status 1).
template <typename T>
__device__ inline void reg_publish(const RegTile<T>& g, T ga, T gb) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (w == 0) {
        lb_store<T>(&g.agg_a[(size_t)g.i * 64 + lane], ga);
        lb_store<T>(&g.agg_b[(size_t)g.i * 64 + lane], gb);
    }
    lb_publish_fence();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&g.st[g.i], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Two-level look-back of the tile (map (ga, gb), already published); on
// return (after a barrier) L.sin_[par] holds the tile's incoming state.
template <typename T>
__device__ __forceinline__ void reg_lookback(const TileArgs<T>& a, const RegTile<T>& g, RegScanLds<T>& L, int par,
                                             T ga, T gb) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int tile = g.i;
    const int sup = tile / SCAN_SUPER, pos = tile % SCAN_SUPER;
    const int s_first = sup * SCAN_SUPER;
    // maps of the tiles before this one in its super-tile (earlier tickets)
    if (tid < pos) lb_wait(g.st, tile - 1 - tid, 1, a.lb_err);
    __syncthreads();
    T ma = 1, mb = 0;
    if (pos > 0) lb_compose_range<T>(g.agg_a, g.agg_b, s_first, tile - 1, w, lane, L.pa, L.pb, ma, mb);
    const bool last = pos == SCAN_SUPER - 1;
    if (last) {  // the super-tile's map
        if (w == 0) {
            lb_store<T>(&g.sagg_a[(size_t)sup * 64 + lane], ga * ma);
            lb_store<T>(&g.sagg_b[(size_t)sup * 64 + lane], ga * mb + gb);
        }
        lb_publish_fence();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(&g.sst[sup], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
                f = lb_wait(g.sst, idx, 1, a.lb_err);
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
        lb_compose_range<T>(g.sagg_a, g.sagg_b, stop, j, w, lane, L.pa, L.pb, xa, xb);
        cb = ca * xb + cb;
        ca = ca * xa;
        if (k >= -1) {
            base = k >= 0 ? lb_load<T>(&g.incl[(size_t)k * 64 + lane]) : (a.s0 && g.row_ok ? a.s0[g.d] : (T)0);
            break;
        }
        j = lo - 1;
        __syncthreads();  // lb_w, lb_k, lb_lo, pa, pb are rewritten
    }
    const T sprev = ca * base + cb;   // state at the end of super-tile sup - 1
    const T sin_t = ma * sprev + mb;  // incoming state of this tile
    if (w == 0) L.sin_[par][lane] = sin_t;
    if (last) {
        if (w == 0) lb_store<T>(&g.incl[(size_t)sup * 64 + lane], ga * sin_t + gb);
        lb_publish_fence();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(&g.sst[sup], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();  // L.sin_[par] is complete
}

// Apply the tile from registers: S written once.
template <typename T, int CPW>
__device__ inline void reg_apply(const TileArgs<T>& a, const RegTile<T>& g, const RegScanLds<T>& L, int par, int w,
                                 int lane, const T (&av)[CPW], const T (&bv)[CPW]) {
    const int64_t D = a.D;
    T s = L.sin_[par][lane];
    for (int q = 0; q < w; ++q) s = L.wa[par][q][lane] * s + L.wb[par][q][lane];
    const int64_t t0 = (int64_t)g.i * (4 * CPW) + (int64_t)w * CPW;
#pragma unroll
    for (int i = 0; i < CPW; ++i) {
        s = av[i] * s + bv[i];
        const int64_t t = t0 + i;
        if (g.row_ok && t < a.T_len) a.S[t * D + g.d] = s;
    }
}

// Tickets of a launch: numTiles tiles of each of the row blocks.
template <typename T>
__device__ inline int reg_tickets(const TileArgs<T>& a) {
    return ((int)a.D + 63) / 64 * a.numTiles;
}

// One tile per workgroup.
template <typename T, int CPW>
__global__ __launch_bounds__(256) void scan_reg_kernel(TileArgs<T> a) {
    __shared__ RegScanLds<T> L;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) L.tk = (int)atomicAdd(a.lb_ticket, 1u);
    __syncthreads();
    const RegTile<T> g = reg_tile(a, L.tk, lane);
    T av[CPW], bv[CPW];
    reg_load<T, CPW>(a, g, w, av, bv);
    T ga, gb;
    reg_map<T, CPW>(L, 0, w, lane, av, bv, ga, gb);
    reg_publish<T>(g, ga, gb);
    reg_lookback<T>(a, g, L, 0, ga, gb);
    reg_apply<T, CPW>(a, g, L, 0, w, lane, av, bv);
}

// Persistent form: a workgroup keeps taking tickets, two tiles in flight.
// Step i (parity p): map and publish tile i (registers X); apply and store
// tile i - 1 (registers Y, its look-back done in the previous step); load
// the next ticket's tile into Y; look back for tile i.  So HBM streams the
// next tiles while the resident tiles wait on their predecessors (with one
// tile per workgroup, the resident tiles load together, then all wait, then
// all store), and publishing tile i (which waits for the wave's outstanding
// memory operations) does not wait for stores just issued.  Progress: the
// lowest unfinished ticket is always being worked on by its holder (a
// workgroup holds a next ticket only while working on a lower one), and
// everything it waits on belongs to lower tickets that have published.
template <typename T, int CPW>
__device__ inline void reg_persist_step(const TileArgs<T>& a, RegScanLds<T>& L, int par, const RegTile<T>& g,
                                        const RegTile<T>& gp, bool has_prev, int w, int lane, const T (&xa)[CPW],
                                        const T (&xb)[CPW], T (&ya)[CPW], T (&yb)[CPW], int& next) {
    __syncthreads();  // L.tk has been read; the previous step is done with L.wa[par]
    if (threadIdx.x == 0) L.tk = (int)atomicAdd(a.lb_ticket, 1u);
    T ga, gb;
    reg_map<T, CPW>(L, par, w, lane, xa, xb, ga, gb);  // its barrier also publishes L.tk
    next = L.tk;
    reg_publish<T>(g, ga, gb);
    if (has_prev) reg_apply<T, CPW>(a, gp, L, par ^ 1, w, lane, ya, yb);
    if (next < reg_tickets(a)) reg_load<T, CPW>(a, reg_tile(a, next, lane), w, ya, yb);
    reg_lookback<T>(a, g, L, par, ga, gb);
}

template <typename T, int CPW>
__global__ __launch_bounds__(256) void scan_reg_persist_kernel(TileArgs<T> a) {
    __shared__ RegScanLds<T> L;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) L.tk = (int)atomicAdd(a.lb_ticket, 1u);
    __syncthreads();
    const int ntk = reg_tickets(a);
    int k = L.tk;
    if (k >= ntk) return;
    T a0[CPW], b0[CPW], a1[CPW], b1[CPW];
    RegTile<T> g = reg_tile(a, k, lane), gp = g;
    reg_load<T, CPW>(a, g, w, a0, b0);
    bool has_prev = false;
    for (;;) {
        int next;
        reg_persist_step<T, CPW>(a, L, 0, g, gp, has_prev, w, lane, a0, b0, a1, b1, next);
        if (next >= ntk) {
            reg_apply<T, CPW>(a, g, L, 0, w, lane, a0, b0);
            break;
        }
        gp = g;
        has_prev = true;
        g = reg_tile(a, next, lane);
        reg_persist_step<T, CPW>(a, L, 1, g, gp, has_prev, w, lane, a1, b1, a0, b0, next);
        if (next >= ntk) {
            reg_apply<T, CPW>(a, g, L, 1, w, lane, a1, b1);
            break;
        }
        gp = g;
        g = reg_tile(a, next, lane);
    }
}

}  // namespace pmcmc
