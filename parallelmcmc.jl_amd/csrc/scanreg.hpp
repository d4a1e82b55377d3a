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
//   2. decoupled look-back on tickets (tiles run in ticket order): publish the
//      tile map, walk back to the nearest inclusive state, the walk split into
//      four contiguous chunks (one per wave) composed in order
//   3. apply from registers: S written once
// Look-back publication and loads are the ones of the LDS-tiled form
// (lb_store / lb_load / lb_publish_fence, bounded spins reported in lb_err).
#pragma once

#include "tile_kernel.hpp"

namespace pmcmc {

template <typename T, int CPW>
__global__ __launch_bounds__(256) void scan_reg64_kernel(TileArgs<T> a) {
    constexpr int D = 64, TC = 4 * CPW;
    __shared__ T wa[4][D], wb[4][D];  // per-wave maps of the tile
    __shared__ T pa[4][D], pb[4][D];  // look-back chunk maps
    __shared__ T sin_[D];
    __shared__ int tk, lb_k, lb_lo;
    __shared__ int lb_w[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) tk = (int)atomicAdd(a.lb_ticket, 1u);
    __syncthreads();
    const int tile = tk;
    const int64_t Tn = a.T_len;
    const int64_t t0 = (int64_t)tile * TC + (int64_t)w * CPW;
    T av[CPW], bv[CPW];
#pragma unroll
    for (int i = 0; i < CPW; ++i) {
        const int64_t t = t0 + i;
        const bool in = t < Tn;
        av[i] = in ? a.A[t * D + lane] : (T)1;
        bv[i] = in ? a.B[t * D + lane] : (T)0;
    }
    T al = 1, be = 0;
#pragma unroll
    for (int i = 0; i < CPW; ++i) {
        be = av[i] * be + bv[i];
        al = av[i] * al;
    }
    wa[w][lane] = al;
    wb[w][lane] = be;
    __syncthreads();
    T ga = 1, gb = 0;  // the tile's map of row `lane`
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        gb = wa[q][lane] * gb + wb[q][lane];
        ga = wa[q][lane] * ga;
    }
    int* st = a.lb_status;
    if (tile == 0) {
        if (w == 0) {
            const T s0 = a.s0 ? a.s0[lane] : (T)0;
            sin_[lane] = s0;
            lb_store<T>(&a.lb_incl[lane], ga * s0 + gb);
        }
        lb_publish_fence();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(&st[0], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        if (w == 0) {
            lb_store<T>(&a.lb_agg_a[(size_t)tile * D + lane], ga);
            lb_store<T>(&a.lb_agg_b[(size_t)tile * D + lane], gb);
        }
        lb_publish_fence();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(&st[tile], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        T ca = 1, cb = 0;  // map of the predecessors walked so far (wave 0)
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
                if (lane == 0) lb_w[w] = m ? (tid + (int)__builtin_ctzll(m)) : -1;
            }
            __syncthreads();
            if (tid == 0) {
                int l = -1;  // nearest predecessor with an inclusive state
                for (int q = 0; q < 4 && l < 0; ++q) l = lb_w[q];
                lb_k = l >= 0 ? j - l : -1;
                lb_lo = j - 255 < 0 ? 0 : j - 255;
            }
            __syncthreads();
            const int k = lb_k, lo = lb_lo;
            const int stop = k >= 0 ? k + 1 : lo;
            // maps of tiles j .. stop, in four contiguous chunks (chunk 0 nearest)
            const int n = j - stop + 1;
            const int chunk = (n + 3) / 4;
            const int hi_q = j - w * chunk;
            const int lo_q = (hi_q - chunk + 1) > stop ? (hi_q - chunk + 1) : stop;
            T xa = 1, xb = 0;
            // 8 loads in flight: deeper unrolling costs registers, hence
            // resident tiles, which is what keeps HBM busy (measured: 32 in
            // flight made the 128-column form 1.7x slower)
#pragma unroll 8
            for (int jj = hi_q; jj >= lo_q; --jj) {
                xb = xa * lb_load<T>(&a.lb_agg_b[(size_t)jj * D + lane]) + xb;
                xa = xa * lb_load<T>(&a.lb_agg_a[(size_t)jj * D + lane]);
            }
            pa[w][lane] = xa;
            pb[w][lane] = xb;
            __syncthreads();
            if (w == 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    cb = ca * pb[q][lane] + cb;
                    ca = ca * pa[q][lane];
                }
                if (k >= 0) sin_[lane] = ca * lb_load<T>(&a.lb_incl[(size_t)k * D + lane]) + cb;
            }
            if (k >= 0) break;
            j = lo - 1;
            __syncthreads();  // lb_w, lb_k, lb_lo, pa, pb are rewritten
        }
        __syncthreads();  // sin_ is complete
        if (w == 0) lb_store<T>(&a.lb_incl[(size_t)tile * D + lane], ga * sin_[lane] + gb);
        lb_publish_fence();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(&st[tile], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    T s = sin_[lane];
    for (int q = 0; q < w; ++q) s = wa[q][lane] * s + wb[q][lane];
#pragma unroll
    for (int i = 0; i < CPW; ++i) {
        s = av[i] * s + bv[i];
        const int64_t t = t0 + i;
        if (t < Tn) a.S[t * D + lane] = s;
    }
}

}  // namespace pmcmc
