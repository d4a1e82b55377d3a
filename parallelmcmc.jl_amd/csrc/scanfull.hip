// Dense-Jacobian affine scan with materialised Jacobians, D <= 64:
//   S[:,t] = J_t S[:,t-1] + B[:,t],  S[:,-1] = s0,
// J_t the t-th D x D column-major matrix (J + t*strideJ, leading dim ldJ).
// The reference has no such scan (it rejects :full, src/DEER/DEER.jl:464-465;
// its diagonal scan is src/DEER/DEERScan.jl:102-173); this is the general form
// of the structured full-Jacobian solve in fullscan.hip, for callers that
// hold their Jacobians.
//
//   reduce  one workgroup per tile of TC columns composes the tile's map
//           [M | r] <- J_t [M | r] + [0 | B_t] from [I | 0]: J_t M on the 16x16x4
//           MFMA (wave w: rows 16 w .. 16 w + 15, all 64 columns), J_t's A
//           fragments loaded from HBM straight into registers (J_{t+1}'s
//           while J_t is used), M in LDS (two buffers: one barrier per
//           column); J_t r from the same fragments on the VALU
//   spine   the tile maps form a dense scan of their own over the tiles
//           (J' = M_k, B' = r_k, from s0); it is solved by the same kernels
//           with 8-column tiles, recursively, down to one tile
//   apply   one workgroup per tile runs its columns from the incoming state
//           (four waves split the k sum of each mat-vec; the columns' J and B
//           stream through a ring of SAPF register slots)
// The structured full-Jacobian solve runs its scan over tile maps through
// scan_full_spine: up to 256 maps in one launch (sfull_spine_fused_kernel:
// 8-map groups, the chain over the groups in the last workgroup to arrive),
// beyond that the recursion above from level 1.
// Padded rows and columns (>= D) are zero in every map, so they stay out of
// every product.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/pmcmc_amd.h"
#include "mfma.hpp"
#include "scanfull.hpp"

namespace pmcmc {

static constexpr int SF = 64;       // padded dimension
static constexpr int SFS = SF + 4;  // LDS row stride (keeps 4-wide reads aligned)
static constexpr int SAPF = 4;      // apply prefetch depth

template <typename T>
struct SFArgs {
    const T* J;
    int64_t ldJ, strideJ;
    const T* B;
    int64_t ldB;
    const T* s0;    // D entries, or null for a zero start
    const T* prev;  // state after each tile of this level (ld SF), from the level above
    T* S;
    int64_t ldS;
    int64_t D, len;
    int TC, nT;
    T* Magg;  // per tile SF x SF, column-major (ld SF)
    T* ragg;  // per tile SF
    const int* skip;  // nonzero => no-op (a converged Newton loop); may be null
};

// LDS row stride of the map in the MFMA reduce: 80 floats puts the 64 lanes of
// a B-fragment read (16 columns x 4 rows) on 64 distinct banks; FP64 keeps
// the two buffers near 64 KB.
template <typename T>
struct SFStride {
    static constexpr int v = sizeof(T) == 4 ? 80 : 66;
};

// Tile `tile`'s map: [M | r] <- J_t [M | r] + [0 | B_t] over its columns,
// written to Magg / ragg.  Called by all 256 threads of the workgroup.
template <typename T>
__device__ __forceinline__ void sfull_reduce_tile(const SFArgs<T>& p, int64_t tile) {
    typedef typename Mfma<T>::acc_t acc_t;
    constexpr int MS = SFStride<T>::v;
    __shared__ __attribute__((aligned(16))) T Ml[2][SF][MS];  // Ml[b][k][j] = M(k, j)
    __shared__ T rl[2][SF];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t D = p.D;
    const int64_t t0 = tile * p.TC;
    const int tc = (int)std::min<int64_t>(p.TC, p.len - t0);
    for (int e = tid; e < SF * SF; e += 256) Ml[0][e >> 6][e & 63] = (e >> 6) == (e & 63) ? (T)1 : (T)0;
    if (tid < SF) rl[0][tid] = 0;
    // lane (li, lk) of wave w holds J_t(i = 16 w + li, k = 4 ks + lk), ks = 0..15,
    // its A fragments, and B_t(i)
    const int i = 16 * w + li;
    T ja[16], jn[16];
    T bc = 0, bn = 0;
    auto fetch = [&](int64_t t, T(&dst)[16], T& bdst) {
        const T* Jt = p.J + t * p.strideJ;
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) {
            const int k = 4 * ks + lk;
            dst[ks] = (i < D && k < D) ? Jt[i + k * p.ldJ] : (T)0;
        }
        bdst = i < D ? p.B[t * p.ldB + i] : (T)0;
    };
    fetch(t0, ja, bc);
    __syncthreads();
    for (int s = 0; s < tc; ++s) {
        const int cur = s & 1;
        if (s + 1 < tc) fetch(t0 + s + 1, jn, bn);
        acc_t acc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[ct] = (acc_t){0, 0, 0, 0};
        T rp = 0;
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) {
            const T a = ja[ks];
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[ct] = Mfma<T>::step(a, Ml[cur][4 * ks + lk][16 * ct + li], acc[ct]);
            rp += a * rl[cur][4 * ks + lk];
        }
        // r_new(i): this lane's k terms, then the four lanes of row i (lk = 0..3)
        rp += __shfl_xor(rp, 16, 64);
        rp += __shfl_xor(rp, 32, 64);
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) Ml[cur ^ 1][16 * w + Mfma<T>::row(lk, r)][16 * ct + li] = acc[ct][r];
        if (lk == 0) rl[cur ^ 1][i] = rp + bc;
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) ja[ks] = jn[ks];
        bc = bn;
    }
    const int fin = tc & 1;
    T* Mo = p.Magg + (size_t)tile * SF * SF;
    for (int e = tid; e < SF * SF; e += 256) Mo[e] = Ml[fin][e & 63][e >> 6];  // column-major: e = i + 64 j
    if (tid < SF) p.ragg[(size_t)tile * SF + tid] = rl[fin][tid];
}

template <typename T>
__global__ __launch_bounds__(256) void sfull_reduce_kernel(SFArgs<T> p) {
    if (p.skip && *(volatile const int*)p.skip) return;
    sfull_reduce_tile<T>(p, blockIdx.x);
}

// Tile `tile`'s columns from its incoming state (s0 for tile 0, else
// prev[tile - 1]), written to S.  Called by all 256 threads of the workgroup.
template <typename T>
__device__ __forceinline__ void sfull_apply_tile(const SFArgs<T>& p, int64_t tile) {
    __shared__ T v[SF];
    __shared__ T part[4][SF];
    const int tid = threadIdx.x, i = tid & 63, q = tid >> 6;
    const int64_t D = p.D;
    const int64_t t0 = tile * p.TC;
    const int tc = (int)std::min<int64_t>(p.TC, p.len - t0);
    if (tid < SF)
        v[i] = i >= D ? (T)0 : (tile == 0 ? (p.s0 ? p.s0[i] : (T)0) : p.prev[(size_t)(tile - 1) * SF + i]);
    T jr[SAPF][16];
    T br[SAPF];
    auto fetch = [&](int s, int sl) {
        const T* Jt = p.J + (t0 + s) * p.strideJ;
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            const int k = 16 * q + kk;
            jr[sl][kk] = (i < D && k < D) ? Jt[i + k * p.ldJ] : (T)0;
        }
        br[sl] = (q == 0 && i < D) ? p.B[(t0 + s) * p.ldB + i] : (T)0;
    };
#pragma unroll
    for (int sl = 0; sl < SAPF; ++sl)
        if (sl < tc) fetch(sl, sl);
    __syncthreads();
    for (int s0 = 0; s0 < tc; s0 += SAPF) {
#pragma unroll
        for (int sl = 0; sl < SAPF; ++sl) {
            const int s = s0 + sl;
            if (s < tc) {
                T a0 = 0, a1 = 0;
#pragma unroll
                for (int kk = 0; kk < 16; kk += 2) {
                    a0 += jr[sl][kk] * v[16 * q + kk];
                    a1 += jr[sl][kk + 1] * v[16 * q + kk + 1];
                }
                const T bc = br[sl];
                if (s + SAPF < tc) fetch(s + SAPF, sl);
                part[q][i] = a0 + a1;
                __syncthreads();
                if (q == 0) {
                    const T nv = ((part[0][i] + part[1][i]) + (part[2][i] + part[3][i])) + bc;
                    v[i] = nv;
                    if (i < D) p.S[(t0 + s) * p.ldS + i] = nv;
                }
                __syncthreads();
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void sfull_apply_kernel(SFArgs<T> p) {
    if (p.skip && *(volatile const int*)p.skip) return;
    sfull_apply_tile<T>(p, blockIdx.x);
}

// The scan over at most FUSE_MAX_GROUPS * 8 tile maps in one launch (the
// recursive form takes five dependent launches at 256 maps).  Workgroup g
// composes maps 8g .. 8g + 7 into its group map; the last workgroup to arrive
// (acquire-release counter) runs the chain over the group maps, writes the
// group end states and releases them with a new epoch; every other workgroup
// waits for that epoch (bounded: hitting the cap sets *err), then applies its
// eight maps from its group's incoming state.  All groups are resident (at
// most FUSE_MAX_GROUPS workgroups), so the wait ends.  ss: {count, epoch},
// zeroed when the scratch is allocated; the counter is reset by the last
// workgroup, the epoch only grows.
static constexpr int FUSE_MAX_GROUPS = 32;
static constexpr unsigned FUSE_SPIN_CAP = 1u << 21;

template <typename T>
__global__ __launch_bounds__(256) void sfull_spine_fused_kernel(SFArgs<T> lvl, SFArgs<T> top, unsigned* ss,
                                                                 unsigned* err) {
    if (lvl.skip && *(volatile const int*)lvl.skip) return;
    __shared__ unsigned e_start;
    __shared__ int is_last;
    if (threadIdx.x == 0) e_start = __hip_atomic_load(&ss[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sfull_reduce_tile<T>(lvl, blockIdx.x);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's map stores are done
    __syncthreads();  // ... for every wave, before thread 0's release below
    if (threadIdx.x == 0) {
        const unsigned prev = __hip_atomic_fetch_add(&ss[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = prev == gridDim.x - 1u;
    }
    __syncthreads();
    const unsigned e = e_start;
    if (is_last) {
        sfull_apply_tile<T>(top, 0);  // group end states
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        if (threadIdx.x == 0) {
            ss[0] = 0u;
            __hip_atomic_store(&ss[1], e + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else {
        if (threadIdx.x == 0) {
            unsigned spins = 0;
            while (__hip_atomic_load(&ss[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == e) {
                if (++spins > FUSE_SPIN_CAP) {
                    atomicOr(err, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __syncthreads();
    }
    sfull_apply_tile<T>(lvl, blockIdx.x);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int level_tc(int level) { return level == 0 ? 64 : 8; }

template <typename T>
static size_t sfull_bytes_t(int64_t T_len, int level0 = 0) {
    size_t bytes = 0;
    int64_t len = T_len;
    for (int level = level0; len > level_tc(level); ++level) {
        const int64_t nT = (len + level_tc(level) - 1) / level_tc(level);
        bytes += (size_t)nT * (SF * SF + 2 * SF) * sizeof(T);
        len = nT;
    }
    return bytes;
}

size_t scan_full_scratch_bytes(int64_t T, bool f64) { return f64 ? sfull_bytes_t<double>(T) : sfull_bytes_t<float>(T); }

static constexpr size_t SPINE_SYNC_BYTES = 256;  // the fused launch's counter and epoch, then the maps

size_t scan_full_spine_scratch_bytes(int64_t nT, bool f64) {
    return SPINE_SYNC_BYTES + (f64 ? sfull_bytes_t<double>(nT, 1) : sfull_bytes_t<float>(nT, 1));
}

template <typename T>
static void sfull_rec(int level, const T* J, int64_t ldJ, int64_t strideJ, const T* B, int64_t ldB, const T* s0,
                      T* S, int64_t ldS, int64_t D, int64_t len, char* scratch, hipStream_t st,
                      const int* skip = nullptr) {
    SFArgs<T> a{};
    a.skip = skip;
    a.J = J;
    a.ldJ = ldJ;
    a.strideJ = strideJ;
    a.B = B;
    a.ldB = ldB;
    a.s0 = s0;
    a.S = S;
    a.ldS = ldS;
    a.D = D;
    a.len = len;
    a.TC = level_tc(level);
    if (len <= a.TC) {
        a.nT = 1;
        a.prev = nullptr;
        hipLaunchKernelGGL(sfull_apply_kernel<T>, dim3(1), dim3(256), 0, st, a);
        return;
    }
    a.nT = (int)((len + a.TC - 1) / a.TC);
    T* Magg = reinterpret_cast<T*>(scratch);
    T* ragg = Magg + (size_t)a.nT * SF * SF;
    T* dend = ragg + (size_t)a.nT * SF;
    char* rest = reinterpret_cast<char*>(dend + (size_t)a.nT * SF);
    a.Magg = Magg;
    a.ragg = ragg;
    hipLaunchKernelGGL(sfull_reduce_kernel<T>, dim3((unsigned)a.nT), dim3(256), 0, st, a);
    // state after every tile: the dense scan of the tile maps
    sfull_rec<T>(level + 1, Magg, SF, SF * SF, ragg, SF, s0, dend, SF, D, a.nT, rest, st, skip);
    a.prev = dend;
    hipLaunchKernelGGL(sfull_apply_kernel<T>, dim3((unsigned)a.nT), dim3(256), 0, st, a);
}

int scan_full(bool f64, const void* J, const void* B, const void* s0, void* S, int64_t D, int64_t T, void* scratch,
              hipStream_t st, std::string* msg) {
    if (T == 0) return PMCMC_OK;
    if (f64)
        sfull_rec<double>(0, (const double*)J, D, D * D, (const double*)B, D, (const double*)s0, (double*)S, D, D, T,
                          (char*)scratch, st);
    else
        sfull_rec<float>(0, (const float*)J, D, D * D, (const float*)B, D, (const float*)s0, (float*)S, D, D, T,
                         (char*)scratch, st);
    if (hipGetLastError() != hipSuccess) {
        *msg = "dense scan: kernel launch failed";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}

template <typename T>
void scan_full_spine(const T* M, const T* r, T* out, int64_t D, int64_t nT, const int* skip, void* scratch,
                     hipStream_t st, unsigned* err) {
    if (nT <= 0) return;
    unsigned* ss = reinterpret_cast<unsigned*>(scratch);  // fused launch: {count, epoch}
    char* maps = (char*)scratch + SPINE_SYNC_BYTES;
    const int64_t G = (nT + level_tc(1) - 1) / level_tc(1);
    if (nT <= level_tc(1) || G > FUSE_MAX_GROUPS || !err) {
        sfull_rec<T>(1, M, SF, SF * SF, r, SF, nullptr, out, SF, D, nT, maps, st, skip);
        return;
    }
    // the level-1 scratch layout of sfull_rec: group maps, group r, group end states
    T* Mg = reinterpret_cast<T*>(maps);
    T* rg = Mg + (size_t)G * SF * SF;
    T* eg = rg + (size_t)G * SF;
    SFArgs<T> lvl{};
    lvl.J = M;
    lvl.ldJ = SF;
    lvl.strideJ = SF * SF;
    lvl.B = r;
    lvl.ldB = SF;
    lvl.s0 = nullptr;
    lvl.prev = eg;
    lvl.S = out;
    lvl.ldS = SF;
    lvl.D = D;
    lvl.len = nT;
    lvl.TC = level_tc(1);
    lvl.nT = (int)G;
    lvl.Magg = Mg;
    lvl.ragg = rg;
    lvl.skip = skip;
    SFArgs<T> top{};
    top.J = Mg;
    top.ldJ = SF;
    top.strideJ = SF * SF;
    top.B = rg;
    top.ldB = SF;
    top.s0 = nullptr;
    top.S = eg;
    top.ldS = SF;
    top.D = D;
    top.len = G;
    top.TC = (int)G;
    top.nT = 1;
    hipLaunchKernelGGL(sfull_spine_fused_kernel<T>, dim3((unsigned)G), dim3(256), 0, st, lvl, top, ss, err);
}

template void scan_full_spine<float>(const float*, const float*, float*, int64_t, int64_t, const int*, void*,
                                     hipStream_t, unsigned*);
template void scan_full_spine<double>(const double*, const double*, double*, int64_t, int64_t, const int*, void*,
                                      hipStream_t, unsigned*);

}  // namespace pmcmc
