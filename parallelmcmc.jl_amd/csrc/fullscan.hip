// Full-Jacobian DEER for the dense Gaussian target N(mean, Lambda^-1), D <= 64.
//
// For H = -Lambda the Jacobian of the surrogate MALA step (the matrix whose
// products are the JVPs of src/MALA/MALA.jl:749-795) is structured:
//   J_t = I - a_t eps Lambda + u_t c_t^T,
//   a_t = accept flag, u_t = g(1-g)(y - x)  (u_t = 0 for the decision-frozen
//   Jacobian, the derivative of the taped step away from decision flips),
//   c_t = B g_y - g_x - (r - B^2 r)/(2 eps),  B = I - eps Lambda,  r = x - y - eps g_y
//       = g_y - eps Lambda g_y - g_x - Lambda r + (eps/2) Lambda (Lambda r).
// The Newton update is solved in correction form: with rho_t = FT_t - S_t,
//   Delta_t = J_t Delta_{t-1} + rho_t,  Delta_{-1} = 0,  S_new = S + damping Delta,
// which equals the reference's S_new_t = J_t S_new_{t-1} + FT_t - J_t Xbar_t
// (src/DEER/DEER.jl:409-419, 470-473) in exact arithmetic and keeps the fixed
// point exact when the Jacobian products are rounded (BF16 MFMA): at the
// fixed point rho = 0 and so Delta = 0 for any J.
//
// Kernels of one iteration:
//   full_prep_kernel   per column: log alpha (FP64), accept flag, u, c, rho
//   full_reduce_*      per tile of TC columns: M = J_e ... J_b (64 x 64) and
//                      r = the tile's Delta from a zero start.  M is held in
//                      registers (fp32 master); Lambda M runs on MFMA with
//                      Lambda's operand fragments in registers and M's from an
//                      LDS copy (FP32 16x16x4 or BF16 16x16x32; FP64 on VALU)
//   spine              Delta at each tile end, Delta_k = M_k Delta_{k-1} + r_k:
//                      a dense scan over the tile maps, solved by the
//                      recursive reduce/apply kernels of scanfull.hip
//                      (8-map tiles, MFMA composition), not a chain of
//                      dependent mat-vecs
//   full_apply_kernel  one wave per tile: Delta through the tile from the
//                      incoming Delta, S_new, convergence maxima, metric
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>

#include "../../include/pmcmc_amd.h"
#include "fullscan.hpp"
#include "scanfull.hpp"

namespace pmcmc {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

static constexpr int FD = FULL_MAX_D;  // padded dimension
static constexpr int FS = FD + 4;      // fp32 LDS column stride
static constexpr int FSB = FD + 8;     // bf16 LDS column stride

template <typename T>
struct FullArgs {
    const T* S_in;
    const T* s0;
    const T* U;
    const T* Y;
    const T* gx;
    const T* gy;
    const double* lpx;
    const double* lpy;
    const float* L32;   // Dp x Dp row-major, zero padded
    const double* L64;  // D x D column-major
    int64_t D, Dp, Tn;
    int64_t col0;  // first column of tile 0 (the columns before it are final)
    int TC, nT;
    double eps, damping, tol_abs, tol_rel;
    int sigmoid;
    T* a;
    T* u;
    T* c;
    T* rho;
    T* Magg;
    T* ragg;
    T* dend;
    T* S_out;
    uint8_t* acc_out;
    SyncState* sy;
    double* tile_smax;  // max |S| of each tile when it last ran (the final value once it is skipped)
    int iter;
};

// First column of this iteration's work: columns below it are final (their
// residual FT - S was exactly 0, so both iterate buffers hold the fixed point
// there).  A multiple of TC.
__device__ inline int64_t live_col(const SyncState* sy) { return (int64_t)*(volatile const int*)&sy->skip_cols; }

// A reduce workgroup whose tile is final: the spine then carries Delta = 0
// through it (its map and its Delta from a zero start are written as 0).
template <typename T>
__device__ inline bool final_tile_zero_map(const FullArgs<T>& p) {
    const int64_t t0 = p.col0 + (int64_t)blockIdx.x * p.TC;
    if (t0 + p.TC > live_col(p.sy)) return false;
    T* Mo = p.Magg + (size_t)blockIdx.x * FD * FD;
    for (int e = threadIdx.x; e < FD * FD; e += blockDim.x) Mo[e] = (T)0;
    if (threadIdx.x < FD) p.ragg[(size_t)blockIdx.x * FD + threadIdx.x] = (T)0;
    return true;
}

__device__ inline float rlane(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
__device__ inline double rlane(double v, int k) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Lambda[r][k] with zero padding beyond D
template <typename T>
__device__ inline T lam(const FullArgs<T>& p, int r, int k) {
    if (r >= p.D || k >= p.D) return (T)0;
    if constexpr (sizeof(T) == 8)
        return p.L64[r + (int64_t)k * p.D];
    else
        return p.L32[(int64_t)r * p.Dp + k];
}

__device__ inline bool converged(const SyncState* sy) { return *(volatile const int*)&sy->converged != 0; }

// ---------------------------------------------------------------------------
// prep: one wave per column (grid-stride over columns), lane d = row d;
// lane d keeps Lambda's row d in registers.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void full_prep_kernel(FullArgs<T> p) {
    if (converged(p.sy)) return;
    const int lane = threadIdx.x & 63;
    const int64_t D = p.D;
    T lrow[FD];  // only the surrogate's c needs Lambda
    if (p.sigmoid) {
#pragma unroll
        for (int k = 0; k < FD; ++k) lrow[k] = lam<T>(p, lane, k);
    }
    const bool in = lane < D;
    const double eps = p.eps;
    const double c_norm = ((double)D / 2.0) * log(4.0 * 3.14159265358979323846 * eps);
    const int64_t nw = (int64_t)gridDim.x * 4;
    for (int64_t t = live_col(p.sy) + (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); t < p.Tn; t += nw) {
        const size_t gi = (size_t)t * D + lane;
        T xv = 0, yv = 0, g0 = 0, g1 = 0, sv = 0;
        if (in) {
            xv = t == 0 ? p.s0[lane] : p.S_in[gi - D];
            yv = p.Y[gi];
            g0 = p.gx[gi];
            g1 = p.gy[gi];
            sv = p.S_in[gi];
        }
        const double e1 = (double)yv - (double)xv - eps * (double)g0;
        const double r = (double)xv - (double)yv - eps * (double)g1;
        double qyx = e1 * e1, qxy = r * r;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            qyx += __shfl_xor(qyx, m, 64);
            qxy += __shfl_xor(qxy, m, 64);
        }
        // log q(x|y) and log q(y|x), src/MALA/MALA.jl:172-193
        const double lq_xy = -0.5 * qxy / (2.0 * eps) - c_norm;
        const double lq_yx = -0.5 * qyx / (2.0 * eps) - c_norm;
        const double la = (p.lpy[t] + lq_xy) - (p.lpx[t] + lq_yx);
        const double lu = log((double)p.U[t]);
        const bool acc = lu < la;
        const double g = 1.0 / (1.0 + exp(-(la - lu)));
        T cv = 0, uv = 0;  // the rank-1 term is absent from the decision-frozen Jacobian
        if (p.sigmoid) {
            const T rT = (T)r;
            T Lg = 0, Lr = 0;
#pragma unroll
            for (int k = 0; k < FD; ++k) {
                Lg += lrow[k] * rlane(g1, k);
                Lr += lrow[k] * rlane(rT, k);
            }
            T LLr = 0;
#pragma unroll
            for (int k = 0; k < FD; ++k) LLr += lrow[k] * rlane(Lr, k);
            const T epsT = (T)eps;
            cv = g1 - epsT * Lg - g0 - Lr + (T)0.5 * epsT * LLr;
            uv = (T)(g * (1.0 - g)) * (yv - xv);
        }
        const T ft = acc ? yv : xv;
        if (in) {
            if (p.sigmoid) {
                p.u[gi] = uv;
                p.c[gi] = cv;
            }
            p.rho[gi] = ft - sv;
        }
        if (lane == 0) {
            p.a[t] = acc ? (T)1 : (T)0;
            if (p.acc_out) p.acc_out[t] = acc ? 1 : 0;
        }
    }
}

// ---------------------------------------------------------------------------
// reduce, MFMA paths (T = float).  The tile's state is the augmented 64 x 80
// matrix [M | Delta | 0]: column tiles 0..3 hold M, column 64 (tile 4) holds
// the tile's Delta, so one step is [M | Delta] <- J_t [M | Delta] + [0 | rho_t]
// and Lambda [M | Delta] is a single 64 x 80 x 64 MFMA product.
// 256 threads; thread (wave w, lane = li + 16 lk) owns entries (i, j) with
// i = 16w + 4lk + r (r < 4), j = 16jt + li (jt < 5): the C-fragment layout of
// a 16x16 MFMA tile, so Lambda [M | Delta] lands on the entries the thread
// holds (fp32 master in registers).  The B operand comes from an LDS copy
// (double-buffered: one barrier per step), Lambda's A fragments stay in
// registers.  FP32 16x16x4 runs its k-slots as k = 16 lk + ks so a lane's 16
// values of one column are contiguous (4 x 128-bit LDS reads).  u, c, rho and
// the accept flags are staged through LDS in chunks of FCH columns, the next
// chunk's global loads in flight while the current one is consumed.
// With SIG (sigmoid surrogate) the rank-1 term u (c^T [M | Delta]) needs
// w = c^T [M | Delta]: its per-wave partials are formed right after the
// update with the next step's c and published with the copy.
// ---------------------------------------------------------------------------
static constexpr int FNJT = 5;          // column tiles of [M | Delta]
static constexpr int FNC = 16 * FNJT;  // 80 columns
static constexpr int FCH = 8;           // steps per staged chunk

template <int MM, bool SIG>  // MM 0: FP32 MFMA 16x16x4, 1: BF16 MFMA 16x16x32
__global__ __launch_bounds__(256) void full_reduce_mfma_kernel(FullArgs<float> p) {
    __shared__ __attribute__((aligned(16))) float Mf[MM == 0 ? 2 * FNC * FS : 4];
    __shared__ __attribute__((aligned(16))) __bf16 Mb[MM == 1 ? 2 * FNC * FSB : 8];
    __shared__ __attribute__((aligned(16))) float part[SIG ? 2 : 1][4][FNC];
    __shared__ __attribute__((aligned(16))) float su[SIG ? 2 : 1][FCH][FD];
    __shared__ __attribute__((aligned(16))) float sc[SIG ? 2 : 1][FCH][FD];
    __shared__ __attribute__((aligned(16))) float sr[2][FCH][FD];
    __shared__ float sa[2][FCH];
    if (converged(p.sy)) return;
    if (final_tile_zero_map(p)) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int li = lane & 15, lk = lane >> 4;
    const int64_t D = p.D;
    const int64_t t0 = p.col0 + (int64_t)blockIdx.x * p.TC;
    const int tc = (int)((p.Tn - t0) < p.TC ? (p.Tn - t0) : p.TC);
    const float epsf = (float)p.eps;
    const int ib = 16 * wave + 4 * lk;  // first owned row

    // Lambda's A fragments (row tile = wave), constant over the tile
    float af[FD / 4];
    v8bf ab[FD / 32];
    if (MM == 0) {
#pragma unroll
        for (int ks = 0; ks < FD / 4; ++ks) af[ks] = lam<float>(p, 16 * wave + li, 16 * lk + ks);
    } else {
#pragma unroll
        for (int ks = 0; ks < FD / 32; ++ks)
#pragma unroll
            for (int e = 0; e < 8; ++e) ab[ks][e] = (__bf16)lam<float>(p, 16 * wave + li, 32 * ks + 8 * lk + e);
    }

    // staging: zero the padded rows once, then chunk loads (register prefetch)
    for (int i = tid; i < 2 * FCH * FD; i += 256) {
        (&sr[0][0][0])[i] = 0.f;
        if (SIG) {
            (&su[0][0][0])[i] = 0.f;
            (&sc[0][0][0])[i] = 0.f;
        }
    }
    constexpr int PER = (FCH * FD + 255) / 256;  // staged elements per thread and array
    float pu[PER], pc[PER], pr[PER], pa = 0.f;
    auto fetch = [&](int chunk) {  // global -> registers
        const int64_t cb = t0 + (int64_t)chunk * FCH;
        const int ncol = (tc - chunk * FCH) < FCH ? (tc - chunk * FCH) : FCH;
        const int n = ncol > 0 ? ncol * (int)D : 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = tid + 256 * q;
            const bool ok = e < n;
            const size_t g = (size_t)cb * D + e;
            pr[q] = ok ? p.rho[g] : 0.f;
            if (SIG) {
                pu[q] = ok ? p.u[g] : 0.f;
                pc[q] = ok ? p.c[g] : 0.f;
            }
        }
        if (tid < ncol) pa = p.a[cb + tid];
    };
    auto stash = [&](int chunk) {  // registers -> LDS buffer chunk & 1
        const int b = chunk & 1;
        const int ncol = (tc - chunk * FCH) < FCH ? (tc - chunk * FCH) : FCH;
        const int n = ncol > 0 ? ncol * (int)D : 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = tid + 256 * q;
            if (e < n) {
                const int col = e / (int)D, row = e - col * (int)D;
                sr[b][col][row] = pr[q];
                if (SIG) {
                    su[b][col][row] = pu[q];
                    sc[b][col][row] = pc[q];
                }
            }
        }
        if (tid < FCH) sa[b][tid] = tid < ncol ? pa : 0.f;
    };
    const int nchunks = (tc + FCH - 1) / FCH;
    __syncthreads();  // zero fill done
    fetch(0);
    stash(0);
    if (nchunks > 1) fetch(1);

    // [M | Delta] = [I | 0] in registers
    float m[FNJT][4];
#pragma unroll
    for (int jt = 0; jt < FNJT; ++jt)
#pragma unroll
        for (int r = 0; r < 4; ++r) m[jt][r] = (ib + r == 16 * jt + li) ? 1.f : 0.f;

    auto put_copy = [&](int b) {
#pragma unroll
        for (int jt = 0; jt < FNJT; ++jt) {
            const int j = 16 * jt + li;
            if (MM == 0) {
                *reinterpret_cast<float4*>(&Mf[(b * FNC + j) * FS + ib]) =
                    make_float4(m[jt][0], m[jt][1], m[jt][2], m[jt][3]);
            } else {
                v4bf v;
                v[0] = (__bf16)m[jt][0];
                v[1] = (__bf16)m[jt][1];
                v[2] = (__bf16)m[jt][2];
                v[3] = (__bf16)m[jt][3];
                *reinterpret_cast<v4bf*>(&Mb[(b * FNC + j) * FSB + ib]) = v;
            }
        }
    };
    // per-wave partials of w = c^T [M | Delta] for the step about to run
    auto put_w = [&](int b, const float* cv) {
#pragma unroll
        for (int jt = 0; jt < FNJT; ++jt) {
            float wp = cv[0] * m[jt][0] + cv[1] * m[jt][1] + cv[2] * m[jt][2] + cv[3] * m[jt][3];
            wp += __shfl_xor(wp, 16, 64);
            wp += __shfl_xor(wp, 32, 64);
            if (lk == 0) part[b][wave][16 * jt + li] = wp;
        }
    };
    put_copy(0);
    if (SIG) {
        __syncthreads();  // chunk 0 staged
        const float4 c4 = *reinterpret_cast<const float4*>(&sc[0][0][ib]);
        const float cv[4] = {c4.x, c4.y, c4.z, c4.w};
        put_w(0, cv);
    }

    for (int s = 0; s < tc; ++s) {
        const int ch = s / FCH, col = s - ch * FCH, sb = ch & 1;
        __syncthreads();  // copy s&1, partials s&1 and chunk ch visible; reads of step s-1 done
        if (col == 0 && ch + 1 < nchunks) {
            stash(ch + 1);  // buffer (ch+1)&1 was last read in chunk ch-1
            if (ch + 2 < nchunks) fetch(ch + 2);
        }
        const int b = s & 1;
        v4f acc[FNJT];
#pragma unroll
        for (int jt = 0; jt < FNJT; ++jt) acc[jt] = (v4f){0.f, 0.f, 0.f, 0.f};
        if (MM == 0) {
#pragma unroll
            for (int jt = 0; jt < FNJT; ++jt) {
                const float* src = &Mf[(b * FNC + 16 * jt + li) * FS + 16 * lk];
                float bv[16];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 v = *reinterpret_cast<const float4*>(src + 4 * q);
                    bv[4 * q] = v.x;
                    bv[4 * q + 1] = v.y;
                    bv[4 * q + 2] = v.z;
                    bv[4 * q + 3] = v.w;
                }
#pragma unroll
                for (int ks = 0; ks < FD / 4; ++ks)
                    acc[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[ks], bv[ks], acc[jt], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int jt = 0; jt < FNJT; ++jt)
#pragma unroll
                for (int ks = 0; ks < FD / 32; ++ks) {
                    const v8bf bv =
                        *reinterpret_cast<const v8bf*>(&Mb[(b * FNC + 16 * jt + li) * FSB + 32 * ks + 8 * lk]);
                    acc[jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ab[ks], bv, acc[jt], 0, 0, 0);
                }
        }
        const float ae = sa[sb][col] * epsf;
        const float4 r4 = *reinterpret_cast<const float4*>(&sr[sb][col][ib]);
        const float rho[4] = {r4.x, r4.y, r4.z, r4.w};
        float uu[4] = {0.f, 0.f, 0.f, 0.f};
        if (SIG) {
            const float4 u4 = *reinterpret_cast<const float4*>(&su[sb][col][ib]);
            uu[0] = u4.x;
            uu[1] = u4.y;
            uu[2] = u4.z;
            uu[3] = u4.w;
        }
#pragma unroll
        for (int jt = 0; jt < FNJT; ++jt) {
            const int j = 16 * jt + li;
            float wj = 0.f;
            if (SIG) wj = part[b][0][j] + part[b][1][j] + part[b][2][j] + part[b][3][j];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float v = m[jt][r] - ae * acc[jt][r];
                if (SIG) v += uu[r] * wj;
                if (jt == FNJT - 1 && li == 0) v += rho[r];  // column 64: Delta
                m[jt][r] = v;
            }
        }
        if (s + 1 < tc) {
            put_copy(b ^ 1);
            if (SIG) {
                const int s1 = s + 1, ch1 = s1 / FCH, col1 = s1 - ch1 * FCH;
                const float4 c4 = *reinterpret_cast<const float4*>(&sc[ch1 & 1][col1][ib]);
                const float cv[4] = {c4.x, c4.y, c4.z, c4.w};
                put_w(b ^ 1, cv);
            }
        }
    }
    // tile aggregates: M column-major 64 x 64, Delta
    float* Mo = p.Magg + (size_t)blockIdx.x * FD * FD;
#pragma unroll
    for (int jt = 0; jt < FNJT - 1; ++jt) {
        const int j = 16 * jt + li;
        *reinterpret_cast<float4*>(&Mo[j * FD + ib]) = make_float4(m[jt][0], m[jt][1], m[jt][2], m[jt][3]);
    }
    if (li == 0)
        *reinterpret_cast<float4*>(&p.ragg[(size_t)blockIdx.x * FD + ib]) =
            make_float4(m[FNJT - 1][0], m[FNJT - 1][1], m[FNJT - 1][2], m[FNJT - 1][3]);
}

// ---------------------------------------------------------------------------
// reduce, FP64 on VALU.  Thread (wave w, lane j) owns M[16w + ii][j], ii < 16;
// the LDS copy is row-major so that lane j reads column j conflict-free.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void full_reduce_f64_kernel(FullArgs<double> p) {
    __shared__ double Mr[FD * FD];
    __shared__ double part[4][FD];
    __shared__ double uvec[FD];
    if (converged(p.sy)) return;
    if (final_tile_zero_map(p)) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t D = p.D;
    const int64_t t0 = p.col0 + (int64_t)blockIdx.x * p.TC;
    const int tc = (int)((p.Tn - t0) < p.TC ? (p.Tn - t0) : p.TC);
    const double eps = p.eps;
    double lrow[FD];
    if (wave == 0) {
#pragma unroll
        for (int k = 0; k < FD; ++k) lrow[k] = lam<double>(p, lane, k);
    }
    double m[16];
#pragma unroll
    for (int ii = 0; ii < 16; ++ii) {
        m[ii] = (16 * wave + ii == lane) ? 1.0 : 0.0;
        Mr[(16 * wave + ii) * FD + lane] = m[ii];
    }
    double rv = 0.0;
    const bool in = lane < D;
    for (int s = 0; s < tc; ++s) {
        const int64_t t = t0 + s;
        const double a = p.a[t];
        const bool sg = in && p.sigmoid;  // u and c exist only for the sigmoid variant
        const double c0 = sg ? p.c[(size_t)t * D + lane] : 0.0;
        const double u0 = sg ? p.u[(size_t)t * D + lane] : 0.0;
        const double r0 = in ? p.rho[(size_t)t * D + lane] : 0.0;
        if (wave == 0) uvec[lane] = u0;
        __syncthreads();  // copy of M and u complete
        double acc[16];
#pragma unroll
        for (int ii = 0; ii < 16; ++ii) acc[ii] = 0.0;
        for (int k = 0; k < D; ++k) {
            const double mk = Mr[k * FD + lane];
#pragma unroll
            for (int ii = 0; ii < 16; ++ii) {
                const int i = 16 * wave + ii;
                const double l = i < D ? p.L64[i + (int64_t)k * D] : 0.0;
                acc[ii] += l * mk;
            }
        }
        // w_j partial over the owned rows: c_i from lane i via readlane
        double wp = 0.0;
#pragma unroll
        for (int ii = 0; ii < 16; ++ii) wp += rlane(c0, 16 * wave + ii) * m[ii];
        part[wave][lane] = wp;
        if (wave == 0) {
            double Lr = 0.0;
#pragma unroll
            for (int k = 0; k < FD; ++k) Lr += lrow[k] * rlane(rv, k);
            double cr = c0 * rv;
#pragma unroll
            for (int mm = 32; mm >= 1; mm >>= 1) cr += __shfl_xor(cr, mm, 64);
            rv = rv - a * eps * Lr + u0 * cr + r0;
        }
        __syncthreads();  // partials visible; all reads of the copy done
        const double wj = part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane];
        const double ae = a * eps;
#pragma unroll
        for (int ii = 0; ii < 16; ++ii) {
            m[ii] = m[ii] - ae * acc[ii] + uvec[16 * wave + ii] * wj;
            Mr[(16 * wave + ii) * FD + lane] = m[ii];
        }
        __syncthreads();  // uvec is rewritten next step
    }
    double* Mo = p.Magg + (size_t)blockIdx.x * FD * FD;
#pragma unroll
    for (int ii = 0; ii < 16; ++ii) Mo[lane * FD + 16 * wave + ii] = m[ii];
    if (wave == 0) p.ragg[(size_t)blockIdx.x * FD + lane] = rv;
}

// ---------------------------------------------------------------------------
// apply: one wave per tile; lane d = row d with Lambda's row d in registers.
// The step's operands stream through a ring of APF register slots.
// ---------------------------------------------------------------------------
static constexpr int APF = 8;

// End of an apply wave: maxima over the wave, the tile's cached max |S| and
// its residual report (live tiles), the iteration's maxima; the last wave
// forms the metric and the next iteration's final prefix.
template <typename T>
__device__ inline void full_apply_finish(const FullArgs<T>& p, double dmax, double smax, bool nz, bool final_tile) {
    const int lane = threadIdx.x;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        dmax = nan_max(dmax, __shfl_xor(dmax, m, 64));
        smax = nan_max(smax, __shfl_xor(smax, m, 64));
    }
    const bool any_nz = __any(nz ? 1 : 0) != 0;
    if (lane == 0) {
        SyncState* sy = p.sy;
        int live_col = -1;
        if (!final_tile) {
            p.tile_smax[blockIdx.x] = smax;
            if (any_nz) live_col = (int)(p.col0 + (int64_t)blockIdx.x * p.TC);
        }
        double dm, smt;
        if (sync_arrive(sy, (unsigned)p.nT, dmax, smax, live_col, &dm, &smt)) {
            // smax_prefix: written by an earlier launch
            const double sm = nan_max(smt, sy->smax_prefix);
            // src/DEER/DEER.jl:567-577
            const double metric = dm / (p.tol_abs + p.tol_rel * sm);
            sy->last_metric = metric;
            sy->last_iter = p.iter;
            if (metric <= 1.0) {
                sy->conv_iter = p.iter;
                sy->converged = 1;
            }
            // tiles before the first one with a nonzero residual are final
            const int lf = sync_take_live_from(sy);
            sy->skip_cols = lf == INT_MAX ? (int)(p.col0 + (int64_t)p.nT * p.TC) : lf;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(64) void full_apply_kernel(FullArgs<T> p) {
    __shared__ __attribute__((aligned(16))) T dvs[FD];  // Delta of the previous step, for the broadcast reads
    if (converged(p.sy)) return;
    const int lane = threadIdx.x;
    const int64_t D = p.D;
    const int64_t tile = blockIdx.x;
    const int64_t t0 = p.col0 + tile * p.TC;
    const int tc = (int)((p.Tn - t0) < p.TC ? (p.Tn - t0) : p.TC);
    if (t0 + p.TC <= live_col(p.sy)) {  // final tile: S_out = S_in there already
        full_apply_finish<T>(p, 0.0, lane == 0 ? p.tile_smax[tile] : 0.0, false, true);
        return;
    }
    T lrow[FD];
#pragma unroll
    for (int k = 0; k < FD; ++k) lrow[k] = lam<T>(p, lane, k);
    const bool in = lane < D;
    const bool sig = p.sigmoid != 0;
    const T epsT = (T)p.eps, damp = (T)p.damping;
    T dv = tile == 0 ? (T)0 : p.dend[(size_t)(tile - 1) * FD + lane];
    double dmax = 0.0, smax = 0.0;
    bool nz = false;  // a nonzero residual in this tile
    T ru[APF], rc[APF], rr[APF], rs[APF], ra[APF];
    auto fetch = [&](int s, int sl) {
        const size_t gi = (size_t)(t0 + s) * D + lane;
        ru[sl] = (T)0;
        rc[sl] = (T)0;
        rr[sl] = (T)0;
        rs[sl] = (T)0;
        if (in) {
            if (sig) {
                ru[sl] = p.u[gi];
                rc[sl] = p.c[gi];
            }
            rr[sl] = p.rho[gi];
            rs[sl] = p.S_in[gi];
        }
        ra[sl] = p.a[t0 + s];
    };
#pragma unroll
    for (int sl = 0; sl < APF; ++sl)
        if (sl < tc) fetch(sl, sl);
    for (int s0 = 0; s0 < tc; s0 += APF) {
#pragma unroll
        for (int sl = 0; sl < APF; ++sl) {
            const int s = s0 + sl;
            if (s < tc) {
                const T u = ru[sl], c = rc[sl], rho = rr[sl], old = rs[sl], a = ra[sl];
                if (s + APF < tc) fetch(s + APF, sl);
                // Lambda Delta: the wave's Delta through LDS (one write, then
                // broadcast vector reads), four independent FMA chains.  The
                // workgroup is one wave and a wave's LDS operations complete in
                // issue order, so only the compiler's order needs pinning (a
                // __syncthreads would also wait for the prefetch ring's loads)
                dvs[lane] = dv;
                __builtin_amdgcn_wave_barrier();
                T L0 = 0, L1 = 0, L2 = 0, L3 = 0;
#pragma unroll
                for (int k = 0; k < FD; k += 4) {
                    L0 += lrow[k] * dvs[k];
                    L1 += lrow[k + 1] * dvs[k + 1];
                    L2 += lrow[k + 2] * dvs[k + 2];
                    L3 += lrow[k + 3] * dvs[k + 3];
                }
                __builtin_amdgcn_wave_barrier();  // reads before the next step's write
                T nd = dv - a * epsT * ((L0 + L1) + (L2 + L3)) + rho;
                if (sig) {
                    T cd = c * dv;
#pragma unroll
                    for (int m = 32; m >= 1; m >>= 1) cd += __shfl_xor(cd, m, 64);
                    nd += u * cd;
                }
                dv = nd;
                nz = nz || rho != (T)0;  // NaN counts as nonzero
                if (in) {
                    const T nv = old + damp * dv;
                    p.S_out[(size_t)(t0 + s) * D + lane] = nv;
                    dmax = nan_max(dmax, fabs((double)nv - (double)old));
                    smax = nan_max(smax, fabs((double)nv));
                }
            }
        }
    }
    full_apply_finish<T>(p, dmax, smax, nz, false);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int full_work_ensure(FullWork* w, int64_t D, int64_t T, bool f64, std::string* msg) {
    if (D > FD) {
        *msg = "jacobian=:full supports D <= 64";
        return PMCMC_ERR_UNSUPPORTED;
    }
    const size_t esz = f64 ? 8 : 4;
    if (w->a && w->T == T && w->D == D && w->esz == esz) return PMCMC_OK;
    full_work_destroy(w);
    w->D = D;
    w->T = T;
    w->esz = esz;
    w->TC = 64;
    w->nT = (int)((T + w->TC - 1) / w->TC);
    const size_t mat = (size_t)D * T * esz;
    bool ok = hipMalloc(&w->a, T * esz) == hipSuccess && hipMalloc(&w->u, mat) == hipSuccess &&
              hipMalloc(&w->c, mat) == hipSuccess && hipMalloc(&w->rho, mat) == hipSuccess &&
              hipMalloc(&w->Magg, (size_t)w->nT * FD * FD * esz) == hipSuccess &&
              hipMalloc(&w->ragg, (size_t)w->nT * FD * esz) == hipSuccess &&
              hipMalloc(&w->dend, (size_t)w->nT * FD * esz) == hipSuccess &&
              hipMalloc(&w->spine, scan_full_spine_scratch_bytes(w->nT, f64)) == hipSuccess &&
              // the fused spine's counter and epoch start at zero
              hipMemset(w->spine, 0, scan_full_spine_scratch_bytes(w->nT, f64)) == hipSuccess &&
              hipMalloc(&w->tile_smax, (size_t)w->nT * sizeof(double)) == hipSuccess &&
              // the spine writes rows < D only; the apply reads all 64 lanes
              hipMemset(w->dend, 0, (size_t)w->nT * FD * esz) == hipSuccess;
    if (!ok) {
        full_work_destroy(w);
        *msg = "full-Jacobian workspace allocation failed";
        return PMCMC_ERR_ALLOC;
    }
    return PMCMC_OK;
}

// max |S| of the tiles of the tiling (col0, TC, nT) that end at or before lo,
// folded into sy->smax_prefix; one wave
__global__ __launch_bounds__(64) void full_fold_prefix_kernel(SyncState* sy, const double* tile_smax, int64_t col0,
                                                               int TC, int nT, int64_t lo) {
    double m = 0.0;
    for (int b = threadIdx.x; b < nT; b += 64)
        if (col0 + (int64_t)(b + 1) * TC <= lo) m = nan_max(m, tile_smax[b]);
#pragma unroll
    for (int k = 32; k >= 1; k >>= 1) m = nan_max(m, __shfl_xor(m, k, 64));
    if (threadIdx.x == 0) sy->smax_prefix = nan_max(sy->smax_prefix, m);
}

void full_geometry_reset(FullWork* w) {
    w->col0 = 0;
    w->TCc = w->TC;
    w->nTc = w->nT;
}

void full_retile(FullWork* w, SyncState* sync, int64_t lo, hipStream_t s) {
    if (lo <= w->col0 || lo >= w->T) return;
    hipLaunchKernelGGL(full_fold_prefix_kernel, dim3(1), dim3(64), 0, s, sync, (const double*)w->tile_smax, w->col0,
                       w->TCc, w->nTc, lo);
    const int64_t live = w->T - lo;
    // per-iteration latency model of the tile kernels (measured at D = 64): the
    // reduce and apply chains cost about 2.5 us per column of a tile, the scan
    // over the tile maps about 5 us up to 8 maps, 35 us up to 64, 70 us beyond
    int best = w->TC;
    double best_us = 1e30;
    for (int tc = 8; tc <= w->TC; tc *= 2) {
        const int64_t n = (live + tc - 1) / tc;
        if (n > w->nT) continue;
        const double us = 2.5 * tc + (n <= 8 ? 5.0 : n <= 64 ? 35.0 : 70.0);
        if (us < best_us) {
            best_us = us;
            best = tc;
        }
    }
    w->col0 = lo;
    w->TCc = best;
    w->nTc = (int)((live + best - 1) / best);
}

void full_work_destroy(FullWork* w) {
    (void)hipFree(w->a);
    (void)hipFree(w->u);
    (void)hipFree(w->c);
    (void)hipFree(w->rho);
    (void)hipFree(w->Magg);
    (void)hipFree(w->ragg);
    (void)hipFree(w->dend);
    (void)hipFree(w->spine);
    (void)hipFree(w->tile_smax);
    *w = FullWork{};
}

template <typename T>
int full_newton_step(FullWork* fw, const GlmWork* gw, const GaussData* g, hipStream_t s, SyncState* sync,
                     int precision, bool sigmoid, double eps, double damping, double tol_abs, double tol_rel, const T* x0,
                     const T* S_in, T* S_out, const T* U, uint8_t* acc_out, int iter, Prof* prof,
                     std::string* msg) {
    FullArgs<T> p{};
    p.S_in = S_in;
    p.s0 = x0;
    p.U = U;
    p.Y = (const T*)gw->Y;
    p.gx = (const T*)gw->gx;
    p.gy = (const T*)gw->gy;
    p.lpx = gw->lpx;
    p.lpy = gw->lpy;
    p.L32 = g->L32;
    p.L64 = g->L64;
    p.D = fw->D;
    p.Dp = g->Dp;
    p.Tn = fw->T;
    p.col0 = fw->col0;
    p.TC = fw->TCc;
    p.nT = fw->nTc;
    p.eps = eps;
    p.damping = damping;
    p.tol_abs = tol_abs;
    p.tol_rel = tol_rel;
    p.sigmoid = sigmoid ? 1 : 0;
    p.a = (T*)fw->a;
    p.u = (T*)fw->u;
    p.c = (T*)fw->c;
    p.rho = (T*)fw->rho;
    p.Magg = (T*)fw->Magg;
    p.ragg = (T*)fw->ragg;
    p.dend = (T*)fw->dend;
    p.S_out = S_out;
    p.acc_out = acc_out;
    p.sy = sync;
    p.tile_smax = fw->tile_smax;
    p.iter = iter;
    int64_t pg = (fw->T + 3) / 4;
    if (pg > 2048) pg = 2048;
    int e = prof ? prof->begin(KT_TILE_REDUCE, iter, s) : -1;
    hipLaunchKernelGGL((full_prep_kernel<T>), dim3((unsigned)pg), dim3(256), 0, s, p);
    if constexpr (sizeof(T) == 8) {
        hipLaunchKernelGGL(full_reduce_f64_kernel, dim3((unsigned)fw->nTc), dim3(256), 0, s, p);
    } else {
        const dim3 g((unsigned)fw->nTc), blk(256);
        if (precision == PMCMC_PREC_BF16_JAC) {
            if (sigmoid)
                hipLaunchKernelGGL((full_reduce_mfma_kernel<1, true>), g, blk, 0, s, p);
            else
                hipLaunchKernelGGL((full_reduce_mfma_kernel<1, false>), g, blk, 0, s, p);
        } else {
            if (sigmoid)
                hipLaunchKernelGGL((full_reduce_mfma_kernel<0, true>), g, blk, 0, s, p);
            else
                hipLaunchKernelGGL((full_reduce_mfma_kernel<0, false>), g, blk, 0, s, p);
        }
    }
    if (prof) prof->end(e, s);
    e = prof ? prof->begin(KT_SCAN_SPINE, iter, s) : -1;
    scan_full_spine<T>(p.Magg, p.ragg, p.dend, fw->D, fw->nTc, &sync->converged, fw->spine, s, &sync->err);
    if (prof) prof->end(e, s);
    e = prof ? prof->begin(KT_TILE_APPLY, iter, s) : -1;
    hipLaunchKernelGGL((full_apply_kernel<T>), dim3((unsigned)fw->nTc), dim3(64), 0, s, p);
    if (prof) prof->end(e, s);
    if (hipGetLastError() != hipSuccess) {
        *msg = "full-Jacobian kernels failed to launch";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}

template int full_newton_step<float>(FullWork*, const GlmWork*, const GaussData*, hipStream_t, SyncState*, int,
                                     bool, double, double, double, double, const float*, const float*, float*,
                                     const float*, uint8_t*, int, Prof*, std::string*);
template int full_newton_step<double>(FullWork*, const GlmWork*, const GaussData*, hipStream_t, SyncState*, int,
                                      bool, double, double, double, double, const double*, const double*, double*,
                                      const double*, uint8_t*, int, Prof*, std::string*);

}  // namespace pmcmc
