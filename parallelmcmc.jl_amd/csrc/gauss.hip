// Dense Gaussian target: the Newton-iteration products as Lambda GEMMs.
//
// One pass over a column set [x - mean | z] (or [y - mean | w]) computes
// P = Lambda V on MFMA (K = D streamed through LDS in 32-wide slices) and, in
// the epilogue, g = -P_x, h = -P_z, logp = (x - mean).g / 2 and, for the pass
// at Xbar, Y = x + eps g + sqrt(2 eps) xi and W = z + eps h.  The products
// feed the same MODE_GEMM tile kernels as the logistic target.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/pmcmc_amd.h"
#include "gauss.hpp"

namespace pmcmc {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

static int64_t round_up_g(int64_t x, int64_t m) { return (x + m - 1) / m * m; }
static int grid_cap_g(int64_t n, int block = 256, int cap = 8192) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

template <typename Tin>
__global__ void gauss_prep_kernel(const Tin* L, const Tin* mu, int64_t D, int64_t Dp, float* L32,
                                  double* L64, float* mu32, double* mu64) {
    const int64_t n_all = Dp * Dp;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_all;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / Dp, c = i % Dp;  // row-major Dp x Dp
        const bool in = r < D && c < D;
        const double v = in ? (double)L[r + c * D] : 0.0;
        L32[i] = (float)v;
        if (in) L64[r + c * D] = v;
        if (c == 0 && r < D && mu) {
            mu32[r] = (float)mu[r];
            mu64[r] = (double)mu[r];
        }
    }
}

int gauss_data_create(GaussData* g, hipStream_t s, int64_t D, bool src_f64, const void* Lam,
                      const void* mean, std::string* msg) {
    g->D = D;
    g->Dp = round_up_g(D, 32);
    bool ok = hipMalloc(&g->L32, g->Dp * g->Dp * sizeof(float)) == hipSuccess &&
              hipMalloc(&g->L64, D * D * sizeof(double)) == hipSuccess;
    if (ok && mean)
        ok = hipMalloc(&g->mu32, D * sizeof(float)) == hipSuccess &&
             hipMalloc(&g->mu64, D * sizeof(double)) == hipSuccess;
    if (!ok) {
        *msg = "gaussian target: device allocation failed";
        return PMCMC_ERR_ALLOC;
    }
    const int grid = grid_cap_g(g->Dp * g->Dp);
    if (src_f64)
        hipLaunchKernelGGL((gauss_prep_kernel<double>), dim3(grid), dim3(256), 0, s, (const double*)Lam,
                           (const double*)mean, D, g->Dp, g->L32, g->L64, g->mu32, g->mu64);
    else
        hipLaunchKernelGGL((gauss_prep_kernel<float>), dim3(grid), dim3(256), 0, s, (const float*)Lam,
                           (const float*)mean, D, g->Dp, g->L32, g->L64, g->mu32, g->mu64);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        *msg = "gaussian target: preparation kernel failed";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}

void gauss_data_destroy(GaussData* g) {
    (void)hipFree(g->L32);
    (void)hipFree(g->L64);
    (void)hipFree(g->mu32);
    (void)hipFree(g->mu64);
    *g = GaussData{};
}

static constexpr int GCB = GAUSS_CB;  // t-columns per column block
static constexpr int GKC = 32;  // K slice

// Epilogue shared by both paths: per column t, given the products at
// (set 0) P0 = Lambda (v0 - mean) and (sets k) Pk = Lambda v_k.
//   pass 0: gx = -P0, W_k = z_k - eps P_k, Y = x + eps gx + sqrt(2 eps) xi, lpx
//   pass 1: gy = -P0, hw_k = -P_k, lpy

// ---------------------------------------------------------------------------
// FP32 MFMA pass over 64 columns per block: KS = 1 probe set (32 states and
// their 32 directions) or KS = 0 (64 states, no directions), four 16-column
// output tiles.  Wave w owns the 16-row tiles w, w + 4, ... of Lambda for all
// four column tiles, so every SIMD gets the same share of the FP32 MFMAs (and,
// with JBF, of the BF16 MFMAs of the direction tiles 2 and 3); the
// log-density of a column sums the four waves' row partials in a fixed order.
// ---------------------------------------------------------------------------
// ldL: Lambda's padded width (its row stride, a multiple of 32), at most
// 16 DPT; rows and columns of the DPT-wide tile beyond it are zero, so a
// narrower target can run in a wider launch class.
template <int DPT, int KS, bool JBF>
__device__ __forceinline__ void gauss_pass_body(const ColSrc<float>& src, const float* L32, const float* mu,
                                                int pass, float eps, const float* Xi, float* gout, float* hout,
                                                double* lpout, float* Yout, float* Wout, const int* skip,
                                                const int64_t cb, const int ldL) {
    constexpr int Dp = DPT * 16;
    constexpr int NC = 2 * GCB;
    constexpr int CPB = NC / (1 + KS);  // trajectory columns per block
    constexpr int SL = GKC + 4;   // Lambda slice [d][k]
    constexpr int SVc = GKC + 4;  // column slice [j][k]
    constexpr int NT = (DPT + 3) / 4;         // row tiles of a wave
    constexpr bool BFD = JBF && KS == 1;      // direction tiles on BF16 MFMA
    constexpr int NF = BFD ? 2 : 4;           // column tiles on FP32 MFMA
    __shared__ float Ls[Dp * SL];
    __shared__ float Vc[NC * SVc];
    __shared__ double lpred[4][NC];

    if (skip && *(volatile const int*)skip) return;
    if (src.col_lo && (cb + 1) * CPB <= (int64_t)*src.col_lo) return;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t D = src.D, Tn = src.Tn;

    v4f acc[NT][4];
#pragma unroll
    for (int q = 0; q < NT; ++q)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[q][ct] = (v4f){0.f, 0.f, 0.f, 0.f};

    // The k-slices are software-pipelined: slice s + 1 (Lambda rows as
    // float4, NC x 32 column values and the one mu entry a thread needs) is
    // loaded into registers while the MFMAs run on slice s in LDS.
    constexpr int NLV = (Dp * (GKC / 4) + 255) / 256;  // float4 of Lambda per thread
    constexpr int NCV = NC * GKC / 256;                 // column values per thread
    float4 lv[NLV];
    float cv[NCV];
    float muv = 0.f;
    auto fetch = [&](int k0) {
#pragma unroll
        for (int m = 0; m < NLV; ++m) {
            const int i = tid + 256 * m;
            const int d = i / (GKC / 4), q = i % (GKC / 4);
            lv[m] = (i < Dp * (GKC / 4) && d < ldL && k0 < ldL)
                        ? *reinterpret_cast<const float4*>(L32 + (size_t)d * ldL + k0 + 4 * q)
                        : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        const int64_t kk = k0 + (tid % GKC);  // the same row for all of this thread's values
#pragma unroll
        for (int m = 0; m < NCV; ++m) {
            const int j = (tid + 256 * m) / GKC;
            const int set = j / CPB;
            const int64_t t = cb * CPB + (j % CPB);
            float v = 0.f;
            if (t < Tn && kk < D) {
                if (set == 0) {
                    v = src.x0 ? (t == 0 ? src.x0[kk] : src.V[(size_t)(t - 1) * D + kk]) : src.V[(size_t)t * D + kk];
                } else {
                    const size_t gi = (size_t)t * D + kk;
                    v = src.Z ? src.Z[(size_t)(set - 1) * D * Tn + gi]
                              : rademacher(src.keys[set - 1], gi + (size_t)src.zoff);
                }
            }
            cv[m] = v;
        }
        muv = (mu && kk < D) ? mu[kk] : 0.f;
    };
    fetch(0);
    for (int k0 = 0; k0 < Dp; k0 += GKC) {
        __syncthreads();  // the previous slice's MFMAs are done with LDS
        // Lambda[:, k0:k0+32] (rows d, contiguous in k in the row-major copy)
#pragma unroll
        for (int m = 0; m < NLV; ++m) {
            const int i = tid + 256 * m;
            if (i < Dp * (GKC / 4)) {
                const int d = i / (GKC / 4), q = i % (GKC / 4);
                float* dst = &Ls[d * SL + 4 * q];
                dst[0] = lv[m].x;
                dst[1] = lv[m].y;
                dst[2] = lv[m].z;
                dst[3] = lv[m].w;
            }
        }
#pragma unroll
        for (int m = 0; m < NCV; ++m) {
            const int i = tid + 256 * m;
            const int j = i / GKC, k = i % GKC;
            // centred states; beyond D value and mu are both zero, and
            // columns past Tn stay zero
            const bool cen = j / CPB == 0 && cb * CPB + (j % CPB) < Tn;
            Vc[j * SVc + k] = cen ? cv[m] - muv : cv[m];
        }
        __syncthreads();
        if (k0 + GKC < Dp) fetch(k0 + GKC);
#pragma unroll
        for (int ks = 0; ks < GKC / 4; ++ks) {
            float bv[NF];
#pragma unroll
            for (int ct = 0; ct < NF; ++ct) bv[ct] = Vc[(16 * ct + li) * SVc + 4 * ks + lk];
#pragma unroll
            for (int q = 0; q < NT; ++q) {
                const int dt = wave + 4 * q;
                if (dt < DPT) {
                    const float av = Ls[(16 * dt + li) * SL + 4 * ks + lk];
#pragma unroll
                    for (int ct = 0; ct < NF; ++ct)
                        acc[q][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[ct], acc[q][ct], 0, 0, 0);
                }
            }
        }
        if (BFD) {
            v8bf b2, b3;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                b2[e] = (__bf16)Vc[(32 + li) * SVc + 8 * lk + e];
                b3[e] = (__bf16)Vc[(48 + li) * SVc + 8 * lk + e];
            }
#pragma unroll
            for (int q = 0; q < NT; ++q) {
                const int dt = wave + 4 * q;
                if (dt < DPT) {
                    v8bf av;
#pragma unroll
                    for (int e = 0; e < 8; ++e) av[e] = (__bf16)Ls[(16 * dt + li) * SL + 8 * lk + e];
                    acc[q][2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b2, acc[q][2], 0, 0, 0);
                    acc[q][3] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b3, acc[q][3], 0, 0, 0);
                }
            }
        }
    }

    // epilogue: lane holds P[d = 16 dt + 4 lk + r][column 16 ct + li] for its row tiles dt
    const double sq2e = sqrt(2.0 * (double)eps);
    double lpc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const int j = 16 * ct + li;
        const int set = j / CPB;
        const int64_t t = cb * CPB + (j % CPB);
        if (t >= Tn) continue;
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            const int dt = wave + 4 * q;
            if (dt >= DPT) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t d = 16 * dt + 4 * lk + r;
                if (d >= D) continue;
                const size_t gi = (size_t)t * D + d;
                const float g = -acc[q][ct][r];
                if (set == 0) {
                    const double x = src.get(0, t, d);
                    const double xm = x - (mu ? (double)mu[d] : 0.0);
                    lpc[ct] += 0.5 * xm * (double)g;
                    gout[gi] = g;
                    if (pass == 0) Yout[gi] = (float)(x + (double)eps * g + sq2e * (double)Xi[gi]);
                } else {
                    if (pass == 0)
                        Wout[gi] = (float)(src.get(1, t, d) + (double)eps * g);
                    else
                        hout[gi] = g;
                }
            }
        }
    }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        lpc[ct] += __shfl_xor(lpc[ct], 16, 64);
        lpc[ct] += __shfl_xor(lpc[ct], 32, 64);
        if (lk == 0) lpred[wave][16 * ct + li] = lpc[ct];
    }
    __syncthreads();
    if (tid < NC) {
        const int j = tid;
        const int64_t t = cb * CPB + (j % CPB);
        if (j / CPB == 0 && t < Tn) lpout[t] = (lpred[0][j] + lpred[1][j]) + (lpred[2][j] + lpred[3][j]);
    }
}

template <int DPT, int KS, bool JBF>
__global__ __launch_bounds__(256) void gauss_pass_mfma_kernel(ColSrc<float> src, const float* L32,
                                                              const float* mu, int pass, float eps,
                                                              const float* Xi, float* gout, float* hout,
                                                              double* lpout, float* Yout, float* Wout,
                                                              const int* skip) {
    gauss_pass_body<DPT, KS, JBF>(src, L32, mu, pass, eps, Xi, gout, hout, lpout, Yout, Wout, skip, blockIdx.x,
                                  DPT * 16);
}

// Batched over chains of padded width up to 16 DPT (each chain's own width
// in its args): work item blockIdx.x is (chain, column block) = items[blockIdx.x].
template <int DPT, bool JBF>
__global__ __launch_bounds__(256) void gauss_pass_mfma_batched(const GaussPassArgs* __restrict__ args, const int2* __restrict__ items) {
    const int2 w = items[blockIdx.x];
    const GaussPassArgs& a = args[w.x];
    if ((int64_t)w.y * GCB >= a.src.Tn) return;  // beyond the chain's current window
    gauss_pass_body<DPT, 1, JBF>(a.src, a.L32, a.mu, a.pass, a.eps, a.Xi, a.gout, a.hout, a.lpout, a.Yout, a.Wout,
                                 a.skip, w.y, a.ldL);
}

// ---------------------------------------------------------------------------
// Batched pass over one 64-row slab of Lambda: work item (chain, column
// block, slab) computes rows [64 slab, 64 slab + 64) of P = Lambda [v - mean |
// dir] for the block's 32 states and 32 directions, K = the chain's own
// padded width (any width in one launch).  Wave w owns row tile 4 slab + w
// for all four column tiles.  Splitting the rows over workgroups keeps a
// column block's latency at a quarter of the whole-column pass's when few
// chains are live (a chain's window then has only a handful of blocks); the
// log-density of a column is written per slab (lpout[slab * lp_stride + t])
// and summed over the slabs, in slab order, by its readers.
// ---------------------------------------------------------------------------
template <bool JBF>
__device__ __forceinline__ void gauss_slab_body(const GaussPassArgs& a, const int64_t cb, const int slab) {
    // Lambda's slab rows stay in registers for the whole K: lane (li, lk) of
    // wave w holds row r0 + 16 w + li at k = 16 j + 4 lk + e (j < 16, e < 4),
    // loaded as 16 float4 all in flight at once.  The block's 32 states
    // (minus the mean) and 32 directions sit in LDS, k contiguous per column
    // (stride SV = 4 mod 64 words, so the 16-byte B reads of a 16-lane group
    // hit distinct banks); both are loaded as contiguous float4 spans in the
    // same round trip as A.  The 16x16x4 MFMA step (j, e) takes k = 16 j +
    // 4 lk + e from every lane; even and odd j accumulate separately (four
    // independent chains for the states).  The epilogue goes through LDS so
    // that every output and operand access is a coalesced run of rows.
    constexpr int CPB = GCB;  // 32 states + 32 directions per block
    constexpr int SV = 260;
    constexpr int SP = 65;    // stride of the staged products [column][row]
    __shared__ __align__(16) float Vs[CPB * SV];
    __shared__ __align__(16) float Vd[CPB * SV];
    __shared__ double lpw[CPB];

    const ColSrc<float>& src = a.src;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t D = src.D, Tn = src.Tn;
    const int ldL = a.ldL;    // a multiple of 32, <= 256
    const int nj = ldL / 16;  // float4 columns of A per lane
    const int r0 = 64 * slab;
    const int row = r0 + 16 * wave + li;
    const float* mu = a.mu;
    const int64_t t0 = cb * CPB;
    const int nc = (int)(Tn - t0 < CPB ? Tn - t0 : CPB);  // >= 1 (items beyond Tn exit before)

    pmcmc_f4v A[16];
    {
        // every load in flight at once: the K steps j < nj only (the products
        // read no others; nj is even, so the BF16 pairs stay inside).  Rows
        // past ldL read row 0: they only make product rows past D, which are
        // never stored.
        const int rowc = row < ldL ? row : 0;
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (j < nj) A[j] = gld4(a.L32 + (size_t)rowc * ldL + 16 * j + 4 * lk);
    }
    // the epilogue's operands: the states' x and tape values of the wave's
    // columns (lane = row of the slab), in flight with Lambda (clamped
    // addresses; values of dead lanes unused)
    const int Di = (int)D;
    const int d = r0 + lane;
    const bool drow = d < Di;
    const int dc = drow ? d : 0;
    const bool p0 = a.pass == 0;
    float xv[8], xiv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int jc = wave + 4 * q;
        const int t = (int)t0 + (jc < nc ? jc : 0);
        const float* xp = src.x0 ? (t == 0 ? src.x0 + dc : src.V + ((size_t)(t - 1) * Di + dc))
                                 : src.V + ((size_t)t * Di + dc);
        xv[q] = *gp(xp);
        xiv[q] = gp(p0 ? a.Xi : src.V)[(size_t)t * Di + dc];
    }
    // loaded unconditionally (any valid row when there is no mean) and
    // selected and converted in the epilogue: doing either here would wait
    // for every load issued so far
    const float muf = gp(mu ? mu : src.V)[dc];
    // states: column t is V[:, t] (or Xbar: x0, then V[:, t - 1]); given
    // directions: Z's columns.  Both spans are issued with Lambda's loads and
    // written to LDS after the probe hashing, one memory round trip in all.
    SpanRegs sps, spz;
    const bool xfirst = src.x0 && t0 == 0;
    float x0v = 0.f;
    if (xfirst) {
        x0v = gp(src.x0)[tid < Di ? tid : 0];
        span_issue(sps, src.V, Di, nc - 1, tid);
    } else {
        span_issue(sps, src.V + (size_t)(src.x0 ? t0 - 1 : t0) * D, Di, nc, tid);
    }
    if (src.Z) span_issue(spz, src.Z + (size_t)t0 * D, Di, nc, tid);
    const float muk = (mu && tid < Di) ? gp(mu)[tid] : 0.f;
    // directions without a matrix: the probe of the block's absolute columns
    if (!src.Z && tid < Di)
        for (int col = 0; col < nc; ++col)
            Vd[col * SV + tid] = rademacher(src.keys[0], (size_t)(t0 + col) * D + tid + (size_t)src.zoff);
    if (xfirst) {
        if (tid < Di) Vs[tid] = x0v;
        span_store<SV>(Vs + SV, sps, Di, tid);
    } else {
        span_store<SV>(Vs, sps, Di, tid);
    }
    if (src.Z) span_store<SV>(Vd, spz, Di, tid);
    // zeros where the MFMAs read beyond the data: rows k in [D, ldL), columns >= nc
    if (tid < ldL)
        for (int col = 0; col < CPB; ++col)
            if (tid >= Di || col >= nc) {
                Vs[col * SV + tid] = 0.f;
                Vd[col * SV + tid] = 0.f;
            }
    __syncthreads();
    if (mu) {  // the states, centred (each thread its own row k = tid)
        if (tid < Di)
            for (int col = 0; col < nc; ++col) Vs[col * SV + tid] -= muk;
        __syncthreads();
    }

    v4f acc[4][2];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) acc[ct][0] = acc[ct][1] = (v4f){0.f, 0.f, 0.f, 0.f};
    // states: column tiles 0, 1 on FP32 MFMA (the loops over A are unrolled,
    // so that A stays in registers)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if (j < nj) {
            const float av[4] = {A[j].x, A[j].y, A[j].z, A[j].w};
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) {
                const float4 b = *reinterpret_cast<const float4*>(&Vs[(16 * ct + li) * SV + 16 * j + 4 * lk]);
                const float bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    acc[ct][j & 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[e], bv[e], acc[ct][j & 1], 0, 0, 0);
            }
        }
    }
    // directions: column tiles 2, 3 (BF16 MFMA, 32 k per step: the lane's
    // eight values of two float4 columns, in the same k order for A and B)
    if (JBF) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (2 * q >= nj) continue;
            v8bf av, b0, b1;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const pmcmc_f4v f = A[2 * q + h];
                const float4 c0 = *reinterpret_cast<const float4*>(&Vd[li * SV + 16 * (2 * q + h) + 4 * lk]);
                const float4 c1 = *reinterpret_cast<const float4*>(&Vd[(16 + li) * SV + 16 * (2 * q + h) + 4 * lk]);
                const float fv[4] = {f.x, f.y, f.z, f.w};
                const float v0[4] = {c0.x, c0.y, c0.z, c0.w};
                const float v1[4] = {c1.x, c1.y, c1.z, c1.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    av[4 * h + e] = (__bf16)fv[e];
                    b0[4 * h + e] = (__bf16)v0[e];
                    b1[4 * h + e] = (__bf16)v1[e];
                }
            }
            acc[2][q & 1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b0, acc[2][q & 1], 0, 0, 0);
            acc[3][q & 1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b1, acc[3][q & 1], 0, 0, 0);
        }
    } else {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            if (j >= nj) continue;
            const float av[4] = {A[j].x, A[j].y, A[j].z, A[j].w};
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) {
                const float4 b = *reinterpret_cast<const float4*>(&Vd[(16 * ct + li) * SV + 16 * j + 4 * lk]);
                const float bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    acc[2 + ct][j & 1] =
                        __builtin_amdgcn_mfma_f32_16x16x4f32(av[e], bv[e], acc[2 + ct][j & 1], 0, 0, 0);
            }
        }
    }
    __syncthreads();  // the MFMAs' reads of Vs are done: Vs now stages the products
    // lane holds P[row 16 wave + 4 lk + r][column 16 ct + li] (columns 0-31:
    // the states, 32-63: their directions); staged as Pt[column][row]
    float* Pt = Vs;
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const v4f p = acc[ct][0] + acc[ct][1];
#pragma unroll
        for (int r = 0; r < 4; ++r) Pt[(16 * ct + li) * SP + 16 * wave + 4 * lk + r] = p[r];
    }
    __syncthreads();

    // outputs, one column per wave and round, lanes over the slab's 64 rows:
    // the wave's columns j = wave + 4 q (q < 8: states, q >= 8: their
    // directions); the states' x and tape values were loaded with Lambda
    const double eps = (double)a.eps;
    const double sq2e = sqrt(2.0 * eps);
    const double mud = mu ? (double)muf : 0.0;
    // the eight state columns' log-density sums over the slab's rows run as
    // one interleaved butterfly after the loop (each column's own order of
    // additions unchanged), not as eight chains of lane exchanges in turn
    double lpq[8];
    // every LDS operand of the loop read first (one round trip; the empty asm
    // keeps the compiler from sinking each read into its column's branch)
    float gq[16], wq[8];
#pragma unroll
    for (int q = 0; q < 16; ++q) gq[q] = -Pt[(wave + 4 * q) * SP + lane];
#pragma unroll
    for (int q = 0; q < 8; ++q) wq[q] = Vd[(wave + 4 * q) * SV + dc];
    asm volatile("" ::: "memory");
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int j = wave + 4 * q;  // column of the block: 0-31 states, 32-63 directions
        const int jc = j & (CPB - 1);
        const int t = (int)t0 + jc;
        const bool live = jc < nc;
        const float g = gq[q];
        const size_t gi = (size_t)t * Di + d;
        if (q < 8) {
            double lpp = 0.0;
            if (live && drow) {
                const double x = (double)xv[q];
                lpp = 0.5 * (x - mud) * (double)g;
                gp(a.gout)[gi] = g;
                if (p0) gp(a.Yout)[gi] = (float)(x + eps * g + sq2e * (double)xiv[q]);
            }
            lpq[q] = lpp;
        } else if (live && drow) {
            if (p0)
                gp(a.Wout)[gi] = (float)((double)wq[q - 8] + eps * g);
            else
                gp(a.hout)[gi] = g;
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
        for (int q = 0; q < 8; ++q) lpq[q] += __shfl_xor(lpq[q], m, 64);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 8; ++q) lpw[wave + 4 * q] = lpq[q];
    __syncthreads();
    if (tid < nc) gp(a.lpout)[(size_t)slab * a.lp_stride + t0 + tid] = lpw[tid];
    if (a.flops && tid == 0) {
        const int64_t rows = D - r0 < 64 ? D - r0 : 64;
        if (rows > 0) atomicAdd(a.flops, (unsigned long long)(2 * rows * D * 2 * nc));
    }
}

// work item blockIdx.x is (chain, column block, slab) = items[blockIdx.x]
template <bool JBF>
__global__ __launch_bounds__(256) void gauss_slab_batched(const GaussPassArgs* __restrict__ args, const int4* __restrict__ items) {
    const int4 w = items[blockIdx.x];
    const GaussPassArgs& a = args[w.x];
    if (a.skip && *(volatile const int*)a.skip) return;
    if ((int64_t)w.y * GCB >= a.src.Tn) return;  // beyond the chain's current window
    gauss_slab_body<JBF>(a, w.y, w.z);
}

void launch_gauss_slab_batched(bool jbf, const GaussPassArgs* args, const int4* items, int nitems, hipStream_t s) {
    if (nitems <= 0) return;
    if (jbf)
        hipLaunchKernelGGL((gauss_slab_batched<true>), dim3((unsigned)nitems), dim3(256), 0, s, args, items);
    else
        hipLaunchKernelGGL((gauss_slab_batched<false>), dim3((unsigned)nitems), dim3(256), 0, s, args, items);
}

// ---------------------------------------------------------------------------
// plain path (FP64, or several probes): thread per (set, t, d) output
// ---------------------------------------------------------------------------
template <typename T>
__global__ void gauss_plain_kernel(ColSrc<T> src, const double* L64, const double* mu, int pass,
                                   double eps, const T* Xi, T* gout, T* hout, T* Yout, T* Wout,
                                   const int* skip) {
    if (skip && *(volatile const int*)skip) return;
    const int64_t D = src.D, Tn = src.Tn;
    const int K = src.K;
    const int64_t total = (int64_t)(1 + K) * Tn * D;
    const double sq2e = sqrt(2.0 * eps);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = i % D, st = i / D, t = st % Tn;
        const int set = (int)(st / Tn);
        double p = 0.0;
        for (int64_t k = 0; k < D; ++k) {
            double v = src.get(set, t, k);
            if (set == 0 && mu) v -= mu[k];
            p += L64[d + k * D] * v;
        }
        const double g = -p;
        const size_t gi = (size_t)t * D + d;
        if (set == 0) {
            gout[gi] = (T)g;
            if (pass == 0) Yout[gi] = (T)(src.get(0, t, d) + eps * g + sq2e * (double)Xi[gi]);
        } else {
            const size_t ki = (size_t)(set - 1) * D * Tn + gi;
            if (pass == 0)
                Wout[ki] = (T)(src.get(set, t, d) + eps * g);
            else
                hout[ki] = (T)g;
        }
    }
}

// logp of the set-0 columns from the stored gradient: (v - mean).g / 2
template <typename T>
__global__ __launch_bounds__(256) void gauss_lp_kernel(ColSrc<T> src, const double* mu, const T* g,
                                                       double* lpout, const int* skip) {
    if (skip && *(volatile const int*)skip) return;
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= src.Tn) return;
    double s = 0.0;
    for (int64_t d = lane; d < src.D; d += 64) {
        const double xm = src.get(0, t, d) - (mu ? mu[d] : 0.0);
        s += 0.5 * xm * (double)g[(size_t)t * src.D + d];
    }
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if (lane == 0) lpout[t] = s;
}

template <int DPT>
static void launch_gauss_mfma(const ColSrc<float>& src, const GaussData* g, int pass, double eps,
                              const float* Xi, float* gout, float* hout, double* lpout, float* Yout,
                              float* Wout, const int* skip, bool jbf, hipStream_t s) {
    if (src.K == 0) {
        const dim3 grid((unsigned)((src.Tn + 2 * GCB - 1) / (2 * GCB)));
        hipLaunchKernelGGL((gauss_pass_mfma_kernel<DPT, 0, false>), grid, dim3(256), 0, s, src, g->L32, g->mu32,
                           pass, (float)eps, Xi, gout, hout, lpout, Yout, Wout, skip);
        return;
    }
    const dim3 grid((unsigned)((src.Tn + GCB - 1) / GCB));
    if (jbf)
        hipLaunchKernelGGL((gauss_pass_mfma_kernel<DPT, 1, true>), grid, dim3(256), 0, s, src, g->L32, g->mu32,
                           pass, (float)eps, Xi, gout, hout, lpout, Yout, Wout, skip);
    else
        hipLaunchKernelGGL((gauss_pass_mfma_kernel<DPT, 1, false>), grid, dim3(256), 0, s, src, g->L32, g->mu32,
                           pass, (float)eps, Xi, gout, hout, lpout, Yout, Wout, skip);
}

template <typename T>
static int gauss_pass(const GaussData* g, hipStream_t s, const ColSrc<T>& src, int pass, double eps,
                      const T* Xi, T* gout, T* hout, double* lpout, T* Yout, T* Wout, const int* skip,
                      Prof* prof, int iter, bool jbf, std::string* msg) {
    const int pe = prof ? prof->begin(KT_GLM_PASS, iter, s) : -1;
    if (sizeof(T) == 4 && src.K <= 1 && g->Dp <= 256) {
        const ColSrc<float>& fs = reinterpret_cast<const ColSrc<float>&>(src);
        float* go = (float*)gout;
        float* ho = (float*)hout;
        float* yo = (float*)Yout;
        float* wo = (float*)Wout;
        const float* xi = (const float*)Xi;
        switch (g->Dp / 16) {
#define PMCMC_GDPT(n) \
    case n: launch_gauss_mfma<n>(fs, g, pass, eps, xi, go, ho, lpout, yo, wo, skip, jbf, s); break;
            PMCMC_GDPT(2) PMCMC_GDPT(4) PMCMC_GDPT(6) PMCMC_GDPT(8) PMCMC_GDPT(10) PMCMC_GDPT(12)
            PMCMC_GDPT(14) PMCMC_GDPT(16)
#undef PMCMC_GDPT
            default:
                *msg = "gaussian target: unsupported D";
                return PMCMC_ERR_UNSUPPORTED;
        }
    } else {
        hipLaunchKernelGGL((gauss_plain_kernel<T>), dim3(grid_cap_g((1 + src.K) * src.Tn * src.D)), dim3(256), 0, s,
                           src, g->L64, g->mu64, pass, eps, Xi, gout, hout, Yout, Wout, skip);
        hipLaunchKernelGGL((gauss_lp_kernel<T>), dim3((unsigned)((src.Tn + 3) / 4)), dim3(256), 0, s, src, g->mu64,
                           (const T*)gout, lpout, skip);
    }
    if (prof) prof->end(pe, s);
    if (hipGetLastError() != hipSuccess) {
        *msg = "gaussian pass: kernel launch failed";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}

template <typename T>
int gauss_newton_products(GlmWork* w, const GaussData* g, hipStream_t s, SyncState* sync,
                          int precision, double eps, const T* x0, const T* S_in, const T* Xi,
                          uint64_t seed, int iter, uint32_t probe_iter, int probes, int64_t D, int64_t Tn,
                          Prof* prof, std::string* msg, const int* col_lo) {
    const bool jbf = precision == PMCMC_PREC_BF16_JAC;
    const int* skip = sync ? &sync->converged : nullptr;
    ColSrc<T> a{};
    a.V = S_in;
    a.x0 = x0;
    a.Z = nullptr;
    a.K = probes;
    a.D = D;
    a.Tn = Tn;
    a.col_lo = col_lo;
    for (int k = 0; k < probes && k < 4; ++k) a.keys[k] = probe_key(seed, probe_iter, (uint32_t)k);
    int rc = gauss_pass<T>(g, s, a, 0, eps, Xi, (T*)w->gx, (T*)w->hw, w->lpx, (T*)w->Y, (T*)w->W, skip, prof,
                           iter, jbf, msg);
    if (rc != PMCMC_OK) return rc;
    ColSrc<T> b{};
    b.V = (const T*)w->Y;
    b.x0 = nullptr;
    b.Z = (const T*)w->W;
    b.K = probes;
    b.D = D;
    b.Tn = Tn;
    b.col_lo = col_lo;
    return gauss_pass<T>(g, s, b, 1, eps, Xi, (T*)w->gy, (T*)w->hw, w->lpy, nullptr, nullptr, skip, prof, iter,
                         jbf, msg);
}
template int gauss_newton_products<float>(GlmWork*, const GaussData*, hipStream_t, SyncState*, int, double,
                                          const float*, const float*, const float*, uint64_t, int, uint32_t, int,
                                          int64_t, int64_t, Prof*, std::string*, const int*);
template int gauss_newton_products<double>(GlmWork*, const GaussData*, hipStream_t, SyncState*, int, double,
                                           const double*, const double*, const double*, uint64_t, int, uint32_t,
                                           int, int64_t, int64_t, Prof*, std::string*, const int*);

template <int DPT>
static void launch_gauss_batched_dpt(bool jbf, const GaussPassArgs* args, const int2* items, int nitems,
                                     hipStream_t s) {
    if (jbf)
        hipLaunchKernelGGL((gauss_pass_mfma_batched<DPT, true>), dim3((unsigned)nitems), dim3(256), 0, s, args,
                           items);
    else
        hipLaunchKernelGGL((gauss_pass_mfma_batched<DPT, false>), dim3((unsigned)nitems), dim3(256), 0, s, args,
                           items);
}

void launch_gauss_pass_batched(int dpt, bool jbf, const GaussPassArgs* args, const int2* items, int nitems,
                               hipStream_t s) {
    if (nitems <= 0) return;
    switch (dpt) {
#define PMCMC_BGDPT(n) \
    case n: launch_gauss_batched_dpt<n>(jbf, args, items, nitems, s); break;
        PMCMC_BGDPT(2) PMCMC_BGDPT(4) PMCMC_BGDPT(6) PMCMC_BGDPT(8) PMCMC_BGDPT(10) PMCMC_BGDPT(12)
        PMCMC_BGDPT(14) PMCMC_BGDPT(16)
#undef PMCMC_BGDPT
        default: break;
    }
}

template <typename T>
int gauss_dir_products(GlmWork* w, const GaussData* g, hipStream_t s, double eps, const T* V, const T* x0,
                       const T* Z, uint32_t key, const T* Xi, int64_t D, int64_t Tn, const int* skip, Prof* prof,
                       int iter, std::string* msg) {
    ColSrc<T> a{};
    a.V = V;
    a.x0 = x0;
    a.Z = Z;
    a.keys[0] = key;
    a.K = 1;
    a.D = D;
    a.Tn = Tn;
    return gauss_pass<T>(g, s, a, 0, eps, Xi, (T*)w->gx, (T*)w->hw, w->lpx, (T*)w->Y, (T*)w->W, skip, prof, iter,
                         false, msg);
}
template int gauss_dir_products<float>(GlmWork*, const GaussData*, hipStream_t, double, const float*, const float*,
                                       const float*, uint32_t, const float*, int64_t, int64_t, const int*, Prof*, int,
                                       std::string*);
template int gauss_dir_products<double>(GlmWork*, const GaussData*, hipStream_t, double, const double*,
                                        const double*, const double*, uint32_t, const double*, int64_t, int64_t,
                                        const int*, Prof*, int, std::string*);

// As glm_hvp_products: hout = H(V) Z = -Lambda Z; gradient and log-density
// of the states to gscr / lpscr.
template <typename T>
int gauss_hvp_products(const GaussData* g, hipStream_t s, const T* V, const T* x0, const T* Z, T* gscr,
                       double* lpscr, T* hout, int64_t D, int64_t Tn, const int* skip, Prof* prof, int iter, bool jbf,
                       std::string* msg) {
    ColSrc<T> a{};
    a.V = V;
    a.x0 = x0;
    a.Z = Z;
    a.K = 1;
    a.D = D;
    a.Tn = Tn;
    return gauss_pass<T>(g, s, a, 1, 0.0, nullptr, gscr, hout, lpscr, nullptr, nullptr, skip, prof, iter, jbf, msg);
}
template int gauss_hvp_products<float>(const GaussData*, hipStream_t, const float*, const float*, const float*, float*,
                                       double*, float*, int64_t, int64_t, const int*, Prof*, int, bool, std::string*);
template int gauss_hvp_products<double>(const GaussData*, hipStream_t, const double*, const double*, const double*,
                                        double*, double*, double*, int64_t, int64_t, const int*, Prof*, int, bool,
                                        std::string*);

template <typename T>
static int gauss_fwd_jvp_w(const GaussData* g, GlmWork& w, hipStream_t s, double eps, const T* X, const T* Xi,
                           const T* U, const T* Z, T* FT, T* Jt, uint8_t* acc, double* la, int64_t D, int64_t Tn,
                           bool jbf, std::string* msg) {
    ColSrc<T> a{};
    a.V = X;
    a.Z = Z;
    a.K = 1;
    a.D = D;
    a.Tn = Tn;
    int rc = gauss_pass<T>(g, s, a, 0, eps, Xi, (T*)w.gx, (T*)w.hw, w.lpx, (T*)w.Y, (T*)w.W, nullptr, nullptr, 0,
                           jbf, msg);
    if (rc == PMCMC_OK) {
        ColSrc<T> b = a;
        b.V = (const T*)w.Y;
        b.Z = (const T*)w.W;
        rc = gauss_pass<T>(g, s, b, 1, eps, Xi, (T*)w.gy, (T*)w.hw, w.lpy, nullptr, nullptr, nullptr, nullptr, 0,
                           jbf, msg);
    }
    if (rc == PMCMC_OK) {
        launch_fwd_jvp_finish<T>(s, X, Z, U, &w, eps, D, Tn, FT, Jt, acc, la);
        if (hipGetLastError() != hipSuccess) {
            *msg = "gaussian finish kernel failed";
            rc = PMCMC_ERR_HIP;
        }
    }
    return rc;
}

static int gauss_work(GlmWork* w, const GaussData* g, hipStream_t s, int64_t D, int64_t Tn, bool f64,
                      std::string* msg) {
    GlmData dummy{};
    dummy.N = 1;
    dummy.Dp = g->Dp;
    return glm_work_ensure(w, &dummy, s, D, Tn, 1, f64, msg);
}

template <typename T>
static int gauss_fwd_jvp_t(const GaussData* g, hipStream_t s, double eps, const T* X, const T* Xi, const T* U,
                           const T* Z, T* FT, T* Jt, uint8_t* acc, double* la, int64_t D, int64_t Tn, bool jbf,
                           std::string* msg) {
    GlmWork w{};
    int rc = gauss_work(&w, g, s, D, Tn, sizeof(T) == 8, msg);
    if (rc == PMCMC_OK) rc = gauss_fwd_jvp_w<T>(g, w, s, eps, X, Xi, U, Z, FT, Jt, acc, la, D, Tn, jbf, msg);
    (void)hipStreamSynchronize(s);
    glm_work_destroy(&w);
    return rc;
}

// K sequential MALA steps of N chains (as glm_mala_run)
template <typename T>
static int gauss_mala_run_t(const GaussData* g, hipStream_t s, double eps, const T* X0, const T* Xi, const T* U,
                            int64_t K, int64_t D, int64_t N, T* S_out, uint8_t* acc, const T* Z, T* Jw, bool jbf,
                            std::string* msg) {
    GlmWork w{};
    int rc = gauss_work(&w, g, s, D, N, sizeof(T) == 8, msg);
    for (int64_t k = 0; k < K && rc == PMCMC_OK; ++k) {
        const T* X = k == 0 ? X0 : S_out + (size_t)(k - 1) * D * N;
        rc = gauss_fwd_jvp_w<T>(g, w, s, eps, X, Xi + (size_t)k * D * N, U + (size_t)k * N, Z,
                                S_out + (size_t)k * D * N, Jw, acc + (size_t)k * N, nullptr, D, N, jbf, msg);
    }
    (void)hipStreamSynchronize(s);
    glm_work_destroy(&w);
    return rc;
}

int gauss_mala_run(const GaussData* g, hipStream_t s, int prec, double eps, const void* X0, const void* Xi,
                   const void* U, int64_t K, int64_t D, int64_t N, void* S_out, uint8_t* acc, const void* Z, void* Jw,
                   std::string* msg) {
    if (prec == PMCMC_PREC_F64)
        return gauss_mala_run_t<double>(g, s, eps, (const double*)X0, (const double*)Xi, (const double*)U, K, D, N,
                                        (double*)S_out, acc, (const double*)Z, (double*)Jw, false, msg);
    return gauss_mala_run_t<float>(g, s, eps, (const float*)X0, (const float*)Xi, (const float*)U, K, D, N,
                                   (float*)S_out, acc, (const float*)Z, (float*)Jw, prec == PMCMC_PREC_BF16_JAC, msg);
}

int gauss_fwd_and_jvp(const GaussData* g, hipStream_t s, int prec, double eps, const void* X,
                      const void* Xi, const void* U, const void* Z, void* FT, void* Jt,
                      uint8_t* accepted, double* logalpha, int64_t D, int64_t T, std::string* msg) {
    if (prec == PMCMC_PREC_F64)
        return gauss_fwd_jvp_t<double>(g, s, eps, (const double*)X, (const double*)Xi, (const double*)U,
                                       (const double*)Z, (double*)FT, (double*)Jt, accepted, logalpha, D, T,
                                       false, msg);
    return gauss_fwd_jvp_t<float>(g, s, eps, (const float*)X, (const float*)Xi, (const float*)U, (const float*)Z,
                                  (float*)FT, (float*)Jt, accepted, logalpha, D, T,
                                  prec == PMCMC_PREC_BF16_JAC, msg);
}

// logp in double: one workgroup per column
template <typename T>
__global__ __launch_bounds__(256) void gauss_logp_kernel(const T* S, const double* L64, const double* mu,
                                                         int64_t D, double* out) {
    __shared__ double xm[1024];
    __shared__ double red[4];
    const int64_t t = blockIdx.x;
    for (int64_t d = threadIdx.x; d < D; d += 256) xm[d] = (double)S[(size_t)t * D + d] - (mu ? mu[d] : 0.0);
    __syncthreads();
    double acc = 0.0;
    for (int64_t d = threadIdx.x; d < D; d += 256) {
        double p = 0.0;
        for (int64_t k = 0; k < D; ++k) p += L64[d + k * D] * xm[k];
        acc += -0.5 * xm[d] * p;
    }
    for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[t] = red[0] + red[1] + red[2] + red[3];
}

int gauss_logdensity(const GaussData* g, hipStream_t s, int prec, const void* S, double* out,
                     int64_t T, std::string* msg) {
    if (g->D > 1024) {
        *msg = "gaussian log-density: D > 1024";
        return PMCMC_ERR_UNSUPPORTED;
    }
    if (prec == PMCMC_PREC_F64)
        hipLaunchKernelGGL((gauss_logp_kernel<double>), dim3((unsigned)T), dim3(256), 0, s, (const double*)S, g->L64,
                           g->mu64, g->D, out);
    else
        hipLaunchKernelGGL((gauss_logp_kernel<float>), dim3((unsigned)T), dim3(256), 0, s, (const float*)S, g->L64,
                           g->mu64, g->D, out);
    if (hipGetLastError() != hipSuccess) {
        *msg = "gaussian log-density kernel failed";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}

}  // namespace pmcmc
