// Logistic-regression target: the per-iteration gradient / HVP products of
// the DEER Newton loop as fused GEMM -> sigmoid -> GEMM^T passes.
//
// For a set of columns V (the states x_t and the probe directions z_t of the
// same t), one pass computes, without materialising any N x T matrix in HBM,
//   l = Xd x_t, q_k = Xd z_{k,t}               (GEMM-A, N x cols)
//   r = y - sigmoid(l), u_k = sigmoid'(l) q_k  (elementwise, on chip)
//   G = Xd^T r, H_k = Xd^T u_k                 (GEMM-B, D x cols)
//   ll_t = sum_n y log s(l) + (1-y) log s(-l)
// (benchmarks/ParallelMCMCBenchmarks/src/models/bayes_logreg.jl:118-203 make
// the same products with 12 separate GEMMs and N x T temporaries).
// Two passes per Newton iteration: at Xbar (with the probes Z) and at the
// proposal Y (with W = Z + eps H(Xbar) Z).
//
// FP32 path: v_mfma_f32_16x16x4_f32 tiles; the design matrix is streamed
// through LDS in 32-row chunks and used by both GEMMs; the column block stays
// in LDS; GEMM-B accumulates in registers across the N range of the
// workgroup; N is split over blockIdx.y and the splits are summed by
// glm_reduce_kernel in a fixed order (deterministic).
// FP64 path: plain kernels (small problems and parity runs).
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/pmcmc_amd.h"
#include "glm.hpp"

namespace pmcmc {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

static int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

static int grid_cap(int64_t n, int block = 256, int cap = 8192) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

__device__ inline float softplusf(float x) { return fmaxf(x, 0.f) + log1pf(expf(-fabsf(x))); }
__device__ inline double softplusd(double x) { return fmax(x, 0.0) + log1p(exp(-fabs(x))); }
__device__ inline float sigmoidf_(float x) {
    if (x >= 0.f) return 1.f / (1.f + expf(-x));
    const float e = expf(x);
    return e / (1.f + e);
}
__device__ inline double sigmoidd_(double x) {
    if (x >= 0.0) return 1.0 / (1.0 + exp(-x));
    const double e = exp(x);
    return e / (1.0 + e);
}

// ---------------------------------------------------------------------------
// data preparation
// ---------------------------------------------------------------------------
template <typename Tin>
__global__ void glm_prep_kernel(const Tin* Xcm, const Tin* y, int64_t N, int64_t D, int64_t Np,
                                int64_t Dp, float* Xrm, float* y32, double* X64, double* y64) {
    const int64_t n_all = Np * Dp;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_all;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = i / Dp, d = i % Dp;
        const bool in = n < N && d < D;
        const double v = in ? (double)Xcm[n + d * N] : 0.0;
        Xrm[i] = (float)v;
        if (in) X64[n + d * N] = v;
        if (d == 0) {
            const double yv = n < N ? (double)y[n] : 0.0;
            y32[n] = (float)yv;
            if (n < N) y64[n] = yv;
        }
    }
}

// BF16 copies of the FP32 design (row-major Np x Dq and transposed Dp x Np),
// rounded to nearest even from the FP32 values, zero padded.
__global__ void glm_prep_bf16_kernel(const float* Xrm, int64_t Np, int64_t Dp, int64_t Dq, uint16_t* Xb,
                                     uint16_t* Xtb) {
    const int64_t n_all = Np * Dq;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_all;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = i / Dq, d = i % Dq;
        const __bf16 v = (__bf16)(d < Dp ? Xrm[n * Dp + d] : 0.f);
        const uint16_t b = __builtin_bit_cast(uint16_t, v);
        Xb[i] = b;
        if (d < Dp) Xtb[d * Np + n] = b;
    }
}

int glm_data_create(GlmData* g, hipStream_t s, int64_t N, int64_t D, bool src_f64, const void* Xd,
                    const void* y, std::string* msg) {
    g->N = N;
    g->D = D;
    g->Dp = round_up(D, 16);
    g->Dq = round_up(D, 32);
    const int64_t Np = round_up(N, 32);
    g->Np = Np;
    bool ok = hipMalloc(&g->Xrm, Np * g->Dp * sizeof(float)) == hipSuccess &&
              hipMalloc(&g->Xb, Np * g->Dq * sizeof(uint16_t)) == hipSuccess &&
              hipMalloc(&g->Xtb, g->Dp * Np * sizeof(uint16_t)) == hipSuccess &&
              hipMalloc(&g->y, Np * sizeof(float)) == hipSuccess &&
              hipMalloc(&g->Xd64, N * D * sizeof(double)) == hipSuccess &&
              hipMalloc(&g->y64, N * sizeof(double)) == hipSuccess;
    if (!ok) {
        *msg = "logistic target: device allocation failed";
        return PMCMC_ERR_ALLOC;
    }
    const int grid = grid_cap(Np * g->Dp);
    if (src_f64)
        hipLaunchKernelGGL((glm_prep_kernel<double>), dim3(grid), dim3(256), 0, s, (const double*)Xd,
                           (const double*)y, N, D, Np, g->Dp, g->Xrm, g->y, g->Xd64, g->y64);
    else
        hipLaunchKernelGGL((glm_prep_kernel<float>), dim3(grid), dim3(256), 0, s, (const float*)Xd,
                           (const float*)y, N, D, Np, g->Dp, g->Xrm, g->y, g->Xd64, g->y64);
    hipLaunchKernelGGL(glm_prep_bf16_kernel, dim3(grid_cap(Np * g->Dq)), dim3(256), 0, s, g->Xrm, Np, g->Dp,
                       g->Dq, g->Xb, g->Xtb);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        *msg = "logistic target: preparation kernel failed";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}

void glm_data_destroy(GlmData* g) {
    (void)hipFree(g->Xrm);
    (void)hipFree(g->Xb);
    (void)hipFree(g->Xtb);
    (void)hipFree(g->Xcm);
    (void)hipFree(g->Xd64);
    (void)hipFree(g->y);
    (void)hipFree(g->y64);
    (void)hipFree(g->P64);
    *g = GlmData{};
}

template <typename S>
__global__ void glm_prior_kernel(const S* P, int64_t n, double* out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (double)P[i];
}

int glm_data_set_prior(GlmData* g, hipStream_t s, bool src_f64, const void* P, std::string* msg) {
    const int64_t n = g->D * g->D;
    if (!g->P64 && hipMalloc(&g->P64, (size_t)n * sizeof(double)) != hipSuccess) {
        *msg = "logistic target: prior allocation failed";
        return PMCMC_ERR_ALLOC;
    }
    if (src_f64)
        hipLaunchKernelGGL((glm_prior_kernel<double>), dim3(grid_cap(n)), dim3(256), 0, s, (const double*)P, n,
                           g->P64);
    else
        hipLaunchKernelGGL((glm_prior_kernel<float>), dim3(grid_cap(n)), dim3(256), 0, s, (const float*)P, n,
                           g->P64);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        *msg = "logistic target: prior copy failed";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}

// ---------------------------------------------------------------------------
// work buffers
// ---------------------------------------------------------------------------
static constexpr int CB = GLM_CB;  // t-columns per column block of the MFMA pass
static constexpr int NB = 32;   // design rows per LDS chunk

static int glm_splits(int64_t N, int64_t T, int64_t target = 768) {
    // enough workgroups to fill 256 CUs three times (three fit per CU; a
    // batch's pass sets a smaller target), chunks of >= 8 row blocks each
    const int64_t nblk = (T + CB - 1) / CB;
    int64_t ns = (target + nblk - 1) / nblk;
    const int64_t nchunks = (N + NB - 1) / NB;
    if (ns > nchunks / 8) ns = nchunks / 8;
    if (ns < 1) ns = 1;
    if (ns > 64) ns = 64;
    return (int)ns;
}

void glm_split_plan(const GlmData* g, int64_t Tn, int* ns, int64_t* rows_per_split, int target_wgs) {
    *ns = glm_splits(g->N, Tn, target_wgs);
    const int64_t nch = (g->N + NB - 1) / NB;
    *rows_per_split = ((nch + *ns - 1) / *ns) * NB;
}

void launch_glm_pass_batched(int dpt, bool jbf, const GlmPassArgs* args, const int4* items, int nitems,
                             hipStream_t s);
void launch_glm_reduce_batched(const GlmReduceArgs* args, const int4* items, int nitems, hipStream_t s);

int glm_work_ensure(GlmWork* w, const GlmData* g, hipStream_t, int64_t D, int64_t T, int probes,
                    bool f64, std::string* msg) {
    const size_t esz = f64 ? 8 : 4;
    if (w->Y && w->D == D && w->T == T && w->probes == probes && w->esz == esz) return PMCMC_OK;
    glm_work_destroy(w);
    w->D = D;
    w->T = T;
    w->probes = probes;
    w->esz = esz;
    const size_t mat = (size_t)D * T * esz;
    const int ns = f64 ? 1 : glm_splits(g->N, T);
    // partial sums: [split][set][T][Dp] floats (f32) or, for the f64 path,
    // the N x T residual matrices [set][T][N] plus [set][T][Dp] sums
    size_t part = f64 ? ((size_t)(1 + probes) * T * (g->N + g->Dp)) * sizeof(double)
                      : (size_t)ns * (1 + probes) * T * g->Dp * sizeof(float);
    part += (size_t)ns * T * sizeof(double) * (f64 ? (size_t)g->N : 1);  // log-lik partials
    w->part_bytes = part;
    bool ok = hipMalloc(&w->Y, mat) == hipSuccess && hipMalloc(&w->gx, mat) == hipSuccess &&
              hipMalloc(&w->gy, mat) == hipSuccess &&
              hipMalloc(&w->W, mat * probes) == hipSuccess &&
              hipMalloc(&w->hw, mat * probes) == hipSuccess &&
              hipMalloc(&w->lpx, T * sizeof(double)) == hipSuccess &&
              hipMalloc(&w->lpy, T * sizeof(double)) == hipSuccess &&
              hipMalloc(&w->part, part) == hipSuccess;
    if (!ok) {
        glm_work_destroy(w);
        *msg = "logistic workspace allocation failed";
        return PMCMC_ERR_ALLOC;
    }
    return PMCMC_OK;
}

void glm_work_destroy(GlmWork* w) {
    (void)hipFree(w->Y);
    (void)hipFree(w->gx);
    (void)hipFree(w->gy);
    (void)hipFree(w->W);
    (void)hipFree(w->hw);
    (void)hipFree(w->lpx);
    (void)hipFree(w->lpy);
    (void)hipFree(w->part);
    *w = GlmWork{};
}

// The block's state columns t0 .. t0 + CB - 1 of a column source (set 0:
// V[:, t], or with x0 the iterate shifted by one step) into LDS columns
// Vs[j * SV + d], zero beyond T and for d >= D up to Dp: the columns are one
// contiguous span of V, read with 16-byte loads all in flight at once (one
// load round trip, where a load per element would cost one per element a
// thread stages).  Needs a barrier before Vs is read.
template <int SV>
__device__ __forceinline__ void stage_state_cols(float* Vs, const ColSrc<float>& src, int64_t t0, int CB, int Dp,
                                                 int tid) {
    const int D = (int)src.D;
    const int64_t T = src.Tn;
    const int nc = t0 < T ? (int)(T - t0 < CB ? T - t0 : CB) : 0;
    if (nc > 0) {
        if (src.x0) {
            if (t0 == 0) {
                for (int d = tid; d < D; d += 256) Vs[d] = src.x0[d];
                span_to_lds<SV>(Vs + SV, src.V, D, nc - 1, tid);
            } else {
                span_to_lds<SV>(Vs, src.V + (size_t)(t0 - 1) * D, D, nc, tid);
            }
        } else {
            span_to_lds<SV>(Vs, src.V + (size_t)t0 * D, D, nc, tid);
        }
    }
    // zeros: rows D .. Dp - 1 of every column, every row of columns past T
    for (int i = tid; i < CB * Dp; i += 256) {
        const int j = i / Dp, d = i % Dp;
        if (d >= D || j >= nc) Vs[j * SV + d] = 0.f;
    }
}

// ---------------------------------------------------------------------------
// FP32 MFMA fused pass
// ---------------------------------------------------------------------------
// grid: (ceil(T/CB), NS); block 256 = 4 waves.
// LDS: Vs[NC][Dp+4] column block, Xs[NB][Dp+4] design chunk, Rs[NB][NC+4]
// residual tile, Ys[NB] the chunk's labels (53.4 KB at D = 100: three
// workgroups per CU).  The next design chunk and its labels are fetched into
// registers while the current one is multiplied, then stored over it once
// both GEMMs are done with it: three barriers per chunk, global latency off
// the critical path.  MFMA step 4 j + i of a 16-wide k block takes
// k = 16 j + 4 lk + i from lane group lk: the GEMM-A reads each lane's four
// k values of Xs and Vs as one 16-byte read, and the GEMM-B's lane groups
// read rows 4 apart of Xs and Rs, which fall in distinct banks.  The BF16
// pass sums its states' products in the same order.  Wave w owns column tile
// w (and w + 4, ...) for every d tile.
template <int DPT, int K>
__device__ __forceinline__ void glm_pass_body(const ColSrc<float>& src, const float* Xrm, const float* y32,
                                              int64_t N, int64_t rows_per_split, float* part, double* llpart,
                                              const int* skip, unsigned long long* flops, const int64_t cb,
                                              const int split) {
    constexpr int Dp = DPT * 16;
    constexpr int NC = CB * (1 + K);
    constexpr int SV = Dp + 4;
    constexpr int SX = Dp + 4;
    constexpr int SR = NC + 4;  // residual tile, [row][column]
    constexpr int NCT = NC / 16;             // column tiles
    static_assert(NCT % 4 == 0, "column tiles must divide over the 4 waves");
    constexpr int CT_PER_WAVE = NCT / 4;
    constexpr int F4 = NB * Dp / 4;          // float4 per design chunk
    constexpr int PF = (F4 + 255) / 256;     // float4 per thread per chunk
    __shared__ __align__(16) float Vs[NC * SV];
    __shared__ __align__(16) float Xs[NB * SX];
    __shared__ __align__(16) float Rs[NB * SR];
    __shared__ float Ys[NB];  // the chunk's labels
    // the log-likelihood reduction reuses Rs after the chunk loop (whose last
    // barrier ends every read of it)
    double (*llred)[16] = reinterpret_cast<double (*)[16]>(Rs);

    if (skip && *(volatile const int*)skip) return;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int li = lane & 15;  // row (A) / column (B, C) within a tile
    const int lk = lane >> 4;  // k index within a 4-step (A, B); row group (C)
    const int64_t D = src.D, T = src.Tn;
    const int64_t n_begin = (int64_t)split * rows_per_split;
    int64_t n_end = n_begin + rows_per_split;
    if (n_end > N) n_end = N;
    const int nchunk = n_end > n_begin ? (int)((n_end - n_begin + NB - 1) / NB) : 0;

    // design chunk c -> registers (rows beyond N are zero in Xrm, padded to 32)
    float4 pf[PF];
    float pfy = 0.f;
    auto fetch = [&](int c) {
        const int64_t n0 = n_begin + (int64_t)c * NB;
        const float4* src4 = reinterpret_cast<const float4*>(Xrm + (size_t)n0 * Dp);
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int idx = tid + 256 * i;
            if (idx < F4) pf[i] = src4[idx];
        }
        // rows past N read row N - 1 (their products are masked): no write to
        // pfy before the load, which would wait for the loads just issued
        if (tid < NB) pfy = y32[n0 + tid < N ? n0 + tid : N - 1];
    };
    auto stash = [&]() {
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int idx = tid + 256 * i;
            if (idx < F4) {
                const int e = 4 * idx;
                const int r = e / Dp, d = e % Dp;
                float2* dst = reinterpret_cast<float2*>(&Xs[r * SX + d]);
                dst[0] = make_float2(pf[i].x, pf[i].y);
                dst[1] = make_float2(pf[i].z, pf[i].w);
            }
        }
        if (tid < NB) Ys[tid] = pfy;
    };

    if (nchunk > 0) fetch(0);
    // 1. column block into LDS (zero padded): the states as one span, the
    // direction sets element by element
    stage_state_cols<SV>(Vs, src, cb * CB, CB, Dp, tid);
    for (int i = tid; i < (NC - CB) * Dp; i += 256) {
        const int j = CB + i / Dp, d = i % Dp;
        const int set = j / CB;
        const int64_t t = cb * CB + (j % CB);
        float v = 0.f;
        if (t < T && d < D)
            v = src.Z ? (float)src.get(set, t, d)
                      : rademacher(src.keys[set - 1], (uint64_t)t * (uint64_t)D + (uint64_t)d + (uint64_t)src.zoff);
        Vs[j * SV + d] = v;
    }
    if (nchunk > 0) stash();
    __syncthreads();

    v4f acc[CT_PER_WAVE][DPT];
#pragma unroll
    for (int c = 0; c < CT_PER_WAVE; ++c)
#pragma unroll
        for (int dt = 0; dt < DPT; ++dt) acc[c][dt] = (v4f){0.f, 0.f, 0.f, 0.f};
    double ll = 0.0;

    const int rt = wave >> 1;  // GEMM-A row tile
    const int cx = wave & 1;   // GEMM-A x column tile

    for (int c = 0; c < nchunk; ++c) {
        const int64_t n0 = n_begin + (int64_t)c * NB;
        if (c + 1 < nchunk) fetch(c + 1);
        // GEMM-A: logits of the x tile and the matching probe tiles
        v4f cl[1 + K];
#pragma unroll
        for (int s = 0; s <= K; ++s) cl[s] = (v4f){0.f, 0.f, 0.f, 0.f};
        // step 4 j + i takes k = 16 j + 4 lk + i from lane group lk (16-byte reads)
#pragma unroll
        for (int j = 0; j < Dp / 16; ++j) {
            const float4 a4 = *reinterpret_cast<const float4*>(&Xs[(16 * rt + li) * SX + 16 * j + 4 * lk]);
            const float av[4] = {a4.x, a4.y, a4.z, a4.w};
#pragma unroll
            for (int s = 0; s <= K; ++s) {
                const int col = s * CB + 16 * cx + li;
                const float4 b4 = *reinterpret_cast<const float4*>(&Vs[col * SV + 16 * j + 4 * lk]);
                const float bv[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) cl[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bv[i], cl[s], 0, 0, 0);
            }
        }
        // elementwise: residuals and sigmoid'-weighted directions into LDS.
        // With e = exp(-|l|): softplus(l) = max(l,0) + log(1+e),
        // sigmoid(l) = (l >= 0 ? 1 : e) / (1+e), log-lik = y l - softplus(l).
        v4f rv[1 + K];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * rt + 4 * lk + r;
            const int64_t n = n0 + row;
            const bool valid = n < N;
            const float l = cl[0][r];
            const float e = __expf(-fabsf(l));
            const float inv = 1.f / (1.f + e);
            const float sg = l >= 0.f ? inv : e * inv;
            const float yv = Ys[row];
            if (valid) ll += (double)(yv * l - (fmaxf(l, 0.f) + __logf(1.f + e)));
            rv[0][r] = valid ? (yv - sg) : 0.f;
            const float wgt = sg * (1.f - sg);
#pragma unroll
            for (int s = 1; s <= K; ++s) rv[s][r] = valid ? wgt * cl[s][r] : 0.f;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int s = 0; s <= K; ++s) Rs[(16 * rt + 4 * lk + r) * SR + s * CB + 16 * cx + li] = rv[s][r];
        __syncthreads();
        // GEMM-B: accumulate Xd^T R; step 4 j + i takes chunk row k = 16 j + 4 lk + i
        // from lane group lk (the residual tile as 16-byte reads)
#pragma unroll
        for (int ks = 0; ks < NB / 4; ++ks) {
            const int kr = 16 * (ks >> 2) + 4 * lk + (ks & 3);  // rows 4 apart: distinct banks
            float xa[DPT];
#pragma unroll
            for (int dt = 0; dt < DPT; ++dt) xa[dt] = Xs[kr * SX + 16 * dt + li];
#pragma unroll
            for (int cc = 0; cc < CT_PER_WAVE; ++cc) {
                // NCT is a multiple of 4: every wave owns CT_PER_WAVE tiles (no branch,
                // so the accumulators stay in place across k steps)
                const int ct = wave + 4 * cc;
                const float bv = Rs[kr * SR + 16 * ct + li];
#pragma unroll
                for (int dt = 0; dt < DPT; ++dt)
                    acc[cc][dt] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[dt], bv, acc[cc][dt], 0, 0, 0);
            }
        }
        if (c + 1 < nchunk) {
            __syncthreads();  // both GEMMs are done with this chunk
            stash();
        }
        __syncthreads();
    }

    // partial sums out: part[split][set][t][Dp]
#pragma unroll
    for (int cc = 0; cc < CT_PER_WAVE; ++cc) {
        const int ct = wave + 4 * cc;
        if (ct >= NCT) continue;
        const int j = 16 * ct + li;
        const int set = j / CB;
        const int64_t t = cb * CB + (j % CB);
        if (t >= T) continue;
        float* dst = part + (((size_t)split * (1 + K) + set) * T + t) * Dp;
#pragma unroll
        for (int dt = 0; dt < DPT; ++dt)
#pragma unroll
            for (int r = 0; r < 4; ++r) dst[16 * dt + 4 * lk + r] = acc[cc][dt][r];
    }
    // log-likelihood of column 16*cx + li: lanes li, li+16, li+32, li+48 and waves cx, cx+2
    ll += __shfl_xor(ll, 16, 64);
    ll += __shfl_xor(ll, 32, 64);
    if (lk == 0) llred[wave][li] = ll;
    __syncthreads();
    if (tid < CB) {
        const int cxx = tid / 16, l2 = tid % 16;
        const double v = llred[cxx][l2] + llred[cxx + 2][l2];
        const int64_t t = cb * CB + tid;
        if (t < T) llpart[(size_t)split * T + t] = v;
    }
    if (flops && tid == 0 && n_end > n_begin && cb * CB < T) {
        const int64_t nc = T - cb * CB < CB ? T - cb * CB : CB;
        atomicAdd(flops, (unsigned long long)(8 * (n_end - n_begin) * D * nc));
    }
}

// ---------------------------------------------------------------------------
// BF16-Jacobian pass: the states' products (logits, residual, gradient) on
// the FP32 MFMA as in the FP32 pass; the directions' products on the
// 16x16x32 BF16 MFMA with operands that are BF16 already:
//   * the design's BF16 copies (row-major Xb for GEMM-A, transposed Xtb for
//     GEMM-B), 16-byte loads straight from global memory (L2): GEMM-A's at
//     the start of its chunk, ahead of the FP32 logit MFMAs that hide their
//     latency, GEMM-B's a chunk ahead;
//   * the block's 32 direction columns, rounded once into registers (each
//     wave's GEMM-A B operand for every k step);
//   * the directions' sigmoid'-weighted logits u, written to LDS as BF16 in
//     [column][row] order, so that a lane's eight k values are one 16-byte
//     read.
// The FP32 design chunk, its labels and the 32 state columns go through LDS
// as in the FP32 pass, with the same 16-byte k runs and the same summation
// order, so the states' products equal the FP32 pass's bit for bit; each
// wave keeps its GEMM-A state column (the same for every chunk) in registers
// up to D = 128.  The directions' values and summation order are those of
// rounding the FP32 operands to BF16 inside the products.
// Waves: GEMM-A, wave (rt, cx) owns row tile rt and column tile cx of the
// states and of the directions; GEMM-B, wave w owns d tiles w, w + 4, ...
// for all four column tiles (two of states, two of directions).
// ---------------------------------------------------------------------------
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

template <int DPT>
__device__ __forceinline__ void glm_pass_bf16_body(const ColSrc<float>& src, const float* Xrm,
                                                   const uint16_t* Xb16, const uint16_t* Xtb16, int64_t Np,
                                                   const float* y32, int64_t N, int64_t rows_per_split,
                                                   float* part, double* llpart, const int* skip,
                                                   unsigned long long* flops, const int64_t cb, const int split) {
    constexpr int Dp = DPT * 16;
    constexpr int Dq = (Dp + 31) / 32 * 32;  // K of the BF16 GEMM-A, in steps of 32
    constexpr int KQ = Dq / 32;
    constexpr int SV = Dp + 4;
    constexpr int SX = Dp + 4;
    constexpr int SR = NB + 4;  // states' residual, FP32 [column][row]: 16-byte row reads hit distinct banks
    constexpr int SU = NB + 8;  // directions' u, BF16 [column][row]: 80-byte rows, so the
                                // 16-byte reads of 8 consecutive lanes cover distinct banks
    constexpr int F4 = NB * Dp / 4;
    constexpr int PF = (F4 + 255) / 256;
    constexpr int DPW = (DPT + 3) / 4;
    __shared__ __align__(16) float Vs[CB * SV];
    __shared__ __align__(16) float Xs[NB * SX];
    __shared__ __align__(16) float Rs[CB * SR];
    __shared__ __align__(16) __bf16 Ub[CB * SU];
    __shared__ float Ys[NB];  // the chunk's labels
    __shared__ double llred[4][16];

    if (skip && *(volatile const int*)skip) return;
    const __bf16* Xb = reinterpret_cast<const __bf16*>(Xb16);
    const __bf16* Xtb = reinterpret_cast<const __bf16*>(Xtb16);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int li = lane & 15;
    const int lk = lane >> 4;
    const int64_t D = src.D, T = src.Tn;
    const int64_t n_begin = (int64_t)split * rows_per_split;
    int64_t n_end = n_begin + rows_per_split;
    if (n_end > N) n_end = N;
    const int nchunk = n_end > n_begin ? (int)((n_end - n_begin + NB - 1) / NB) : 0;
    const int rt = wave >> 1;  // GEMM-A row tile
    const int cx = wave & 1;   // GEMM-A column tile

    float4 pf[PF];
    float pfy = 0.f;
    auto fetch = [&](int c) {
        const int64_t n0 = n_begin + (int64_t)c * NB;
        const float4* src4 = reinterpret_cast<const float4*>(Xrm + (size_t)n0 * Dp);
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int idx = tid + 256 * i;
            if (idx < F4) pf[i] = src4[idx];
        }
        // rows past N read row N - 1 (their products are masked): no write to
        // pfy before the load, which would wait for the loads just issued
        if (tid < NB) pfy = y32[n0 + tid < N ? n0 + tid : N - 1];
    };
    auto stash = [&]() {
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int idx = tid + 256 * i;
            if (idx < F4) {
                const int e = 4 * idx;
                const int r = e / Dp, d = e % Dp;
                float2* dst = reinterpret_cast<float2*>(&Xs[r * SX + d]);
                dst[0] = make_float2(pf[i].x, pf[i].y);
                dst[1] = make_float2(pf[i].z, pf[i].w);
            }
        }
        if (tid < NB) Ys[tid] = pfy;
    };
    // BF16 design operands of chunk c: GEMM-A rows 16 rt + li, k = 32 q + 8 lk
    // + e (Xb); GEMM-B d tile rows 16 dt + li, k = rows 8 lk + e of the chunk
    // (Xtb; d tiles clamped, a clamped duplicate is not stored)
    v8bf xa[KQ], xt[DPW];
    auto fetch_xa = [&](int c) {
        const int64_t n0 = n_begin + (int64_t)c * NB;
        const __bf16* ra = Xb + (size_t)(n0 + 16 * rt + li) * Dq + 8 * lk;
#pragma unroll
        for (int q = 0; q < KQ; ++q) xa[q] = *reinterpret_cast<const v8bf*>(ra + 32 * q);
    };
    auto fetch_xt = [&](int c) {
        const int64_t n0 = n_begin + (int64_t)c * NB;
#pragma unroll
        for (int m = 0; m < DPW; ++m) {
            const int dt = min(wave + 4 * m, DPT - 1);
            xt[m] = *reinterpret_cast<const v8bf*>(Xtb + (size_t)(16 * dt + li) * Np + n0 + 8 * lk);
        }
    };
    if (nchunk > 0) {
        fetch(0);
        fetch_xt(0);
    }
    // states of the block into LDS (zero padded)
    stage_state_cols<SV>(Vs, src, cb * CB, CB, Dp, tid);
    // directions of column 16 cx + li, rounded once: the GEMM-A B operand
    v8bf zb[KQ];
    {
        const int64_t t = cb * CB + 16 * cx + li;
        if (!src.Z) {
            // the probe: +-1, exact in BF16; hashed straight from its element index
            const uint64_t base = (uint64_t)t * (uint64_t)D + (uint64_t)src.zoff;
#pragma unroll
            for (int q = 0; q < KQ; ++q)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int d = 32 * q + 8 * lk + e;
                    zb[q][e] = (__bf16)((t < T && d < D) ? rademacher(src.keys[0], base + (uint64_t)d) : 0.f);
                }
        } else {
#pragma unroll
            for (int q = 0; q < KQ; ++q)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int64_t d = 32 * q + 8 * lk + e;
                    zb[q][e] = (__bf16)((t < T && d < D) ? (float)src.get(1, t, d) : 0.f);
                }
        }
    }
    if (nchunk > 0) stash();
    __syncthreads();
    // GEMM-A B operand: the block's state columns, the same for every chunk
    constexpr int KA = Dp / 4;
    constexpr bool VREG = DPT <= 8;
    float vreg[VREG ? KA : 1];
    if constexpr (VREG) {
#pragma unroll
        for (int j = 0; j < KA / 4; ++j) {
            const float4 v4 = *reinterpret_cast<const float4*>(&Vs[(16 * cx + li) * SV + 16 * j + 4 * lk]);
            vreg[4 * j] = v4.x;
            vreg[4 * j + 1] = v4.y;
            vreg[4 * j + 2] = v4.z;
            vreg[4 * j + 3] = v4.w;
        }
    }

    v4f accx[DPW][2], accp[DPW][2];
#pragma unroll
    for (int m = 0; m < DPW; ++m)
#pragma unroll
        for (int c = 0; c < 2; ++c) accx[m][c] = accp[m][c] = (v4f){0.f, 0.f, 0.f, 0.f};
    double ll = 0.0;

    for (int c = 0; c < nchunk; ++c) {
        const int64_t n0 = n_begin + (int64_t)c * NB;
        const bool more = c + 1 < nchunk;
        if (more) fetch(c + 1);
        fetch_xa(c);
        // GEMM-A: state logits (FP32) and direction logits (BF16)
        v4f cl = (v4f){0.f, 0.f, 0.f, 0.f}, cq = cl;
        {
            // every operand read issued before the products: one LDS latency per
            // chunk instead of one per pair of MFMAs
            // step 4 j + i takes k = 16 j + 4 lk + i from lane group lk (16-byte reads)
            float ar[KA], br[VREG ? 1 : KA];
#pragma unroll
            for (int j = 0; j < KA / 4; ++j) {
                const float4 a4 = *reinterpret_cast<const float4*>(&Xs[(16 * rt + li) * SX + 16 * j + 4 * lk]);
                ar[4 * j] = a4.x;
                ar[4 * j + 1] = a4.y;
                ar[4 * j + 2] = a4.z;
                ar[4 * j + 3] = a4.w;
                if constexpr (!VREG) {
                    const float4 b4 = *reinterpret_cast<const float4*>(&Vs[(16 * cx + li) * SV + 16 * j + 4 * lk]);
                    br[4 * j] = b4.x;
                    br[4 * j + 1] = b4.y;
                    br[4 * j + 2] = b4.z;
                    br[4 * j + 3] = b4.w;
                }
            }
#pragma unroll
            for (int ks = 0; ks < KA; ++ks) {
                float bv;
                if constexpr (VREG) bv = vreg[ks]; else bv = br[ks];
                cl = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[ks], bv, cl, 0, 0, 0);
            }
        }
#pragma unroll
        for (int q = 0; q < KQ; ++q) cq = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[q], zb[q], cq, 0, 0, 0);
        // elementwise: the states' residual (FP32) and the directions' u (BF16)
        v4bf u4;
        v4f r4;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * rt + 4 * lk + r;
            const int64_t n = n0 + row;
            const bool valid = n < N;
            const float l = cl[r];
            const float e = __expf(-fabsf(l));
            const float inv = 1.f / (1.f + e);
            const float sg = l >= 0.f ? inv : e * inv;
            const float yv = Ys[row];
            if (valid) ll += (double)(yv * l - (fmaxf(l, 0.f) + __logf(1.f + e)));
            r4[r] = valid ? (yv - sg) : 0.f;
            u4[r] = (__bf16)(valid ? sg * (1.f - sg) * cq[r] : 0.f);
        }
        *reinterpret_cast<v4f*>(&Rs[(16 * cx + li) * SR + 16 * rt + 4 * lk]) = r4;
        *reinterpret_cast<v4bf*>(&Ub[(16 * cx + li) * SU + 16 * rt + 4 * lk]) = u4;
        __syncthreads();
        // GEMM-B: Xd^T [r | u] over the chunk's 32 rows; the residual operands
        // (the same for every d tile) and each d tile's design values are read
        // before its products
        // (step 4 j + i takes chunk row k = 16 j + 4 lk + i from lane group lk)
        v4f rr[NB / 16][2];
#pragma unroll
        for (int j = 0; j < NB / 16; ++j)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc)
                rr[j][cc] = *reinterpret_cast<const v4f*>(&Rs[(16 * cc + li) * SR + 16 * j + 4 * lk]);
#pragma unroll
        for (int m = 0; m < DPW; ++m) {
            const int dt = min(wave + 4 * m, DPT - 1);
            float xr[NB / 4];
#pragma unroll
            for (int ks = 0; ks < NB / 4; ++ks) xr[ks] = Xs[(16 * (ks >> 2) + 4 * lk + (ks & 3)) * SX + 16 * dt + li];
#pragma unroll
            for (int ks = 0; ks < NB / 4; ++ks)
#pragma unroll
                for (int cc = 0; cc < 2; ++cc)
                    accx[m][cc] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[ks], rr[ks >> 2][cc][ks & 3], accx[m][cc], 0, 0, 0);
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                const v8bf bv = *reinterpret_cast<const v8bf*>(&Ub[(16 * cc + li) * SU + 8 * lk]);
                accp[m][cc] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xt[m], bv, accp[m][cc], 0, 0, 0);
            }
        }
        if (more) fetch_xt(c + 1);
        if (more) {
            __syncthreads();  // both GEMMs are done with this chunk
            stash();
        }
        __syncthreads();
    }

    // partial sums out: part[split][set][t][Dp]
#pragma unroll
    for (int m = 0; m < DPW; ++m) {
        const int dt = wave + 4 * m;
        if (dt >= DPT) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int set = c >> 1;
            const int64_t t = cb * CB + 16 * (c & 1) + li;
            if (t >= T) continue;
            float* dst = part + (((size_t)split * 2 + set) * T + t) * Dp;
            const v4f v = c < 2 ? accx[m][c & 1] : accp[m][c & 1];
#pragma unroll
            for (int r = 0; r < 4; ++r) dst[16 * dt + 4 * lk + r] = v[r];
        }
    }
    ll += __shfl_xor(ll, 16, 64);
    ll += __shfl_xor(ll, 32, 64);
    if (lk == 0) llred[wave][li] = ll;
    __syncthreads();
    if (tid < CB) {
        const int cxx = tid / 16, l2 = tid % 16;
        const double v = llred[cxx][l2] + llred[cxx + 2][l2];
        const int64_t t = cb * CB + tid;
        if (t < T) llpart[(size_t)split * T + t] = v;
    }
    if (flops && tid == 0 && n_end > n_begin && cb * CB < T) {
        const int64_t nc = T - cb * CB < CB ? T - cb * CB : CB;
        atomicAdd(flops, (unsigned long long)(8 * (n_end - n_begin) * D * nc));
    }
}

// The BF16 form is held to two waves per SIMD: at the widest designs its
// registers would otherwise allow only one, with no second wave to run while
// the first waits at a barrier.
template <int DPT, int K>
__global__ __launch_bounds__(256) void glm_pass_mfma_kernel(ColSrc<float> src, const float* Xrm, const float* y32,
                                                            int64_t N, int64_t rows_per_split, float* part,
                                                            double* llpart, const int* skip) {
    glm_pass_body<DPT, K>(src, Xrm, y32, N, rows_per_split, part, llpart, skip, nullptr, blockIdx.x, blockIdx.y);
}

template <int DPT>
__global__ __launch_bounds__(256, 2) void glm_pass_bf16_kernel(ColSrc<float> src, const float* Xrm,
                                                               const uint16_t* Xb, const uint16_t* Xtb, int64_t Np,
                                                               const float* y32, int64_t N, int64_t rows_per_split,
                                                               float* part, double* llpart, const int* skip) {
    glm_pass_bf16_body<DPT>(src, Xrm, Xb, Xtb, Np, y32, N, rows_per_split, part, llpart, skip, nullptr, blockIdx.x,
                            blockIdx.y);
}

// Batched over chains of one padded width: work item blockIdx.x is (chain,
// column block, N split, split count) = items[blockIdx.x]; the splits cover
// the chain's row chunks evenly (the split count is chosen per item list).
__device__ inline int64_t batched_split_rows(int64_t N, int ns) {
    const int64_t nch = (N + NB - 1) / NB;
    return (nch + ns - 1) / ns * NB;
}

template <int DPT, int K>
__global__ __launch_bounds__(256) void glm_pass_mfma_batched(const GlmPassArgs* __restrict__ args, const int4* __restrict__ items) {
    const int4 w = items[blockIdx.x];
    const GlmPassArgs& a = args[w.x];
    if ((int64_t)w.y * CB >= a.src.Tn) return;  // beyond the chain's current window
    glm_pass_body<DPT, K>(a.src, a.Xrm, a.y32, a.N, batched_split_rows(a.N, w.w), a.part, a.llpart, a.skip, a.flops,
                          w.y, w.z);
}

template <int DPT>
__global__ __launch_bounds__(256, 2) void glm_pass_bf16_batched(const GlmPassArgs* __restrict__ args, const int4* __restrict__ items) {
    const int4 w = items[blockIdx.x];
    const GlmPassArgs& a = args[w.x];
    if ((int64_t)w.y * CB >= a.src.Tn) return;  // beyond the chain's current window
    glm_pass_bf16_body<DPT>(a.src, a.Xrm, a.Xb, a.Xtb, a.Np, a.y32, a.N, batched_split_rows(a.N, w.w), a.part,
                            a.llpart, a.skip, a.flops, w.y, w.z);
}

// Sum the splits, add the prior terms, and form the next pass's columns.
//   pass 0 (at Xbar): gx = G0 - x/s^2 ; hz_k = -G_k - z_k/s^2 ;
//                     Y = x + eps gx + sqrt(2 eps) xi ; W_k = z_k + eps hz_k ;
//                     lpx = ll - |x|^2 / (2 s^2)
//   pass 1 (at Y):    gy = G0 - y/s^2 ; hw_k = -G_k - w_k/s^2 ; lpy likewise
// One wave per column.
template <typename T, typename P>
__device__ __forceinline__ void glm_reduce_body(const ColSrc<T>& src, const P* part, const double* llpart, int ns,
                                                int64_t Dp, int64_t llstride, int pass, double inv_s2, double eps,
                                                const double* Pm, const T* Xi, T* gout, T* hout, double* lpout,
                                                T* Yout, T* Wout, const int* skip, const int64_t blk) {
    if (skip && *(volatile const int*)skip) return;
    const int lane = threadIdx.x & 63;
    const int64_t t = blk * 4 + (threadIdx.x >> 6);
    const int64_t D = src.D, Tn = src.Tn;
    const int K = src.K;
    if (t >= Tn) return;
    const double sq2e = sqrt(2.0 * eps);
    double nrm = 0.0;
    // prior term of the gradient / HVP: P v (dense prior) or v / s^2
    auto prior = [&](int set, int64_t d, double v) -> double {
        if (!Pm) return v * inv_s2;
        double a = 0.0;
        for (int64_t k = 0; k < D; ++k) a += Pm[(size_t)d * D + k] * src.get(set, t, k);
        return a;
    };
    // The splits' partial sums of a set, added in split order.  Loaded in
    // groups of 8 splits, all in flight together, with the first direction
    // set's alongside the states' (a loop with a runtime trip count would
    // wait for each load in turn); clamped indices, so no load is branched.
    constexpr int SG = 8;
    auto split_sums = [&](int64_t d, double* G0, double* G1) {
        double a0 = 0.0, a1 = 0.0;
        for (int s0 = 0; s0 < ns; s0 += SG) {
            P v0[SG], v1[SG];
#pragma unroll
            for (int j = 0; j < SG; ++j) {
                const size_t sj = (size_t)(s0 + j < ns ? s0 + j : ns - 1);
                v0[j] = part[((sj * (1 + K) + 0) * Tn + t) * Dp + d];
                if (K > 0) v1[j] = part[((sj * (1 + K) + 1) * Tn + t) * Dp + d];
            }
#pragma unroll
            for (int j = 0; j < SG; ++j)
                if (s0 + j < ns) {
                    a0 += (double)v0[j];
                    if (K > 0) a1 += (double)v1[j];
                }
        }
        *G0 = a0;
        *G1 = a1;
    };
    for (int64_t d = lane; d < D; d += 64) {
        const double x = src.get(0, t, d);
        double G, H1;
        split_sums(d, &G, &H1);
        const double px = prior(0, d, x);
        const double g = G - px;
        const size_t gi = (size_t)t * D + d;
        gout[gi] = (T)g;
        nrm += x * px;
        double y = 0.0;
        if (pass == 0) {
            y = x + eps * g + sq2e * (double)Xi[gi];
            Yout[gi] = (T)y;
        }
        for (int k = 0; k < K; ++k) {
            double H = H1;
            if (k > 0) {
                H = 0.0;
                for (int s = 0; s < ns; ++s)
                    H += (double)part[(((size_t)s * (1 + K) + 1 + k) * Tn + t) * Dp + d];
            }
            const double z = src.get(1 + k, t, d);
            const double h = -H - prior(1 + k, d, z);
            hout[(size_t)k * D * Tn + gi] = (T)h;
            if (pass == 0) Wout[(size_t)k * D * Tn + gi] = (T)(z + eps * h);
        }
    }
    for (int m = 32; m >= 1; m >>= 1) nrm += __shfl_xor(nrm, m, 64);
    if (lane == 0) {
        double ll = 0.0;
        for (int s0 = 0; s0 < ns; s0 += SG) {
            double v[SG];
#pragma unroll
            for (int j = 0; j < SG; ++j) v[j] = llpart[(size_t)(s0 + j < ns ? s0 + j : ns - 1) * llstride + t];
#pragma unroll
            for (int j = 0; j < SG; ++j)
                if (s0 + j < ns) ll += v[j];
        }
        lpout[t] = ll - 0.5 * nrm;  // nrm = x' P x (x' x / s^2 for the isotropic prior)
    }
}

template <typename T, typename P>
__global__ __launch_bounds__(256) void glm_reduce_kernel(ColSrc<T> src, const P* part, const double* llpart,
                                                         int ns, int64_t Dp, int64_t llstride, int pass,
                                                         double inv_s2, double eps, const double* Pm,
                                                         const T* Xi, T* gout, T* hout, double* lpout, T* Yout,
                                                         T* Wout, const int* skip) {
    glm_reduce_body<T, P>(src, part, llpart, ns, Dp, llstride, pass, inv_s2, eps, Pm, Xi, gout, hout, lpout, Yout,
                          Wout, skip, blockIdx.x);
}

// Batched: work item blockIdx.x is (chain, block of 4 columns, the pass's
// split count) = items[blockIdx.x].
__global__ __launch_bounds__(256) void glm_reduce_batched(const GlmReduceArgs* __restrict__ args, const int4* __restrict__ items) {
    const int4 w = items[blockIdx.x];
    const GlmReduceArgs& a = args[w.x];
    glm_reduce_body<float, float>(a.src, a.part, a.llpart, w.z, a.Dp, a.llstride, a.pass, a.inv_s2, a.eps, a.P,
                                  a.Xi, a.gout, a.hout, a.lpout, a.Yout, a.Wout, a.skip, w.y);
}

// ---------------------------------------------------------------------------
// FP64 plain pass: residual matrices R[set][t][n] then Xd^T R.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void glm_naive_fwd_kernel(ColSrc<T> src, const double* X64, const double* y64, int64_t N,
                                     double* R, double* llnt, const int* skip) {
    if (skip && *(volatile const int*)skip) return;
    const int64_t D = src.D, Tn = src.Tn;
    const int K = src.K;
    const int64_t total = N * Tn;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = i % N, t = i / N;
        double l = 0.0;
        for (int64_t d = 0; d < D; ++d) l += X64[n + d * N] * src.get(0, t, d);
        const double sg = sigmoidd_(l);
        const double yv = y64[n];
        R[((size_t)0 * Tn + t) * N + n] = yv - sg;
        llnt[(size_t)t * N + n] = -yv * softplusd(-l) - (1.0 - yv) * softplusd(l);
        for (int k = 0; k < K; ++k) {
            double q = 0.0;
            for (int64_t d = 0; d < D; ++d) q += X64[n + d * N] * src.get(1 + k, t, d);
            R[((size_t)(1 + k) * Tn + t) * N + n] = sg * (1.0 - sg) * q;
        }
    }
}

// sums[set][t][Dp] = Xd^T R[set][t]; llsum[t] = sum_n llnt[t][n]
template <typename T>
__global__ void glm_naive_bwd_kernel(int K, int64_t D, int64_t Dp, int64_t Tn, const double* X64,
                                     int64_t N, const double* R, const double* llnt, double* sums,
                                     double* llsum, const int* skip) {
    if (skip && *(volatile const int*)skip) return;
    const int64_t total = (int64_t)(1 + K) * Tn * D;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = i % D;
        const int64_t st = i / D;  // set * Tn + t
        const double* r = R + (size_t)st * N;
        double acc = 0.0;
        for (int64_t n = 0; n < N; ++n) acc += X64[n + d * N] * r[n];
        sums[(size_t)st * Dp + d] = acc;
        if (d == 0 && st < Tn) {
            double s = 0.0;
            const double* lr = llnt + (size_t)st * N;
            for (int64_t n = 0; n < N; ++n) s += lr[n];
            llsum[st] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// one pass (dispatch)
// ---------------------------------------------------------------------------
template <int DPT>
static void launch_mfma(const ColSrc<float>& src, const GlmData* g, int ns, int64_t rows_per_split,
                        float* part, double* llpart, const int* skip, bool jbf, hipStream_t s) {
    dim3 grid((unsigned)((src.Tn + CB - 1) / CB), (unsigned)ns);
    if (jbf)
        hipLaunchKernelGGL((glm_pass_bf16_kernel<DPT>), grid, dim3(256), 0, s, src, g->Xrm, g->Xb, g->Xtb, g->Np,
                           g->y, g->N, rows_per_split, part, llpart, skip);
    else
        hipLaunchKernelGGL((glm_pass_mfma_kernel<DPT, 1>), grid, dim3(256), 0, s, src, g->Xrm, g->y, g->N,
                           rows_per_split, part, llpart, skip);
}

template <typename T>
static int glm_pass(const GlmData* g, GlmWork* w, hipStream_t s, const ColSrc<T>& src, int pass,
                    double eps, double prior_scale, const T* Xi, T* gout, T* hout, double* lpout,
                    T* Yout, T* Wout, const int* skip, Prof* prof, int iter, bool jbf,
                    std::string* msg) {
    const int64_t Tn = src.Tn;
    const int K = src.K;
    const double inv_s2 = 1.0 / (prior_scale * prior_scale);
    const int rgrid = (int)((Tn + 3) / 4);
    const bool mfma = sizeof(T) == 4 && K == 1 && g->Dp <= 256;
    if (mfma) {
        const int ns = glm_splits(g->N, Tn);
        const int64_t nch = (g->N + NB - 1) / NB;
        const int64_t rows_per_split = ((nch + ns - 1) / ns) * NB;
        float* part = (float*)w->part;
        double* llpart = (double*)((char*)w->part + (size_t)ns * (1 + K) * Tn * g->Dp * sizeof(float));
        const ColSrc<float>& fs = reinterpret_cast<const ColSrc<float>&>(src);
        const int pe = prof ? prof->begin(KT_GLM_PASS, iter, s) : -1;
        switch (g->Dp / 16) {
#define PMCMC_DPT(n) \
    case n: launch_mfma<n>(fs, g, ns, rows_per_split, part, llpart, skip, jbf, s); break;
            PMCMC_DPT(1) PMCMC_DPT(2) PMCMC_DPT(3) PMCMC_DPT(4) PMCMC_DPT(5) PMCMC_DPT(6)
            PMCMC_DPT(7) PMCMC_DPT(8) PMCMC_DPT(9) PMCMC_DPT(10) PMCMC_DPT(11) PMCMC_DPT(12)
            PMCMC_DPT(13) PMCMC_DPT(14) PMCMC_DPT(15) PMCMC_DPT(16)
#undef PMCMC_DPT
            default:
                *msg = "logistic target: unsupported D";
                return PMCMC_ERR_UNSUPPORTED;
        }
        if (prof) prof->end(pe, s);
        const int pr = prof ? prof->begin(KT_GLM_REDUCE, iter, s) : -1;
        hipLaunchKernelGGL((glm_reduce_kernel<T, float>), dim3(rgrid), dim3(256), 0, s, src, part, llpart, ns,
                           g->Dp, Tn, pass, inv_s2, eps, g->P64, Xi, gout, hout, lpout, Yout, Wout, skip);
        if (prof) prof->end(pr, s);
    } else {
        // plain path: residual matrices in the part buffer
        double* R = (double*)w->part;
        double* sums = R + (size_t)(1 + K) * Tn * g->N;
        double* llnt = sums + (size_t)(1 + K) * Tn * g->Dp;
        double* llsum = llnt + (size_t)Tn * g->N;
        const size_t need = ((size_t)(1 + K) * Tn * (g->N + g->Dp) + (size_t)Tn * g->N + Tn) * sizeof(double);
        if (need > w->part_bytes) {
            (void)hipFree(w->part);
            w->part = nullptr;
            if (hipMalloc(&w->part, need) != hipSuccess) {
                *msg = "logistic workspace allocation failed";
                return PMCMC_ERR_ALLOC;
            }
            w->part_bytes = need;
            R = (double*)w->part;
            sums = R + (size_t)(1 + K) * Tn * g->N;
            llnt = sums + (size_t)(1 + K) * Tn * g->Dp;
            llsum = llnt + (size_t)Tn * g->N;
        }
        hipLaunchKernelGGL((glm_naive_fwd_kernel<T>), dim3(grid_cap(g->N * Tn)), dim3(256), 0, s, src, g->Xd64,
                           g->y64, g->N, R, llnt, skip);
        hipLaunchKernelGGL((glm_naive_bwd_kernel<T>), dim3(grid_cap((1 + K) * Tn * src.D)), dim3(256), 0, s, K,
                           src.D, g->Dp, Tn, g->Xd64, g->N, R, llnt, sums, llsum, skip);
        hipLaunchKernelGGL((glm_reduce_kernel<T, double>), dim3(rgrid), dim3(256), 0, s, src, sums, llsum, 1,
                           g->Dp, Tn, pass, inv_s2, eps, g->P64, Xi, gout, hout, lpout, Yout, Wout, skip);
    }
    if (hipGetLastError() != hipSuccess) {
        *msg = "logistic pass: kernel launch failed";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}

template <typename T>
int glm_newton_products(GlmWork* w, const GlmData* g, hipStream_t s, SyncState* sync, int precision,
                        double eps, double prior_scale, const T* x0, const T* S_in, const T* Xi,
                        uint64_t seed, int iter, uint32_t probe_iter, int probes, int64_t D, int64_t Tn,
                        Prof* prof, std::string* msg) {
    const bool jbf = precision == PMCMC_PREC_BF16_JAC;
    const int* skip = sync ? &sync->converged : nullptr;
    ColSrc<T> a{};
    a.V = S_in;
    a.x0 = x0;
    a.Z = nullptr;
    a.K = probes;
    a.D = D;
    a.Tn = Tn;
    for (int k = 0; k < probes && k < 4; ++k) a.keys[k] = probe_key(seed, probe_iter, (uint32_t)k);
    int rc = glm_pass<T>(g, w, s, a, 0, eps, prior_scale, Xi, (T*)w->gx, (T*)w->hw, w->lpx, (T*)w->Y,
                         (T*)w->W, skip, prof, iter, jbf, msg);
    if (rc != PMCMC_OK) return rc;
    ColSrc<T> b{};
    b.V = (const T*)w->Y;
    b.x0 = nullptr;
    b.Z = (const T*)w->W;
    b.K = probes;
    b.D = D;
    b.Tn = Tn;
    return glm_pass<T>(g, w, s, b, 1, eps, prior_scale, Xi, (T*)w->gy, (T*)w->hw, w->lpy, nullptr, nullptr,
                       skip, prof, iter, jbf, msg);
}
template int glm_newton_products<float>(GlmWork*, const GlmData*, hipStream_t, SyncState*, int,
                                        double, double, const float*, const float*, const float*,
                                        uint64_t, int, uint32_t, int, int64_t, int64_t, Prof*, std::string*);
template int glm_newton_products<double>(GlmWork*, const GlmData*, hipStream_t, SyncState*, int,
                                         double, double, const double*, const double*,
                                         const double*, uint64_t, int, uint32_t, int, int64_t, int64_t, Prof*,
                                         std::string*);

template <int DPT>
static void launch_glm_batched_dpt(bool jbf, const GlmPassArgs* args, const int4* items, int nitems,
                                   hipStream_t s) {
    if (jbf)
        hipLaunchKernelGGL((glm_pass_bf16_batched<DPT>), dim3((unsigned)nitems), dim3(256), 0, s, args, items);
    else
        hipLaunchKernelGGL((glm_pass_mfma_batched<DPT, 1>), dim3((unsigned)nitems), dim3(256), 0, s, args, items);
}

void launch_glm_pass_batched(int dpt, bool jbf, const GlmPassArgs* args, const int4* items, int nitems,
                             hipStream_t s) {
    if (nitems <= 0) return;
    switch (dpt) {
#define PMCMC_BDPT(n) \
    case n: launch_glm_batched_dpt<n>(jbf, args, items, nitems, s); break;
        PMCMC_BDPT(1) PMCMC_BDPT(2) PMCMC_BDPT(3) PMCMC_BDPT(4) PMCMC_BDPT(5) PMCMC_BDPT(6) PMCMC_BDPT(7)
        PMCMC_BDPT(8) PMCMC_BDPT(9) PMCMC_BDPT(10) PMCMC_BDPT(11) PMCMC_BDPT(12) PMCMC_BDPT(13) PMCMC_BDPT(14)
        PMCMC_BDPT(15) PMCMC_BDPT(16)
#undef PMCMC_BDPT
        default: break;
    }
}

void launch_glm_reduce_batched(const GlmReduceArgs* args, const int4* items, int nitems, hipStream_t s) {
    if (nitems > 0) hipLaunchKernelGGL(glm_reduce_batched, dim3((unsigned)nitems), dim3(256), 0, s, args, items);
}

// One pass at the states V (shifted by x0 when given) with one direction set
// (Z, or the probe with `key` when Z is null): w->gx = grad, w->W = dir + eps H dir.
template <typename T>
int glm_dir_products(GlmWork* w, const GlmData* g, hipStream_t s, double eps, double prior_scale, const T* V,
                     const T* x0, const T* Z, uint32_t key, const T* Xi, int64_t D, int64_t Tn, const int* skip,
                     Prof* prof, int iter, std::string* msg) {
    ColSrc<T> a{};
    a.V = V;
    a.x0 = x0;
    a.Z = Z;
    a.keys[0] = key;
    a.K = 1;
    a.D = D;
    a.Tn = Tn;
    return glm_pass<T>(g, w, s, a, 0, eps, prior_scale, Xi, (T*)w->gx, (T*)w->hw, w->lpx, (T*)w->Y, (T*)w->W,
                       skip, prof, iter, false, msg);
}
template int glm_dir_products<float>(GlmWork*, const GlmData*, hipStream_t, double, double, const float*,
                                     const float*, const float*, uint32_t, const float*, int64_t, int64_t,
                                     const int*, Prof*, int, std::string*);
template int glm_dir_products<double>(GlmWork*, const GlmData*, hipStream_t, double, double, const double*,
                                      const double*, const double*, uint32_t, const double*, int64_t, int64_t,
                                      const int*, Prof*, int, std::string*);

// One pass at the states V (shifted by x0 when given) with one direction set
// Z: hout = H(V) Z, the Hessian-vector products of the log-density; the
// states' gradient and log-density go to gscr / lpscr (not needed by the
// caller, written by the pass).
template <typename T>
int glm_hvp_products(GlmWork* w, const GlmData* g, hipStream_t s, double prior_scale, const T* V, const T* x0,
                     const T* Z, T* gscr, double* lpscr, T* hout, int64_t D, int64_t Tn, const int* skip, Prof* prof,
                     int iter, bool jbf, std::string* msg) {
    ColSrc<T> a{};
    a.V = V;
    a.x0 = x0;
    a.Z = Z;
    a.K = 1;
    a.D = D;
    a.Tn = Tn;
    return glm_pass<T>(g, w, s, a, 1, 0.0, prior_scale, nullptr, gscr, hout, lpscr, nullptr, nullptr, skip, prof,
                       iter, jbf, msg);
}
template int glm_hvp_products<float>(GlmWork*, const GlmData*, hipStream_t, double, const float*, const float*,
                                     const float*, float*, double*, float*, int64_t, int64_t, const int*, Prof*, int,
                                     bool, std::string*);
template int glm_hvp_products<double>(GlmWork*, const GlmData*, hipStream_t, double, const double*, const double*,
                                      const double*, double*, double*, double*, int64_t, int64_t, const int*, Prof*,
                                      int, bool, std::string*);

// ---------------------------------------------------------------------------
// standalone batched MALA step + JVP, and log-densities
// ---------------------------------------------------------------------------
// One wave per column: logalpha, accept, FT, Jt from the two passes' products.
template <typename T>
__global__ __launch_bounds__(256) void glm_finish_kernel(const T* X, const T* Z, const T* U, const T* Y,
                                                         const T* gx, const T* W, const T* gy,
                                                         const T* hw, const double* lpx,
                                                         const double* lpy, double eps, int64_t D,
                                                         int64_t Tn, T* FT, T* Jt, uint8_t* acc_out,
                                                         double* la_out) {
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= Tn) return;
    const double c_norm = ((double)D / 2.0) * log(4.0 * 3.14159265358979323846 * eps);
    double qyx = 0, qxy = 0, d1 = 0, d2 = 0, d3 = 0;
    for (int64_t d = lane; d < D; d += 64) {
        const size_t gi = (size_t)t * D + d;
        const double x = X[gi], y = Y[gi], g0 = gx[gi], g1 = gy[gi], z = Z[gi], w = W[gi], h = hw[gi];
        const double e1 = y - x - eps * g0, r = x - y - eps * g1, dr = z - w - eps * h;
        qyx += e1 * e1;
        qxy += r * r;
        d1 += g1 * w;
        d2 += g0 * z;
        d3 += r * dr;
    }
    for (int m = 32; m >= 1; m >>= 1) {
        qyx += __shfl_xor(qyx, m, 64);
        qxy += __shfl_xor(qxy, m, 64);
        d1 += __shfl_xor(d1, m, 64);
        d2 += __shfl_xor(d2, m, 64);
        d3 += __shfl_xor(d3, m, 64);
    }
    const double la = (lpy[t] + (-0.5 * qxy / (2.0 * eps) - c_norm)) - (lpx[t] + (-0.5 * qyx / (2.0 * eps) - c_norm));
    const double lu = log((double)U[t]);
    const double g = 1.0 / (1.0 + exp(-(la - lu)));
    const bool acc = lu < la;
    const double dg = g * (1.0 - g) * (d1 - d2 - d3 / (2.0 * eps));
    if (lane == 0) {
        if (acc_out) acc_out[t] = acc ? 1 : 0;
        if (la_out) la_out[t] = la;
    }
    for (int64_t d = lane; d < D; d += 64) {
        const size_t gi = (size_t)t * D + d;
        const double x = X[gi], y = Y[gi], z = Z[gi], w = W[gi];
        FT[gi] = (T)(acc ? y : x);
        Jt[gi] = (T)((acc ? w : z) + (y - x) * dg);
    }
}

template <typename T>
void launch_fwd_jvp_finish(hipStream_t s, const T* X, const T* Z, const T* U, const GlmWork* w,
                           double eps, int64_t D, int64_t Tn, T* FT, T* Jt, uint8_t* acc,
                           double* la) {
    hipLaunchKernelGGL((glm_finish_kernel<T>), dim3((unsigned)((Tn + 3) / 4)), dim3(256), 0, s, X, Z, U,
                       (const T*)w->Y, (const T*)w->gx, (const T*)w->W, (const T*)w->gy, (const T*)w->hw,
                       w->lpx, w->lpy, eps, D, Tn, FT, Jt, acc, la);
}
template void launch_fwd_jvp_finish<float>(hipStream_t, const float*, const float*, const float*,
                                           const GlmWork*, double, int64_t, int64_t, float*, float*,
                                           uint8_t*, double*);
template void launch_fwd_jvp_finish<double>(hipStream_t, const double*, const double*, const double*,
                                            const GlmWork*, double, int64_t, int64_t, double*, double*,
                                            uint8_t*, double*);

template <typename T>
static int fwd_jvp_w(const GlmData* g, GlmWork& w, hipStream_t s, double eps, double prior_scale, const T* X,
                     const T* Xi, const T* U, const T* Z, T* FT, T* Jt, uint8_t* acc, double* la, int64_t D,
                     int64_t Tn, bool jbf, std::string* msg) {
    int rc = PMCMC_OK;
    ColSrc<T> a{};
    a.V = X;
    a.x0 = nullptr;
    a.Z = Z;
    a.K = 1;
    a.D = D;
    a.Tn = Tn;
    rc = glm_pass<T>(g, &w, s, a, 0, eps, prior_scale, Xi, (T*)w.gx, (T*)w.hw, w.lpx, (T*)w.Y, (T*)w.W, nullptr,
                     nullptr, 0, jbf, msg);
    if (rc == PMCMC_OK) {
        ColSrc<T> b = a;
        b.V = (const T*)w.Y;
        b.Z = (const T*)w.W;
        rc = glm_pass<T>(g, &w, s, b, 1, eps, prior_scale, Xi, (T*)w.gy, (T*)w.hw, w.lpy, nullptr, nullptr, nullptr,
                         nullptr, 0, jbf, msg);
    }
    if (rc == PMCMC_OK) {
        hipLaunchKernelGGL((glm_finish_kernel<T>), dim3((unsigned)((Tn + 3) / 4)), dim3(256), 0, s, X, Z, U,
                           (const T*)w.Y, (const T*)w.gx, (const T*)w.W, (const T*)w.gy, (const T*)w.hw,
                           w.lpx, w.lpy, eps, D, Tn, FT, Jt, acc, la);
        if (hipGetLastError() != hipSuccess) {
            *msg = "logistic finish kernel failed";
            rc = PMCMC_ERR_HIP;
        }
    }
    return rc;
}

template <typename T>
static int fwd_jvp_t(const GlmData* g, hipStream_t s, double eps, double prior_scale, const T* X,
                     const T* Xi, const T* U, const T* Z, T* FT, T* Jt, uint8_t* acc, double* la,
                     int64_t D, int64_t Tn, bool jbf, std::string* msg) {
    GlmWork w{};
    int rc = glm_work_ensure(&w, g, s, D, Tn, 1, sizeof(T) == 8, msg);
    if (rc == PMCMC_OK) rc = fwd_jvp_w<T>(g, w, s, eps, prior_scale, X, Xi, U, Z, FT, Jt, acc, la, D, Tn, jbf, msg);
    hipStreamSynchronize(s);
    glm_work_destroy(&w);
    return rc;
}

// K sequential MALA steps of N chains (columns): step k reads the states of
// step k-1 (X0 first) and writes column block k of S_out (D x N each); one
// work allocation for all steps, no synchronisation between them.  Z and Jw:
// D x N scratch (Z zero: the JVP half of the passes is not used).
template <typename T>
static int mala_run_t(const GlmData* g, hipStream_t s, double eps, double prior_scale, const T* X0, const T* Xi,
                      const T* U, int64_t K, int64_t D, int64_t N, T* S_out, uint8_t* acc, const T* Z, T* Jw,
                      bool jbf, std::string* msg) {
    GlmWork w{};
    int rc = glm_work_ensure(&w, g, s, D, N, 1, sizeof(T) == 8, msg);
    for (int64_t k = 0; k < K && rc == PMCMC_OK; ++k) {
        const T* X = k == 0 ? X0 : S_out + (size_t)(k - 1) * D * N;
        rc = fwd_jvp_w<T>(g, w, s, eps, prior_scale, X, Xi + (size_t)k * D * N, U + (size_t)k * N, Z,
                          S_out + (size_t)k * D * N, Jw, acc + (size_t)k * N, nullptr, D, N, jbf, msg);
    }
    hipStreamSynchronize(s);
    glm_work_destroy(&w);
    return rc;
}

int glm_mala_run(const GlmData* g, hipStream_t s, int prec, double eps, double prior_scale, const void* X0,
                 const void* Xi, const void* U, int64_t K, int64_t D, int64_t N, void* S_out, uint8_t* acc,
                 const void* Z, void* Jw, std::string* msg) {
    if (prec == PMCMC_PREC_F64)
        return mala_run_t<double>(g, s, eps, prior_scale, (const double*)X0, (const double*)Xi, (const double*)U, K,
                                  D, N, (double*)S_out, acc, (const double*)Z, (double*)Jw, false, msg);
    return mala_run_t<float>(g, s, eps, prior_scale, (const float*)X0, (const float*)Xi, (const float*)U, K, D, N,
                             (float*)S_out, acc, (const float*)Z, (float*)Jw, prec == PMCMC_PREC_BF16_JAC, msg);
}

int glm_fwd_and_jvp(const GlmData* g, hipStream_t s, int prec, double eps, double prior_scale,
                    const void* X, const void* Xi, const void* U, const void* Z, void* FT, void* Jt,
                    uint8_t* accepted, double* logalpha, int64_t D, int64_t T, std::string* msg) {
    if (prec == PMCMC_PREC_F64)
        return fwd_jvp_t<double>(g, s, eps, prior_scale, (const double*)X, (const double*)Xi,
                                 (const double*)U, (const double*)Z, (double*)FT, (double*)Jt, accepted,
                                 logalpha, D, T, false, msg);
    return fwd_jvp_t<float>(g, s, eps, prior_scale, (const float*)X, (const float*)Xi, (const float*)U,
                            (const float*)Z, (float*)FT, (float*)Jt, accepted, logalpha, D, T,
                            prec == PMCMC_PREC_BF16_JAC, msg);
}

// logp of every column: one workgroup per column, threads over n, in double.
template <typename T>
__global__ __launch_bounds__(256) void glm_logp_kernel(const T* S, const double* X64, const double* y64,
                                                       int64_t N, int64_t D, double inv_s2, const double* Pm,
                                                       double* out) {
    __shared__ double red[4];
    const int64_t t = blockIdx.x;
    const T* x = S + (size_t)t * D;
    double acc = 0.0;
    for (int64_t n = threadIdx.x; n < N; n += 256) {
        double l = 0.0;
        for (int64_t d = 0; d < D; ++d) l += X64[n + d * N] * (double)x[d];
        const double yv = y64[n];
        acc += -yv * softplusd(-l) - (1.0 - yv) * softplusd(l);
    }
    double nrm = 0.0;
    for (int64_t d = threadIdx.x; d < D; d += 256) {
        double px = (double)x[d] * inv_s2;
        if (Pm) {
            px = 0.0;
            for (int64_t k = 0; k < D; ++k) px += Pm[(size_t)d * D + k] * (double)x[k];
        }
        nrm += (double)x[d] * px;
    }
    acc -= 0.5 * nrm;
    for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[t] = red[0] + red[1] + red[2] + red[3];
}

int glm_logdensity(const GlmData* g, hipStream_t s, int prec, double prior_scale, const void* S,
                   double* out, int64_t T, std::string* msg) {
    const double inv_s2 = 1.0 / (prior_scale * prior_scale);
    if (prec == PMCMC_PREC_F64)
        hipLaunchKernelGGL((glm_logp_kernel<double>), dim3((unsigned)T), dim3(256), 0, s, (const double*)S,
                           g->Xd64, g->y64, g->N, g->D, inv_s2, g->P64, out);
    else
        hipLaunchKernelGGL((glm_logp_kernel<float>), dim3((unsigned)T), dim3(256), 0, s, (const float*)S,
                           g->Xd64, g->y64, g->N, g->D, inv_s2, g->P64, out);
    if (hipGetLastError() != hipSuccess) {
        *msg = "logistic log-density kernel failed";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}

}  // namespace pmcmc
