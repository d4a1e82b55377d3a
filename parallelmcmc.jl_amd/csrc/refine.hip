// FP64 recompute of the accept decision near the uniform threshold, for the
// GEMM targets in FP32 / BF16-Jacobian precision.
//
// The MALA accept test (src/MALA/MALA.jl:589-597) compares log u with log
// alpha.  In the FP32 path log alpha is formed in double from FP32 GEMM
// products (log-densities, gradients, the proposal Y); their rounding moves it
// by up to ~5e-5 on the logistic workloads, which decides the test only when
// log alpha is that close to log u.  Per Newton iteration:
//
//   flag    one wave per column re-forms log alpha from the FP32 products
//           (as the tile kernel does), writes dec[t] = -1 and appends the
//           columns with |log alpha - log u| < band to a compact list;
//   pass 0  FP64 gradient / log-density partials at the state Xbar_t of every
//           listed column, split over row ranges of X (logistic) or column
//           ranges of Lambda (Gaussian) so one column's evaluation is spread
//           over many workgroups;
//   fin 0   fixed-order sum of the partials, FP64 proposal
//           y = x + eps g(x) + sqrt(2 eps) xi;
//   pass 1  the same partials at y;
//   fin 1   FP64 log alpha and the decision (dec[t] bit 0; bit 1 flags a
//           margin below the tile kernel's hysteresis), which overrides the
//           FP32 decision in the tile kernel.
//
// The state x is the iterate's FP32 value, exact in double; the target data
// are the FP64 copies.  Each pass reads one counter, so with nothing flagged
// the four FP64 launches exit at once.
//
// Decisions matter only at the fixed point, so the FP64 work is gated: it
// starts once the previous iteration's update metric is within `gate`
// tolerance units, or at iteration `gate_iters` (FP32 decisions of columns
// closer to the threshold than their ~1e-5 error flip as the iterate's last
// ulps settle, which can keep the metric up for ever), and then stays on (a
// gate that opened and closed could alternate FP32 and FP64 decisions of one
// column for ever).  The tile kernel
// counts an iteration as converged only once the gate is on, so the returned
// trajectory always carries FP64-checked decisions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/pmcmc_amd.h"
#include "refine.hpp"

namespace pmcmc {

static constexpr int RNT = 256;
static constexpr int RMAXD = 256;
// list capacity beyond the expected count: as many columns as 32 MiB of partials hold
static constexpr int64_t REFINE_PART_BUDGET = 32ll << 20;

__device__ inline double softplus_r(double x) { return fmax(x, 0.0) + log1p(exp(-fabs(x))); }
__device__ inline double sigmoid_r(double x) {
    if (x >= 0.0) return 1.0 / (1.0 + exp(-x));
    const double e = exp(x);
    return e / (1.0 + e);
}

__device__ inline double wave_sum_r(double v) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// sum over the 256-thread block; every thread gets the total.  The trailing
// barrier also orders all earlier LDS reads before the caller's next writes.
__device__ inline double block_sum(double v, double* red) {
    v = wave_sum_r(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const double s = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return s;
}

// Xbar_t[d]: x0 for t = 0, else column t-1 of the iterate
__device__ inline double state_at(const RefineArgs& a, int64_t t, int64_t d, int64_t D) {
    return t == 0 ? (double)a.x0[d] : (double)a.S_in[(size_t)(t - 1) * D + d];
}

__device__ inline bool skipped(const RefineArgs& a) { return a.skip && *(volatile const int*)a.skip; }

// log-density of column t from its slab partials, summed in slab order
__device__ inline double lp_at(const RefineArgs& a, const double* lp, int64_t t) {
    double v = lp[t];
    for (int s = 1; s < a.lp_nslab; ++s) v += lp[(size_t)s * a.lp_stride + t];
    return v;
}

__device__ inline int iter_of(const RefineArgs& a) { return a.iter_dev ? *a.iter_dev + 1 : a.iter; }

// log alpha of column t from the FP32 products (one wave); lists t when it is
// within the band of log u
__device__ void refine_flag_column(const RefineArgs& a, int8_t* dec, int* cnt, int* list, int lmax, int par,
                                   int64_t D, int64_t t, int lane) {
    const double eps = a.eps;
    double qyx = 0.0, qxy = 0.0;
    if (D <= 256) {
        // at most four rows per lane: every load of the column in flight at once
        float xv[4], yv[4], gxv[4], gyv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t d = lane + 64 * i;
            if (d < D) {
                const size_t gi = (size_t)t * D + d;
                xv[i] = t == 0 ? a.x0[d] : a.S_in[(size_t)(t - 1) * D + d];
                yv[i] = a.Y[gi];
                gxv[i] = a.gx[gi];
                gyv[i] = a.gy[gi];
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (lane + 64 * i < D) {
                const double x = (double)xv[i], y = (double)yv[i];
                const double e1 = y - x - eps * (double)gxv[i];
                const double e2 = x - y - eps * (double)gyv[i];
                qyx += e1 * e1;
                qxy += e2 * e2;
            }
        }
    } else {
        for (int64_t d = lane; d < D; d += 64) {
            const size_t gi = (size_t)t * D + d;
            const double x = state_at(a, t, d, D);
            const double y = (double)a.Y[gi];
            const double e1 = y - x - eps * (double)a.gx[gi];
            const double e2 = x - y - eps * (double)a.gy[gi];
            qyx += e1 * e1;
            qxy += e2 * e2;
        }
    }
    qyx = wave_sum_r(qyx);
    qxy = wave_sum_r(qxy);
    if (lane == 0) {
        const double c_norm = ((double)D / 2.0) * log(4.0 * 3.14159265358979323846 * eps);
        const double la = (lp_at(a, a.lpy, t) + (-0.5 * qxy / (2.0 * eps) - c_norm)) -
                          (lp_at(a, a.lpx, t) + (-0.5 * qyx / (2.0 * eps) - c_norm));
        const double lu = log((double)a.U[t]);
        dec[t] = -1;
        if (fabs(la - lu) < a.band) {
            if (list) {
                const int j = atomicAdd(&cnt[par], 1);
                if (j < lmax) list[j] = (int)t;
            } else {
                dec[t] = -2;  // flagged for the caller's own FP64 decision
            }
        }
    }
}

// Columns FLAG_COLS blk .. FLAG_COLS (blk + 1) - 1, one wave per column at a
// time (a workgroup per handful of columns made the launch dispatch-bound).
static constexpr int FLAG_COLS = 32;
__device__ void refine_flag_body(const RefineArgs& a, int8_t* dec, int* cnt, int* list, int lmax, int par, int64_t D,
                                 int64_t T, int64_t blk) {
    if (skipped(a)) return;
    // every block reaches the same answer: cnt[2] only turns on when the gate
    // condition (fixed during this launch) holds
    const bool on = cnt[2] != 0 || *a.last_metric <= a.gate || iter_of(a) >= a.gate_iters;
    if (blk == 0 && threadIdx.x == 0) {
        cnt[par ^ 1] = 0;
        if (on) cnt[2] = 1;
    }
    const int lane = threadIdx.x & 63;
    const int64_t t_end = (blk + 1) * FLAG_COLS < T ? (blk + 1) * FLAG_COLS : T;
    const int64_t lo = a.final_cols ? (int64_t)*(volatile const int*)a.final_cols : 0;
    for (int64_t t = blk * FLAG_COLS + (threadIdx.x >> 6); t < t_end; t += RNT / 64) {
        if (t < lo) continue;  // final column: its decision is settled
        if (!on) {
            if (lane == 0) dec[t] = -1;
            continue;
        }
        refine_flag_column(a, dec, cnt, list, lmax, par, D, t, lane);
    }
}

// Logistic partials of listed column j at v = Xbar_t (phase 0) or at its FP64
// proposal (phase 1), over rows [split * rows, ...) of X: z = X v,
// r = y - sigmoid(z), grad partial X^T r, log-likelihood partial.  A block
// takes 64-row chunks (lane = row); wave w owns the columns d = w + 4q of X,
// holds that chunk's X[n, d] in registers (all Q loads in flight at once) for
// both products, and reduces its X^T r accumulators across lanes once at the
// end.  Q >= ceil(D / 4).
template <int Q>
__device__ __forceinline__ void refine_pass_glm_body(const RefineTarget& tg, const RefineArgs& a, const RefineWork& w, int par,
                                     int phase, int64_t split, int slot, int nslots) {
    if (skipped(a) || split >= w.nsplit) return;
    __shared__ double v[RMAXD], zpart[RNT / 64][64], red[4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t D = w.D, N = tg.N;
    const int n_ent = min(w.cnt[par], w.lmax);
    const int64_t r0 = split * w.rows;
    const int64_t r1 = std::min(N, r0 + w.rows);
    for (int j = slot; j < n_ent; j += nslots) {
        const int64_t t = w.list[j];
        for (int64_t d = tid; d < RMAXD; d += RNT)
            v[d] = d >= D ? 0.0 : (phase == 0 ? state_at(a, t, d, D) : w.y[(size_t)j * D + d]);
        __syncthreads();
        double gacc[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) gacc[q] = 0.0;
        double ll = 0.0;
        for (int64_t n0 = r0; n0 < r1; n0 += 64) {
            const int64_t n = n0 + lane;
            const bool row = n < r1;
            const int64_t nc = row ? n : r0;  // padding lanes read a valid row, weight 0
            double xr[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int64_t d = wv + 4 * q;
                const int64_t dc = d < D ? d : 0;
                xr[q] = tg.X64[nc + dc * N] * (d < D ? 1.0 : 0.0);
            }
            double zp = 0.0;
#pragma unroll
            for (int q = 0; q < Q; ++q) zp += xr[q] * v[wv + 4 * q];
            zpart[wv][lane] = zp;
            __syncthreads();
            const double z = (zpart[0][lane] + zpart[1][lane]) + (zpart[2][lane] + zpart[3][lane]);
            double r = 0.0;
            if (row) {
                const double yv = tg.y64[n];
                r = yv - sigmoid_r(z);
                if (wv == 0) ll += -yv * softplus_r(-z) - (1.0 - yv) * softplus_r(z);
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) gacc[q] += xr[q] * r;
            __syncthreads();  // zpart is rewritten by the next chunk
        }
        double* out = w.part + ((size_t)j * w.nsplit + split) * (D + 1);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double sq = wave_sum_r(gacc[q]);
            const int64_t d = wv + 4 * q;
            if (lane == 0 && d < D) out[d] = sq;
        }
        const double sl = block_sum(ll, red);
        if (tid == 0) out[D] = sl;
    }
}

// Gaussian partials over columns [split * rows, ...) of Lambda:
// grad partial -Lambda[:, k0:k1] (v - mu)[k0:k1]; the log-density
// 0.5 (v - mu).grad is linear in the partials
__device__ void refine_pass_gauss_body(const RefineTarget& tg, const RefineArgs& a, const RefineWork& w, int par,
                                       int phase, int64_t split, int slot, int nslots) {
    if (skipped(a) || split >= w.nsplit) return;
    __shared__ double v[RMAXD], red[4];
    const int tid = threadIdx.x;
    const int64_t D = w.D;
    const int n_ent = min(w.cnt[par], w.lmax);
    const int64_t k0 = split * w.rows;
    const int64_t k1 = std::min(D, k0 + w.rows);
    for (int j = slot; j < n_ent; j += nslots) {
        const int64_t t = w.list[j];
        for (int64_t d = tid; d < D; d += RNT)
            v[d] = (phase == 0 ? state_at(a, t, d, D) : w.y[(size_t)j * D + d]) - (tg.mu64 ? tg.mu64[d] : 0.0);
        __syncthreads();
        double* out = w.part + ((size_t)j * w.nsplit + split) * (D + 1);
        double lpp = 0.0;
        for (int64_t d = tid; d < D; d += RNT) {
            double p = 0.0;
            for (int64_t k = k0; k < k1; ++k) p += tg.L64[d + k * D] * v[k];
            out[d] = -p;
            lpp += 0.5 * v[d] * (-p);
        }
        const double sl = block_sum(lpp, red);
        if (tid == 0) out[D] = sl;
    }
}

// fixed-order sum of the partials of listed column j (wave w takes the splits
// w, w + 4, ..., eight loads in flight; the four waves combine in order);
// phase 0 forms the FP64 proposal, phase 1 the FP64 log alpha and the decision
__device__ void refine_finish_body(const RefineTarget& tg, const RefineArgs& a, const RefineWork& w, int par,
                                   int phase, int slot, int nslots) {
    if (skipped(a)) return;
    __shared__ double gs[RMAXD + 1], wpart[RNT / 64][64], red[4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t D = w.D;
    const int ns = w.nsplit;
    const int n_ent = min(w.cnt[par], w.lmax);
    const double eps = a.eps;
    const double sq2e = sqrt(2.0 * eps);
    for (int j = slot; j < n_ent; j += nslots) {
        const int64_t t = w.list[j];
        const double* pj = w.part + (size_t)j * ns * (D + 1);
        for (int64_t dc = 0; dc <= D; dc += 64) {  // index D is the log-density term
            const int64_t d = dc + lane;
            double acc = 0.0;
            if (d <= D) {
                int s = wv;
                for (; s + 28 < ns; s += 32) {
                    double p[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) p[u] = pj[(size_t)(s + 4 * u) * (D + 1) + d];
#pragma unroll
                    for (int u = 0; u < 8; ++u) acc += p[u];
                }
                for (; s < ns; s += 4) acc += pj[(size_t)s * (D + 1) + d];
            }
            wpart[wv][lane] = acc;
            __syncthreads();
            if (wv == 0 && d <= D) gs[d] = (wpart[0][lane] + wpart[1][lane]) + (wpart[2][lane] + wpart[3][lane]);
            __syncthreads();
        }
        double lp = gs[D];
        double nrm = 0.0, qyx = 0.0, qxy = 0.0;
        for (int64_t d = tid; d < D; d += RNT) {
            double g = gs[d];
            const double x = state_at(a, t, d, D);
            const size_t jd = (size_t)j * D + d;
            if (phase == 0) {
                if (tg.kind == 0) {
                    // prior: P x (dense precision) or x / s^2
                    double px = x * tg.inv_s2;
                    if (tg.P64) {
                        px = 0.0;
                        for (int64_t k = 0; k < D; ++k) px += tg.P64[(size_t)d * D + k] * state_at(a, t, k, D);
                    }
                    g -= px;
                    nrm += x * px;
                }
                w.g0[jd] = g;
                w.y[jd] = x + eps * g + sq2e * (double)a.Xi[(size_t)t * D + d];
            } else {
                const double y = w.y[jd];
                if (tg.kind == 0) {
                    double py = y * tg.inv_s2;
                    if (tg.P64) {
                        py = 0.0;
                        for (int64_t k = 0; k < D; ++k) py += tg.P64[(size_t)d * D + k] * w.y[(size_t)j * D + k];
                    }
                    g -= py;
                    nrm += y * py;
                }
                const double e1 = y - x - eps * w.g0[jd];
                const double e2 = x - y - eps * g;
                qyx += e1 * e1;
                qxy += e2 * e2;
            }
        }
        lp -= 0.5 * block_sum(nrm, red);  // x' P x (x' x / s^2 isotropic); 0 for the Gaussian
        if (phase == 0) {
            if (tid == 0) w.lp0[j] = lp;
        } else {
            qyx = block_sum(qyx, red);
            qxy = block_sum(qxy, red);
            if (tid == 0) {
                const double c_norm = ((double)D / 2.0) * log(4.0 * 3.14159265358979323846 * eps);
                const double la = (lp + (-0.5 * qxy / (2.0 * eps) - c_norm)) -
                                  (w.lp0[j] + (-0.5 * qyx / (2.0 * eps) - c_norm));
                const double lu = log((double)a.U[t]);
                w.dec[t] = (int8_t)((lu < la ? 1 : 0) | (fabs(la - lu) < a.hyst ? 2 : 0));
            }
        }
    }
}


__global__ __launch_bounds__(RNT) void refine_flag_kernel(RefineArgs a, RefineWork w, int par) {
    refine_flag_body(a, w.dec, w.cnt, w.list, w.lmax, par, w.D, w.T, blockIdx.x);
}
template <int Q>
__global__ __launch_bounds__(RNT) void refine_pass_glm_kernel(RefineTarget tg, RefineArgs a, RefineWork w, int par,
                                                             int phase) {
    refine_pass_glm_body<Q>(tg, a, w, par, phase, blockIdx.x, blockIdx.y, gridDim.y);
}
__global__ __launch_bounds__(RNT) void refine_pass_gauss_kernel(RefineTarget tg, RefineArgs a, RefineWork w, int par,
                                                               int phase) {
    refine_pass_gauss_body(tg, a, w, par, phase, blockIdx.x, blockIdx.y, gridDim.y);
}
__global__ __launch_bounds__(RNT) void refine_finish_kernel(RefineTarget tg, RefineArgs a, RefineWork w, int par,
                                                           int phase) {
    refine_finish_body(tg, a, w, par, phase, blockIdx.x, gridDim.x);
}

// batched: blockIdx.z picks the chain from a live-chain list
__global__ __launch_bounds__(RNT) void refine_flag_batched(const RefineArgs* ras, const RefineWork* rws,
                                                          const int* ids, int par) {
    const int c = ids[blockIdx.z];
    const RefineWork& w = rws[c];
    if ((int64_t)blockIdx.x * FLAG_COLS >= w.T) return;
    refine_flag_body(ras[c], w.dec, w.cnt, w.list, w.lmax, par, w.D, w.T, blockIdx.x);
}
__global__ __launch_bounds__(RNT) void refine_pass_glm_batched(const RefineTarget* tgs, const RefineArgs* ras,
                                                              const RefineWork* rws, const int* ids, int par,
                                                              int phase) {
    const int c = ids[blockIdx.z];
    refine_pass_glm_body<RMAXD / 4>(tgs[c], ras[c], rws[c], par, phase, blockIdx.x, blockIdx.y, gridDim.y);
}
__global__ __launch_bounds__(RNT) void refine_pass_gauss_batched(const RefineTarget* tgs, const RefineArgs* ras,
                                                                const RefineWork* rws, const int* ids, int par,
                                                                int phase) {
    const int c = ids[blockIdx.z];
    refine_pass_gauss_body(tgs[c], ras[c], rws[c], par, phase, blockIdx.x, blockIdx.y, gridDim.y);
}
__global__ __launch_bounds__(RNT) void refine_finish_batched(const RefineTarget* tgs, const RefineArgs* ras,
                                                            const RefineWork* rws, const int* ids, int par,
                                                            int phase) {
    const int c = ids[blockIdx.z];
    refine_finish_body(tgs[c], ras[c], rws[c], par, phase, blockIdx.x, gridDim.x);
}

// Dense-Gaussian chains of a batched solve: flag and FP64 decision in one
// launch.  A workgroup owns DIRECT_COLS columns: it flags them from the FP32
// products (one wave per column) and then decides each flagged column in
// FP64 itself, the two D x D mat-vecs of Lambda (x and the FP64 proposal) one
// row per thread, so no list, split partials or further launches are needed.
// The same FP64 quantities as the split path, summed in another order.
static constexpr int DIRECT_COLS = 8;
__global__ __launch_bounds__(RNT) void refine_gauss_direct_batched(const RefineTarget* tgs, const RefineArgs* ras,
                                                                  const RefineWork* rws, const int* ids, int par) {
    const int c = ids[blockIdx.z];
    const RefineArgs& a = ras[c];
    const RefineWork& w = rws[c];
    const RefineTarget& tg = tgs[c];
    if (skipped(a)) return;
    const int64_t T = w.T, D = w.D;
    const int64_t t0 = (int64_t)blockIdx.x * DIRECT_COLS;
    if (t0 >= T) return;
    const int64_t t1 = t0 + DIRECT_COLS < T ? t0 + DIRECT_COLS : T;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // the gate (fixed during this launch), as in refine_flag_body
    const bool on = w.cnt[2] != 0 || *a.last_metric <= a.gate || iter_of(a) >= a.gate_iters;
    if (blockIdx.x == 0 && tid == 0 && on) w.cnt[2] = 1;
    __shared__ double v[RMAXD], g0[RMAXD], yv[RMAXD], red[4];
    __shared__ int flagged[DIRECT_COLS];
    if (tid < DIRECT_COLS) flagged[tid] = 0;
    __syncthreads();
    const int64_t lo = a.final_cols ? (int64_t)*(volatile const int*)a.final_cols : 0;
    for (int64_t t = t0 + wv; t < t1; t += RNT / 64) {
        if (t < lo) continue;
        if (!on) {
            if (lane == 0) w.dec[t] = -1;
            continue;
        }
        refine_flag_column(a, w.dec, nullptr, nullptr, 0, par, D, t, lane);
        if (lane == 0) flagged[t - t0] = w.dec[t] == -2;  // the lane's own store
    }
    __syncthreads();
    const double eps = a.eps;
    const double sq2e = sqrt(2.0 * eps);
    for (int64_t t = t0; t < t1; ++t) {
        if (!flagged[t - t0]) continue;  // the same for every thread of the block
        for (int64_t d = tid; d < D; d += RNT) v[d] = state_at(a, t, d, D) - (tg.mu64 ? tg.mu64[d] : 0.0);
        __syncthreads();
        double lpp = 0.0;
        for (int64_t d = tid; d < D; d += RNT) {
            double p = 0.0;
            for (int64_t k = 0; k < D; ++k) p += tg.L64[d + k * D] * v[k];
            const double g = -p;
            g0[d] = g;
            lpp += 0.5 * v[d] * g;
            yv[d] = state_at(a, t, d, D) + eps * g + sq2e * (double)a.Xi[(size_t)t * D + d];
        }
        const double lp0 = block_sum(lpp, red);  // its barriers also order the reads of v before the rewrite
        for (int64_t d = tid; d < D; d += RNT) v[d] = yv[d] - (tg.mu64 ? tg.mu64[d] : 0.0);
        __syncthreads();
        double lq = 0.0, qyx = 0.0, qxy = 0.0;
        for (int64_t d = tid; d < D; d += RNT) {
            double p = 0.0;
            for (int64_t k = 0; k < D; ++k) p += tg.L64[d + k * D] * v[k];
            const double g = -p;
            lq += 0.5 * v[d] * g;
            const double x = state_at(a, t, d, D), y = yv[d];
            const double e1 = y - x - eps * g0[d];
            const double e2 = x - y - eps * g;
            qyx += e1 * e1;
            qxy += e2 * e2;
        }
        const double lp1 = block_sum(lq, red);
        qyx = block_sum(qyx, red);
        qxy = block_sum(qxy, red);
        if (tid == 0) {
            const double c_norm = ((double)D / 2.0) * log(4.0 * 3.14159265358979323846 * eps);
            const double la = (lp1 + (-0.5 * qxy / (2.0 * eps) - c_norm)) - (lp0 + (-0.5 * qyx / (2.0 * eps) - c_norm));
            const double lu = log((double)a.U[t]);
            w.dec[t] = (int8_t)((lu < la ? 1 : 0) | (fabs(la - lu) < a.hyst ? 2 : 0));
        }
        __syncthreads();
    }
}

__global__ void refine_zero_kernel(int* cnt) {
    if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
}

void refine_work_destroy(RefineWork* w) {
    hipFree(w->dec);
    hipFree(w->cnt);
    hipFree(w->list);
    hipFree(w->part);
    hipFree(w->g0);
    hipFree(w->y);
    hipFree(w->lp0);
    *w = RefineWork{};
}

int refine_work_ensure(RefineWork* w, const RefineTarget& tg, int64_t T, double band, std::string* msg) {
    if (tg.D < 1 || tg.D > RMAXD || T < 1 || (tg.kind == 0 && tg.N < 1)) {
        *msg = "decision refinement: unsupported target shape";
        return PMCMC_ERR_UNSUPPORTED;
    }
    int64_t rows, nsplit;
    if (tg.kind == 0) {
        // 64-row chunks, at most 64 splits
        rows = 64 * std::max<int64_t>(1, (tg.N + 64 * 64 - 1) / (64 * 64));
        nsplit = (tg.N + rows - 1) / rows;
    } else {
        rows = 32;
        nsplit = (tg.D + rows - 1) / rows;
    }
    const int64_t per_entry = nsplit * (tg.D + 1) * (int64_t)sizeof(double);
    const int64_t lm = std::min(T, std::max((int64_t)std::ceil(16.0 * band * (double)T), REFINE_PART_BUDGET / per_entry));
    const int64_t N = tg.kind == 0 ? tg.N : 0;
    if (w->dec && w->D == tg.D && w->T == T && w->N == N && w->kind == tg.kind && w->lmax == lm &&
        w->nsplit == nsplit)
        return PMCMC_OK;
    refine_work_destroy(w);
    bool ok = hipMalloc(&w->dec, (size_t)T) == hipSuccess;
    ok = ok && hipMalloc(&w->cnt, 4 * sizeof(int)) == hipSuccess;
    ok = ok && hipMalloc(&w->list, (size_t)lm * sizeof(int)) == hipSuccess;
    ok = ok && hipMalloc(&w->part, (size_t)lm * nsplit * (tg.D + 1) * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&w->g0, (size_t)lm * tg.D * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&w->y, (size_t)lm * tg.D * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&w->lp0, (size_t)lm * sizeof(double)) == hipSuccess;
    if (!ok) {
        refine_work_destroy(w);
        *msg = "decision refinement: buffer allocation failed";
        return PMCMC_ERR_ALLOC;
    }
    w->D = tg.D;
    w->T = T;
    w->N = N;
    w->kind = tg.kind;
    w->lmax = (int)lm;
    w->nsplit = (int)nsplit;
    w->rows = rows;
    return PMCMC_OK;
}

int refine_reset(RefineWork* w, hipStream_t s, std::string* msg) {
    hipLaunchKernelGGL(refine_zero_kernel, dim3(1), dim3(64), 0, s, w->cnt);
    if (hipGetLastError() != hipSuccess) {
        *msg = "decision refinement: reset launch failed";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}

int refine_decisions(RefineWork* w, const RefineTarget& tg, const RefineArgs& a, int iter, hipStream_t s,
                     std::string* msg) {
    const int par = iter & 1;
    hipLaunchKernelGGL(refine_flag_kernel, dim3((unsigned)((w->T + FLAG_COLS - 1) / FLAG_COLS)), dim3(RNT),
                       0, s, a, *w, par);
    // 16 slots of listed columns (2 band T ~ 8 expected at T = 4096, band 1e-3);
    // more are taken in turn by the same workgroups
    const unsigned gy = (unsigned)std::min(w->lmax, 16);
    const dim3 pg((unsigned)w->nsplit, gy);
    for (int phase = 0; phase < 2; ++phase) {
        if (tg.kind == 1)
            hipLaunchKernelGGL(refine_pass_gauss_kernel, pg, dim3(RNT), 0, s, tg, a, *w, par, phase);
        else if (w->D <= 32)
            hipLaunchKernelGGL(refine_pass_glm_kernel<8>, pg, dim3(RNT), 0, s, tg, a, *w, par, phase);
        else if (w->D <= 64)
            hipLaunchKernelGGL(refine_pass_glm_kernel<16>, pg, dim3(RNT), 0, s, tg, a, *w, par, phase);
        else if (w->D <= 128)
            hipLaunchKernelGGL(refine_pass_glm_kernel<32>, pg, dim3(RNT), 0, s, tg, a, *w, par, phase);
        else
            hipLaunchKernelGGL(refine_pass_glm_kernel<64>, pg, dim3(RNT), 0, s, tg, a, *w, par, phase);
        hipLaunchKernelGGL(refine_finish_kernel, dim3(gy), dim3(RNT), 0, s, tg, a, *w, par, phase);
    }
    if (hipGetLastError() != hipSuccess) {
        *msg = "decision refinement kernels failed to launch";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}

// ---------------------------------------------------------------------------
// batched solves: one chain's buffers carved from the batch arena
// ---------------------------------------------------------------------------
static void chain_plan(const RefineTarget& tg, int64_t L, double band, int64_t* lm, int64_t* rows, int64_t* nsplit) {
    if (tg.kind == 0) {
        *rows = 64 * std::max<int64_t>(1, (tg.N + 64 * 64 - 1) / (64 * 64));
        *nsplit = (tg.N + *rows - 1) / *rows;
    } else {
        *rows = 32;
        *nsplit = (tg.D + *rows - 1) / *rows;
    }
    *lm = std::min(L, std::max<int64_t>((int64_t)std::ceil(16.0 * band * (double)L), 64));
}

static size_t align256r(size_t x) { return (x + 255) / 256 * 256; }

size_t refine_chain_bytes(const RefineTarget& tg, int64_t L, double band) {
    int64_t lm, rows, nsplit;
    chain_plan(tg, L, band, &lm, &rows, &nsplit);
    return align256r((size_t)L) + align256r(4 * sizeof(int)) + align256r((size_t)lm * sizeof(int)) +
           align256r((size_t)lm * nsplit * (tg.D + 1) * sizeof(double)) +
           2 * align256r((size_t)lm * tg.D * sizeof(double)) + align256r((size_t)lm * sizeof(double));
}

void refine_chain_carve(RefineWork* w, const RefineTarget& tg, int64_t L, double band, char* base) {
    int64_t lm, rows, nsplit;
    chain_plan(tg, L, band, &lm, &rows, &nsplit);
    char* q = base;
    auto take = [&](size_t n) {
        char* r = q;
        q += align256r(n);
        return r;
    };
    *w = RefineWork{};
    w->D = tg.D;
    w->T = L;
    w->N = tg.kind == 0 ? tg.N : 0;
    w->kind = tg.kind;
    w->lmax = (int)lm;
    w->nsplit = (int)nsplit;
    w->rows = rows;
    w->dec = (int8_t*)take((size_t)L);
    w->cnt = (int*)take(4 * sizeof(int));
    w->list = (int*)take((size_t)lm * sizeof(int));
    w->part = (double*)take((size_t)lm * nsplit * (tg.D + 1) * sizeof(double));
    w->g0 = (double*)take((size_t)lm * tg.D * sizeof(double));
    w->y = (double*)take((size_t)lm * tg.D * sizeof(double));
    w->lp0 = (double*)take((size_t)lm * sizeof(double));
}

void refine_decisions_batched(const RefineBatch& b, int par, hipStream_t s) {
    // a handful of columns is listed per chain and iteration: two slots per
    // (chain, split), so that empty workgroups do not dominate the launches
    const unsigned slots = 2;
    // logistic chains: flag pass, then the FP64 gradients over row splits of X
    if (b.n_glm > 0) {
        hipLaunchKernelGGL(refine_flag_batched,
                           dim3((unsigned)((b.max_T + FLAG_COLS - 1) / FLAG_COLS), 1, (unsigned)b.n_glm), dim3(RNT), 0,
                           s, b.ras, b.rws, b.glm, par);
        for (int phase = 0; phase < 2; ++phase) {
            hipLaunchKernelGGL(refine_pass_glm_batched, dim3((unsigned)b.max_split_glm, slots, (unsigned)b.n_glm),
                               dim3(RNT), 0, s, b.tgs, b.ras, b.rws, b.glm, par, phase);
            hipLaunchKernelGGL(refine_finish_batched, dim3(slots, 1, (unsigned)b.n_glm), dim3(RNT), 0, s, b.tgs,
                               b.ras, b.rws, b.glm, par, phase);
        }
    }
    // dense-Gaussian chains: flag and FP64 decision in one launch
    if (b.n_gauss > 0)
        hipLaunchKernelGGL(refine_gauss_direct_batched,
                           dim3((unsigned)((b.max_T + DIRECT_COLS - 1) / DIRECT_COLS), 1, (unsigned)b.n_gauss),
                           dim3(RNT), 0, s, b.tgs, b.ras, b.rws, b.gauss, par);
}

}  // namespace pmcmc
