// Adjoint of the taped MALA recursion at a converged trajectory, accept
// decisions held fixed: the reverse recurrence
//   lambda_t = G_t + J_{t+1}^T lambda_{t+1},  lambda_{T-1} = G_{T-1},
//   J_t = a_t (I + eps H(Xbar_t)) + (1 - a_t) I   (H symmetric: J^T = J),
// is solved like a DEER iteration run backwards in time.  Each sweep forms
// the residual rho_t = G_t + J_{t+1}^T lambda_{t+1} - lambda_t from one batched
// HVP pass at every state, scans Delta_t = A_{t+1} Delta_{t+1} + rho_t with a
// diagonal A (the Hutchinson estimate of diag J, or the exact diagonal for
// diagonal targets), and sets lambda += Delta.  The recurrence is linear, so
// the fixed point is exact for any A; the reverse scan runs as a forward scan
// over the time-reversed columns.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace pmcmc {

// Diagonal-Gaussian products: gx = -prec (x - mean) at the states (V, shifted
// by x0 when given), W = dir - eps prec dir with dir = Z or the probe `key`.
template <typename T>
__global__ void adj_diag_products_kernel(const T* V, const T* x0, const T* Z, uint32_t key, const T* prec,
                                         const T* mean, double eps, int64_t D, int64_t Tn, T* gx, T* W,
                                         const int* skip) {
    if (skip && *(volatile const int*)skip) return;
    const int64_t n = D * Tn;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = i % D, t = i / D;
        const double x = x0 ? (t == 0 ? (double)x0[d] : (double)V[i - D]) : (double)V[i];
        const double pr = prec ? (double)prec[d] : 1.0;
        const double mu = mean ? (double)mean[d] : 0.0;
        gx[i] = (T)(-pr * (x - mu));
        const double dir = Z ? (double)Z[i] : (double)rademacher(key, (uint64_t)i);
        W[i] = (T)(dir - eps * pr * dir);
    }
}

// Reversed-time scan coefficients: Ar[:, T-1-t] = A_{t+1} (t < T-1), 0 at t = T-1.
// exact: A = a(1 - eps prec) + (1 - a) from prec; otherwise the Hutchinson
// estimate a (z o W_z) + (1 - a) with W_z = z + eps H z from the probe pass.
// Also clears lambda (D x (T+1)).
template <typename T>
__global__ void adj_coef_kernel(const T* Wz, uint32_t key, const T* prec, bool exact, double eps,
                                const uint8_t* acc, int64_t D, int64_t Tn, T* Ar, T* lam) {
    const int64_t n = D * Tn;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n + D;
         i += (int64_t)gridDim.x * blockDim.x) {
        lam[i] = (T)0;
        if (i >= n) continue;
        const int64_t d = i % D, t = i / D;
        double A = 0.0;
        if (t < Tn - 1) {
            const bool a = acc[t + 1] != 0;
            if (!a) {
                A = 1.0;
            } else if (exact) {
                A = 1.0 - eps * (prec ? (double)prec[d] : 1.0);
            } else {
                A = (double)rademacher(key, (uint64_t)i) * (double)Wz[i];
            }
        }
        Ar[(Tn - 1 - t) * D + d] = (T)A;
    }
}

// Br[:, T-1-t] = G_t + J_{t+1}^T lambda_{t+1} - lambda_t, with
// J_{t+1}^T lambda_{t+1} = W_lam[:, t] if a_{t+1} else lambda_{t+1} (0 at t = T-1).
template <typename T>
__global__ void adj_resid_kernel(const T* G, const T* Wl, const T* lam, const uint8_t* acc, int64_t D,
                                 int64_t Tn, T* Br, const int* skip) {
    if (skip && *(volatile const int*)skip) return;
    const int64_t n = D * Tn;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = i % D, t = i / D;
        double jl = 0.0;
        if (t < Tn - 1) jl = acc[t + 1] ? (double)Wl[i] : (double)lam[i + D];
        Br[(Tn - 1 - t) * D + d] = (T)((double)G[i] + jl - (double)lam[i]);
    }
}

// lambda_t += Delta_t (Delta from the reversed scan); convergence maxima and,
// in the last block, the metric (same rule as the forward solve).
template <typename T>
__global__ __launch_bounds__(256) void adj_update_kernel(const T* Dr, T* lam, int64_t D, int64_t Tn, SyncState* sy,
                                                         int iter, double tol_abs, double tol_rel) {
    __shared__ double red[2][4];
    if (*(volatile const int*)&sy->converged) return;
    const int64_t n = D * Tn;
    double dmax = 0.0, smax = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = i % D, t = i / D;
        const T dl = Dr[(Tn - 1 - t) * D + d];
        const T nv = lam[i] + dl;
        lam[i] = nv;
        dmax = nan_max(dmax, fabs((double)dl));
        smax = nan_max(smax, fabs((double)nv));
    }
    for (int m = 32; m >= 1; m >>= 1) {
        dmax = nan_max(dmax, __shfl_xor(dmax, m, 64));
        smax = nan_max(smax, __shfl_xor(smax, m, 64));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wave] = dmax;
        red[1][wave] = smax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m0 = 0.0, m1 = 0.0;
        for (int w = 0; w < 4; ++w) {
            m0 = nan_max(m0, red[0][w]);
            m1 = nan_max(m1, red[1][w]);
        }
        double dm, sm;
        if (sync_arrive(sy, gridDim.x, m0, m1, -1, &dm, &sm)) {
            const double metric = dm / (tol_abs + tol_rel * sm);
            sy->last_metric = metric;
            sy->last_iter = iter;
            if (metric <= 1.0) {
                sy->conv_iter = iter;
                sy->converged = 1;
            }
        }
    }
}

// Per-column gradient terms (one wave per column), with Xbar_t = [x0, S][:, t]:
//   colpart[t]     = a_t lambda_t . (grad(Xbar_t) + xi_t / sqrt(2 eps))
//   colpart[T + t] = a_t eps lambda_t . dgrad/dhyper(Xbar_t), where for the
//                    logistic prior N(0, s^2 I) dgrad/ds = 2 Xbar / s^3
// and dx0 = J_0^T lambda_0 = (a_0 ? W[:, 0] : lambda_0), W = lambda + eps H(Xbar) lambda.
template <typename T>
__global__ __launch_bounds__(256) void adj_grad_kernel(const T* S, const T* x0, const T* Xi, const T* gx,
                                                       const T* W, const T* lam, const uint8_t* acc, double eps,
                                                       int hyper, double scale, int64_t D, int64_t Tn,
                                                       double* colpart, T* dx0) {
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= Tn) return;
    const bool a = acc[t] != 0;
    const double isq = 1.0 / sqrt(2.0 * eps);
    const double hs = hyper ? 2.0 * eps / (scale * scale * scale) : 0.0;
    double se = 0.0, sh = 0.0;
    for (int64_t d = lane; d < D; d += 64) {
        const size_t i = (size_t)t * D + d;
        const double l = (double)lam[i];
        const double xb = t == 0 ? (double)x0[d] : (double)S[i - D];
        se += l * ((double)gx[i] + (double)Xi[i] * isq);
        sh += l * xb * hs;
        if (t == 0 && dx0) dx0[d] = a ? W[i] : lam[i];
    }
    for (int m = 32; m >= 1; m >>= 1) {
        se += __shfl_xor(se, m, 64);
        sh += __shfl_xor(sh, m, 64);
    }
    if (lane == 0) {
        colpart[t] = a ? se : 0.0;
        colpart[Tn + t] = a ? sh : 0.0;
    }
}

// out[k] = sum_t colpart[k T + t] in a fixed order (one workgroup).
__global__ __launch_bounds__(256) void adj_sum_kernel(const double* colpart, int64_t Tn, double* out) {
    __shared__ double red[4];
    for (int k = 0; k < 2; ++k) {
        double s = 0.0;
        for (int64_t t = threadIdx.x; t < Tn; t += 256) s += colpart[k * Tn + t];
        for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) out[k] = (red[0] + red[1]) + (red[2] + red[3]);
        __syncthreads();
    }
}

}  // namespace pmcmc
