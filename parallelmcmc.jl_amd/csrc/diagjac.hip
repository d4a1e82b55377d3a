// Exact-diagonal (:diag) Jacobian for the GEMM targets: the surrogate's
// log-alpha gradient c from two Hessian-vector GEMM passes, and diag(H_x).
#include "diagjac.hpp"

#include "../../include/pmcmc_amd.h"

namespace pmcmc {

int diag_work_ensure(DiagWork* w, int64_t D, int64_t T, size_t esz, std::string* msg) {
    if (w->R && w->D == D && w->T >= T && w->esz == esz) return PMCMC_OK;
    diag_work_destroy(w);
    const size_t m = (size_t)D * T * esz;
    if (hipMalloc(&w->R, m) != hipSuccess || hipMalloc(&w->U, m) != hipSuccess ||
        hipMalloc(&w->G, m) != hipSuccess || hipMalloc(&w->lp, (size_t)T * sizeof(double)) != hipSuccess) {
        diag_work_destroy(w);
        *msg = ":diag workspace allocation failed";
        return PMCMC_ERR_ALLOC;
    }
    w->D = D;
    w->T = T;
    w->esz = esz;
    return PMCMC_OK;
}

void diag_work_destroy(DiagWork* w) {
    (void)hipFree(w->R);
    (void)hipFree(w->U);
    (void)hipFree(w->G);
    (void)hipFree(w->lp);
    *w = DiagWork{};
}

// r = x - y - eps g_y, x the iterate shifted by one step (x0 first)
template <typename T>
__global__ __launch_bounds__(256) void diag_r_kernel(const T* x0, const T* S_in, const T* Y, const T* gy, double eps,
                                                     int64_t D, int64_t n, T* R, const int* skip) {
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double x = i < D ? (double)x0[i] : (double)S_in[i - D];
        R[i] = (T)(x - (double)Y[i] - eps * (double)gy[i]);
    }
}

// u = g_y + (r + eps H_y r) / (2 eps)
template <typename T>
__global__ __launch_bounds__(256) void diag_u_kernel(const T* gy, const T* R, const T* Hr, double eps, int64_t n,
                                                     T* U, const int* skip) {
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double r = (double)R[i];
        U[i] = (T)((double)gy[i] + (r + eps * (double)Hr[i]) / (2.0 * eps));
    }
}

// c = (u + eps H_x u) - g_x - r / (2 eps) into W; with L64 (Gaussian) also
// diag(H_x) = -Lambda_dd into Hd (which holds H_x u until this thread reads it)
template <typename T>
__global__ __launch_bounds__(256) void diag_c_kernel(const T* U, const T* gx, const T* R, double eps, int64_t D,
                                                     int64_t n, const double* L64, T* W, T* Hd, const int* skip) {
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double u = (double)U[i];
        const double hu = (double)Hd[i];
        W[i] = (T)((u + eps * hu) - (double)gx[i] - (double)R[i] / (2.0 * eps));
        if (L64) {
            const int64_t d = i % D;
            Hd[i] = (T)(-L64[d * D + d]);
        }
    }
}

// Logistic diag(H_x)[d, t] = -sum_n s_n (1 - s_n) X_nd^2 - 1/prior^2 (or
// - P_dd with a dense prior precision P),
// s_n = sigmoid(X_n . x_t), x_t the shifted iterate.  A workgroup owns 8
// columns; the design streams through LDS in chunks of 32 rows: first the
// 32 x 8 linear predictors (one per thread, FP64 sums), then each thread
// (row d of the design) accumulates its 8 columns.  D <= 256.
template <typename T>
__global__ __launch_bounds__(256) void glm_hess_diag_kernel(const float* Xrm, const double* Xd64, int64_t N,
                                                            int64_t D, int64_t Dp, const T* x0, const T* S_in,
                                                            int64_t Tn, double inv_s2, const double* P64, T* Hd,
                                                            const int* skip) {
    constexpr int CT = 8, NR = 32, XS = 257;
    __shared__ T xs[CT][256];
    __shared__ T xc[NR][XS];
    __shared__ double wv[NR][CT];
    if (skip && *skip) return;
    const int tid = threadIdx.x;
    const int64_t t0 = (int64_t)blockIdx.x * CT;
    for (int i = tid; i < CT * 256; i += 256) {
        const int c = i >> 8, d = i & 255;
        const int64_t t = t0 + c;
        T v = (T)0;
        if (d < D && t < Tn) v = t == 0 ? x0[d] : S_in[(t - 1) * D + d];
        xs[c][d] = v;
    }
    double acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = 0.0;
    for (int64_t n0 = 0; n0 < N; n0 += NR) {
        __syncthreads();  // the previous chunk's reads are done
        if (sizeof(T) == 4) {
            // row-major float design: rows contiguous over d
            for (int i = tid; i < NR * 256; i += 256) {
                const int r = i >> 8, d = i & 255;
                const int64_t nn = n0 + r;
                xc[r][d] = (nn < N && d < D) ? (T)Xrm[nn * Dp + d] : (T)0;
            }
        } else {
            // column-major double design: contiguous over n
            for (int i = tid; i < NR * 256; i += 256) {
                const int r = i & (NR - 1), d = i >> 5;
                const int64_t nn = n0 + r;
                xc[r][d] = (nn < N && d < D) ? (T)Xd64[nn + d * N] : (T)0;
            }
        }
        __syncthreads();
        {
            const int r = tid & (NR - 1), c = tid >> 5;
            double eta = 0.0;
            for (int d = 0; d < D; ++d) eta += (double)xc[r][d] * (double)xs[c][d];
            const double sg = 1.0 / (1.0 + exp(-eta));
            wv[r][c] = (n0 + r < N) ? sg * (1.0 - sg) : 0.0;
        }
        __syncthreads();
        if (tid < D) {
#pragma unroll 4
            for (int r = 0; r < NR; ++r) {
                const double x = (double)xc[r][tid];
                const double x2 = x * x;
#pragma unroll
                for (int c = 0; c < CT; ++c) acc[c] += wv[r][c] * x2;
            }
        }
    }
    if (tid < D)
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const int64_t t = t0 + c;
            if (t < Tn) Hd[t * D + tid] = (T)(-acc[c] - (P64 ? P64[(size_t)tid * D + tid] : inv_s2));
        }
}

static unsigned grid_of(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (unsigned)(g < 4096 ? (g < 1 ? 1 : g) : 4096);
}

template <typename T>
int gemm_diag_products(DiagWork* dw, GlmWork* gw, bool kind_logistic, const GlmData* glm, const GaussData* gauss,
                       double prior_scale, hipStream_t s, double eps, const T* x0, const T* S_in, int64_t D,
                       int64_t Tn, const int* skip, Prof* prof, int iter, bool jbf, std::string* msg) {
    if (D > 256) {
        *msg = "jacobian=:diag on a GEMM target needs D <= 256";
        return PMCMC_ERR_UNSUPPORTED;
    }
    int rc = diag_work_ensure(dw, D, Tn, sizeof(T), msg);
    if (rc != PMCMC_OK) return rc;
    const int64_t n = D * Tn;
    T* R = (T*)dw->R;
    T* U = (T*)dw->U;
    T* G = (T*)dw->G;
    T* W = (T*)gw->W;
    T* H = (T*)gw->hw;
    hipLaunchKernelGGL((diag_r_kernel<T>), dim3(grid_of(n)), dim3(256), 0, s, x0, S_in, (const T*)gw->Y,
                       (const T*)gw->gy, eps, D, n, R, skip);
    // H_y r
    rc = kind_logistic
             ? glm_hvp_products<T>(gw, glm, s, prior_scale, (const T*)gw->Y, nullptr, R, G, dw->lp, H, D, Tn, skip,
                                   prof, iter, jbf, msg)
             : gauss_hvp_products<T>(gauss, s, (const T*)gw->Y, nullptr, R, G, dw->lp, H, D, Tn, skip, prof, iter,
                                     jbf, msg);
    if (rc != PMCMC_OK) return rc;
    hipLaunchKernelGGL((diag_u_kernel<T>), dim3(grid_of(n)), dim3(256), 0, s, (const T*)gw->gy, R, H, eps, n, U,
                       skip);
    // H_x u at the shifted iterate
    rc = kind_logistic ? glm_hvp_products<T>(gw, glm, s, prior_scale, S_in, x0, U, G, dw->lp, H, D, Tn, skip, prof,
                                             iter, jbf, msg)
                       : gauss_hvp_products<T>(gauss, s, S_in, x0, U, G, dw->lp, H, D, Tn, skip, prof, iter, jbf, msg);
    if (rc != PMCMC_OK) return rc;
    hipLaunchKernelGGL((diag_c_kernel<T>), dim3(grid_of(n)), dim3(256), 0, s, U, (const T*)gw->gx, R, eps, D, n,
                       kind_logistic ? nullptr : gauss->L64, W, H, skip);
    if (kind_logistic)
        hipLaunchKernelGGL((glm_hess_diag_kernel<T>), dim3((unsigned)((Tn + 7) / 8)), dim3(256), 0, s, glm->Xrm,
                           glm->Xd64, glm->N, D, glm->Dp, x0, S_in, Tn, 1.0 / (prior_scale * prior_scale), glm->P64, H,
                           skip);
    if (hipGetLastError() != hipSuccess) {
        *msg = ":diag products: kernel launch failed";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}
template int gemm_diag_products<float>(DiagWork*, GlmWork*, bool, const GlmData*, const GaussData*, double,
                                       hipStream_t, double, const float*, const float*, int64_t, int64_t, const int*,
                                       Prof*, int, bool, std::string*);
template int gemm_diag_products<double>(DiagWork*, GlmWork*, bool, const GlmData*, const GaussData*, double,
                                        hipStream_t, double, const double*, const double*, int64_t, int64_t,
                                        const int*, Prof*, int, bool, std::string*);

}  // namespace pmcmc
