// Exact-diagonal (:diag) Jacobian of the taped MALA step for the GEMM
// targets (logistic regression, dense Gaussian).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "common.hpp"
#include "gauss.hpp"
#include "glm.hpp"

namespace pmcmc {

// Scratch of the :diag products: r, u and the states' gradient / log-density
// that the Hessian-vector passes also write (D x T each, T doubles).
struct DiagWork {
    int64_t D = 0, T = 0;
    size_t esz = 0;
    void* R = nullptr;
    void* U = nullptr;
    void* G = nullptr;
    double* lp = nullptr;
};

int diag_work_ensure(DiagWork* w, int64_t D, int64_t T, size_t esz, std::string* msg);
void diag_work_destroy(DiagWork* w);

// After the Newton products (gx, Y, lpx at Xbar; gy, lpy at Y in `gw`),
// overwrite gw->W (probe slot 0) with c and gw->hw (slot 0) with diag(H_x):
//   c = (I + eps H_x) u - g_x - r / (2 eps),  u = g_y + (I + eps H_y) r / (2 eps),
//   r = x - y - eps g_y,
// the derivative of log alpha of the sigmoid surrogate (the closed form of
// the reference's basis-vector JVPs, src/DEER/DEER.jl:327-342), so that
// diag J = a (1 + eps diag H_x) + (1 - a) + g (1 - g) (y - x) c.
// Two Hessian-vector passes (H_y r, H_x u) and diag H_x: -diag(Lambda) for
// the Gaussian, -sum_n s_n (1 - s_n) X_nd^2 - 1/prior^2 for the logistic.
// glm is used when kind_logistic, gauss otherwise; D <= 256.
template <typename T>
int gemm_diag_products(DiagWork* dw, GlmWork* gw, bool kind_logistic, const GlmData* glm, const GaussData* gauss,
                       double prior_scale, hipStream_t s, double eps, const T* x0, const T* S_in, int64_t D,
                       int64_t Tn, const int* skip, Prof* prof, int iter, bool jbf, std::string* msg);

}  // namespace pmcmc
