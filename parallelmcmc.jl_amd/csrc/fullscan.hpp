// Full-Jacobian DEER iteration for the dense Gaussian target (D <= 64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "common.hpp"
#include "gauss.hpp"
#include "glm.hpp"

namespace pmcmc {

static constexpr int FULL_MAX_D = 64;

struct FullWork {
    int64_t D = 0, T = 0;
    size_t esz = 0;
    int TC = 0, nT = 0;
    void* a = nullptr;     // T accept flags (0/1) of the current iterate
    void* u = nullptr;     // D x T: g(1-g)(y - x)
    void* c = nullptr;     // D x T: gradient of log alpha through the surrogate
    void* rho = nullptr;   // D x T: FT - S (Newton residual)
    void* Magg = nullptr;  // nT x 64 x 64: product of the tile's Jacobians (column-major)
    void* ragg = nullptr;  // nT x 64: the tile's Delta from a zero start
    void* dend = nullptr;  // nT x 64: Delta at each tile end (rows >= D stay zero)
    void* spine = nullptr;  // scratch of the recursive scan over the tile maps
};

int full_work_ensure(FullWork* w, int64_t D, int64_t T, bool f64, std::string* msg);
void full_work_destroy(FullWork* w);

// One Newton iteration given the products at Xbar (GlmWork: Y, gx, gy, lpx,
// lpy): accept flags and the structured Jacobians, the correction scan
// Delta_t = J_t Delta_{t-1} + (FT_t - S_t), S_out = S_in + damping Delta, and
// the convergence metric.  No-op once sync->converged is set.
template <typename T>
int full_newton_step(FullWork* fw, const GlmWork* gw, const GaussData* g, hipStream_t s, SyncState* sync,
                     int precision, bool sigmoid, double eps, double damping, double tol_abs, double tol_rel, const T* x0,
                     const T* S_in, T* S_out, const T* U, uint8_t* acc_out, int iter, Prof* prof,
                     std::string* msg);

}  // namespace pmcmc
