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
    double* tile_smax = nullptr;  // nT: max |S| of each tile when it last ran
    // current tiling: nTc tiles of TCc columns from column col0 (the columns
    // before col0 are final); nTc <= nT
    int64_t col0 = 0;
    int TCc = 0, nTc = 0;
};

int full_work_ensure(FullWork* w, int64_t D, int64_t T, bool f64, std::string* msg);
void full_work_destroy(FullWork* w);
// Tiling of a new solve: TC-column tiles over all columns.
void full_geometry_reset(FullWork* w);
// Between iterations (host side, at a sync): when the final prefix reported by
// the last iteration (skip_cols = lo) has moved past col0, fold the max |S| of
// the tiles it covers into sync->smax_prefix and re-tile [lo, T) with tiles of
// 8..64 columns (at most nT tiles; the length from a latency model of the
// chains and of the scan over the tile maps), so that the per-tile chains
// shrink with the live region.
void full_retile(FullWork* w, SyncState* sync, int64_t lo, hipStream_t s);

// One Newton iteration given the products at Xbar (GlmWork: Y, gx, gy, lpx,
// lpy): accept flags and the structured Jacobians, the correction scan
// Delta_t = J_t Delta_{t-1} + (FT_t - S_t), S_out = S_in + damping Delta, and
// the convergence metric.  No-op once sync->converged is set.  Tiles below
// sync->skip_cols are final and left out (the Gaussian passes take the same
// floor); the iteration sets the next floor to the first tile whose residual
// FT - S is not exactly 0.
template <typename T>
int full_newton_step(FullWork* fw, const GlmWork* gw, const GaussData* g, hipStream_t s, SyncState* sync,
                     int precision, bool sigmoid, double eps, double damping, double tol_abs, double tol_rel, const T* x0,
                     const T* S_in, T* S_out, const T* U, uint8_t* acc_out, int iter, Prof* prof,
                     std::string* msg);

}  // namespace pmcmc
