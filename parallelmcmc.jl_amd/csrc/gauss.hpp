// Dense Gaussian target N(mean, Lambda^-1): grad = -Lambda (x - mean),
// Hv = -Lambda v, logp = -(x - mean)' Lambda (x - mean) / 2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "common.hpp"
#include "glm.hpp"

namespace pmcmc {

static constexpr int GAUSS_CB = 32;  // trajectory columns per workgroup of the Lambda pass (one probe set)

struct GaussData {
    int64_t D = 0, Dp = 0;
    float* L32 = nullptr;    // Dp x Dp row-major, zero padded
    double* L64 = nullptr;   // D x D column-major
    float* mu32 = nullptr;   // null: zero mean
    double* mu64 = nullptr;
};

// One chain's operands of a batched Lambda pass (one probe set).
struct GaussPassArgs {
    ColSrc<float> src;
    const float* L32;
    const float* mu;
    int pass;
    float eps;
    const float* Xi;
    float* gout;
    float* hout;
    double* lpout;
    float* Yout;
    float* Wout;
    const int* skip;
    int ldL;  // Lambda's padded width (row stride of L32), <= 16 * dpt of the launch
    // slab passes: the log-density partial of slab s goes to lpout[s * lp_stride + t]
    int64_t lp_stride;
    // non-null: every workgroup adds its algorithmic flops (2 x rows x D x
    // columns, states and directions) here
    unsigned long long* flops;
};

// Batched FP32 Lambda passes over 64-row slabs (one probe set), any widths in
// one launch: items (chain, column block of GAUSS_CB, slab).
void launch_gauss_slab_batched(bool jbf, const GaussPassArgs* args, const int4* items, int nitems, hipStream_t s);

// Batched FP32 Lambda passes (one probe set) over chains of padded width up to 16 * dpt.
void launch_gauss_pass_batched(int dpt, bool jbf, const GaussPassArgs* args, const int2* items, int nitems,
                               hipStream_t s);

int gauss_data_create(GaussData* g, hipStream_t s, int64_t D, bool src_f64, const void* Lam,
                      const void* mean, std::string* msg);
void gauss_data_destroy(GaussData* g);

// Same products and buffers as glm_newton_products (GlmWork): gx, W, Y, lpx
// at Xbar; gy, hw, lpy at Y.  col_lo (device, may be null): the FP32 MFMA
// passes leave out column blocks wholly below *col_lo (their outputs keep
// older values).
template <typename T>
int gauss_newton_products(GlmWork* w, const GaussData* g, hipStream_t s, SyncState* sync,
                          int precision, double eps, const T* x0, const T* S_in, const T* Xi,
                          uint64_t seed, int iter, uint32_t probe_iter, int probes, int64_t D, int64_t Tn,
                          Prof* prof, std::string* msg, const int* col_lo = nullptr);

// As glm_dir_products: w->gx = grad(states), w->W = dir + eps H dir.
template <typename T>
int gauss_dir_products(GlmWork* w, const GaussData* g, hipStream_t s, double eps, const T* V, const T* x0,
                       const T* Z, uint32_t key, const T* Xi, int64_t D, int64_t Tn, const int* skip, Prof* prof,
                       int iter, std::string* msg);

// As glm_hvp_products: hout = H(V) Z = -Lambda Z.
template <typename T>
int gauss_hvp_products(const GaussData* g, hipStream_t s, const T* V, const T* x0, const T* Z, T* gscr,
                       double* lpscr, T* hout, int64_t D, int64_t Tn, const int* skip, Prof* prof, int iter, bool jbf,
                       std::string* msg);

// K sequential MALA steps of N chains (as glm_mala_run).
int gauss_mala_run(const GaussData* g, hipStream_t s, int prec, double eps, const void* X0, const void* Xi,
                   const void* U, int64_t K, int64_t D, int64_t N, void* S_out, uint8_t* acc, const void* Z, void* Jw,
                   std::string* msg);

int gauss_fwd_and_jvp(const GaussData* g, hipStream_t s, int prec, double eps, const void* X,
                      const void* Xi, const void* U, const void* Z, void* FT, void* Jt,
                      uint8_t* accepted, double* logalpha, int64_t D, int64_t T, std::string* msg);

int gauss_logdensity(const GaussData* g, hipStream_t s, int prec, const void* S, double* out,
                     int64_t T, std::string* msg);

}  // namespace pmcmc
