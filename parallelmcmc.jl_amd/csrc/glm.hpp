// Logistic-regression (GLM) target: data layout and the Newton-iteration
// products computed as GEMMs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "common.hpp"

namespace pmcmc {

static constexpr int GLM_CB = 32;  // trajectory columns per workgroup of the FP32 GEMM pass (one probe set)

struct GlmData {
    int64_t N = 0, D = 0;
    int64_t Dp = 0;         // D padded to the MFMA K granularity
    float* Xrm = nullptr;   // N x Dp row-major (row n contiguous), zero padded
    float* Xcm = nullptr;   // Dp x N: column d of the design contiguous over n
    double* Xd64 = nullptr; // N x D column-major copy for the FP64 path
    // BF16 copies of the design for the directions' GEMMs of the BF16-Jacobian
    // pass (the FP32 values rounded to nearest even, zero padded):
    // Xb row-major Np x Dq (Dq = D rounded up to 32, one 16x16x32 MFMA K step
    // per 32 columns), Xtb transposed Dp x Np (row d contiguous over n)
    uint16_t* Xb = nullptr;
    uint16_t* Xtb = nullptr;
    int64_t Np = 0, Dq = 0;
    float* y = nullptr;
    double* y64 = nullptr;
    // Gaussian prior N(0, P^-1) with a dense precision P (D x D, symmetric,
    // FP64), or null: the isotropic prior N(0, prior_scale^2 I)
    double* P64 = nullptr;
};

// Column sources of a GEMM pass.
// Set 0 holds the states: column t is V[:, t], or with `shift` the iterate
// shifted by one step (x0 for t = 0, V[:, t-1] otherwise) = Xbar.
// Sets 1..K hold directions: Zsrc[k] if given, else the Rademacher probe k.
template <typename T>
struct ColSrc {
    const T* V;
    const T* x0;  // non-null: shifted
    const T* Z;   // K x D x T or null
    uint32_t keys[4];
    int K;
    int64_t D, Tn;
    const int* col_lo;  // device; column blocks wholly below *col_lo are left out (null: none)
    // element offset of column 0 in the probe stream: a window of a longer
    // trajectory keeps the probes of its absolute columns
    int64_t zoff;

    __device__ inline double get(int set, int64_t t, int64_t d) const {
        if (set == 0) {
            if (x0) return t == 0 ? (double)x0[d] : (double)V[(size_t)(t - 1) * D + d];
            return (double)V[(size_t)t * D + d];
        }
        const size_t gi = (size_t)t * D + d;
        if (Z) return (double)Z[(size_t)(set - 1) * D * Tn + gi];
        return (double)rademacher(keys[set - 1], gi + (size_t)zoff);
    }
};

struct GlmWork {
    int64_t D = 0, T = 0;
    int probes = 0;
    size_t esz = 0;
    void* Y = nullptr;   // D x T
    void* gx = nullptr;  // D x T
    void* gy = nullptr;  // D x T
    void* W = nullptr;   // probes x D x T
    void* hw = nullptr;  // probes x D x T
    double* lpx = nullptr;
    double* lpy = nullptr;
    void* part = nullptr;  // split-N partial sums
    size_t part_bytes = 0;
};

// One chain's operands of a batched FP32 GEMM pass (glm_pass_mfma_batched).
struct GlmPassArgs {
    ColSrc<float> src;
    const float* Xrm;
    const uint16_t* Xb;   // BF16 copies (GlmData)
    const uint16_t* Xtb;
    int64_t Np;
    const float* y32;
    int64_t N;
    float* part;  // split partial sums (the split count comes with each work item)
    double* llpart;
    const int* skip;
    // non-null: every workgroup adds its algorithmic flops (8 x rows x D x
    // columns: two GEMMs of the states and of the directions)
    unsigned long long* flops;
};

// One chain's operands of a batched split-sum / epilogue (glm_reduce_batched).
struct GlmReduceArgs {
    ColSrc<float> src;
    const float* part;
    const double* llpart;
    int64_t Dp, llstride;
    int pass;
    double inv_s2, eps;
    const double* P;  // dense prior precision or null (isotropic inv_s2)
    const float* Xi;
    float* gout;
    float* hout;
    double* lpout;
    float* Yout;
    float* Wout;
    const int* skip;
};

// N-split plan of the FP32 GEMM pass over Tn columns (ns splits of
// rows_per_split rows, enough for about target_wgs workgroups); the pass's
// partial sums take ns * 2 * Tn * Dp floats followed by ns * Tn doubles of
// log-likelihood partials.
void glm_split_plan(const GlmData* g, int64_t Tn, int* ns, int64_t* rows_per_split, int target_wgs = 768);
// Batched FP32 passes (one probe set) over chains of padded width 16 * dpt.
void launch_glm_pass_batched(int dpt, bool jbf, const GlmPassArgs* args, const int4* items, int nitems,
                             hipStream_t s);
void launch_glm_reduce_batched(const GlmReduceArgs* args, const int4* items, int nitems, hipStream_t s);

int glm_data_create(GlmData* g, hipStream_t s, int64_t N, int64_t D, bool src_f64, const void* Xd,
                    const void* y, std::string* msg);
void glm_data_destroy(GlmData* g);
// Dense prior precision P (D x D, symmetric) in place of the isotropic prior.
int glm_data_set_prior(GlmData* g, hipStream_t s, bool src_f64, const void* P, std::string* msg);
// K sequential MALA steps of N chains (src/MALA/MALA.jl:487-528 per step):
// Xi, S_out D x (N K), U and acc N K, step k in column block k; Z, Jw D x N
// scratch with Z zero.  Synchronises the stream.
int glm_mala_run(const GlmData* g, hipStream_t s, int prec, double eps, double prior_scale, const void* X0,
                 const void* Xi, const void* U, int64_t K, int64_t D, int64_t N, void* S_out, uint8_t* acc,
                 const void* Z, void* Jw, std::string* msg);
int glm_work_ensure(GlmWork* w, const GlmData* g, hipStream_t s, int64_t D, int64_t T, int probes,
                    bool f64, std::string* msg);
void glm_work_destroy(GlmWork* w);

// Products of one Newton iteration at Xbar = [x0, S_in[:, 0:T-1]]:
//   gx = grad(Xbar), W_k = Z_k + eps H(Xbar) Z_k, Y = Xbar + eps gx + sqrt(2 eps) Xi,
//   gy = grad(Y), hw_k = H(Y) W_k, lpx = logp(Xbar), lpy = logp(Y).
// Early-exits (no-op) once sync->converged is set.
template <typename T>
int glm_newton_products(GlmWork* w, const GlmData* g, hipStream_t s, SyncState* sync, int precision,
                        double eps, double prior_scale, const T* x0, const T* S_in, const T* Xi,
                        uint64_t seed, int iter, uint32_t probe_iter, int probes, int64_t D, int64_t Tn,
                        Prof* prof, std::string* msg);

// logalpha, accept, FT and Jt of the batched MALA step + JVP from the products
// of the two passes (one wave per column).
template <typename T>
void launch_fwd_jvp_finish(hipStream_t s, const T* X, const T* Z, const T* U, const GlmWork* w,
                           double eps, int64_t D, int64_t Tn, T* FT, T* Jt, uint8_t* acc,
                           double* la);

// One pass at the states V (shifted by x0 when given) with one direction set
// (Z, or the Rademacher probe with `key` when Z is null): w->gx = grad(states),
// w->W = dir + eps H(states) dir.  w must hold one probe set.
template <typename T>
int glm_dir_products(GlmWork* w, const GlmData* g, hipStream_t s, double eps, double prior_scale, const T* V,
                     const T* x0, const T* Z, uint32_t key, const T* Xi, int64_t D, int64_t Tn, const int* skip,
                     Prof* prof, int iter, std::string* msg);

// One pass at the states V (shifted by x0 when given) along the directions Z:
// hout = H(V) Z (Hessian of the log-density); the states' gradient and
// log-density are written to gscr / lpscr.
template <typename T>
int glm_hvp_products(GlmWork* w, const GlmData* g, hipStream_t s, double prior_scale, const T* V, const T* x0,
                     const T* Z, T* gscr, double* lpscr, T* hout, int64_t D, int64_t Tn, const int* skip, Prof* prof,
                     int iter, bool jbf, std::string* msg);

int glm_fwd_and_jvp(const GlmData* g, hipStream_t s, int prec, double eps, double prior_scale,
                    const void* X, const void* Xi, const void* U, const void* Z, void* FT, void* Jt,
                    uint8_t* accepted, double* logalpha, int64_t D, int64_t T, std::string* msg);

int glm_logdensity(const GlmData* g, hipStream_t s, int prec, double prior_scale, const void* S,
                   double* out, int64_t T, std::string* msg);

}  // namespace pmcmc
