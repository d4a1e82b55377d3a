// FP64 recompute of accept decisions near the threshold (GEMM targets, FP32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace pmcmc {

// FP64 view of a GEMM target: kind 0 logistic (X64 N x D column-major, y64,
// 1/prior_scale^2), kind 1 dense Gaussian (L64 D x D column-major, mu64 or null).
struct RefineTarget {
    int kind;
    int64_t D, N;
    const double* X64;
    const double* y64;
    double inv_s2;
    const double* P64;  // dense prior precision (D x D) or null
    const double* L64;
    const double* mu64;
};

// One Newton iteration's columns: the iterate (Xbar = [x0, S_in]), tape,
// FP32 products and log-densities; dec[t] receives -1 (keep the FP32
// decision) or, where |log alpha - log u| < band, the FP64 decision in bit 0
// and bit 1 set when the FP64 margin is below hyst.
struct RefineArgs {
    const float* x0;
    const float* S_in;
    const float* Xi;
    const float* U;
    const float* Y;
    const float* gx;
    const float* gy;
    const double* lpx;
    const double* lpy;
    // log-densities in lp_nslab partials (lp[s * lp_stride + t], summed in
    // slab order); 0 or 1: one value per column
    int lp_nslab;
    int64_t lp_stride;
    double eps;
    double band;
    const int* skip;
    // gate: the FP64 work starts at the first iteration whose predecessor's
    // update metric (in tolerance units, the solve's SyncState) is <= gate,
    // or at iteration gate_iters (FP32 decisions near the threshold can keep
    // the metric from ever falling), and stays on for the rest of the solve
    const double* last_metric;
    double gate;
    int iter;
    int gate_iters;
    // batched solves keep the iteration on the device: iter = *iter_dev + 1
    // (the SyncState's last finished iteration) when set
    const int* iter_dev;
    // FP64 margins below this set the tie bit (bit 1) of dec[t]
    double hyst;
    // columns below *final_cols are final (the solve's final prefix, which
    // only moves once the gate is on): their decisions stay as they are and
    // the flag pass leaves them out (null: none)
    const int* final_cols;
};

// Device buffers of the refinement.  Flagged columns are compacted into
// list[0 .. min(ctl[par], lmax)); the counter is double-buffered by iteration
// parity (ctl[0], ctl[1]) so the flag pass of one iteration clears the one the
// next iteration appends to.  ctl[2] is the sticky gate flag.
// Each flagged column's FP64 gradient is split over nsplit workgroups (row
// ranges of X for the logistic target, column ranges of Lambda for the
// Gaussian) whose partials part[j][split][0..D] (D gradient entries and one
// log-density term) are summed in a fixed order.
struct RefineWork {
    int64_t D = 0, T = 0, N = 0;
    int kind = 0;
    int lmax = 0, nsplit = 0;
    int64_t rows = 0;       // rows of X (logistic) or columns of Lambda (Gaussian) per split
    int8_t* dec = nullptr;  // T: -1 keep, else bit 0 accept, bit 1 tie
    int* cnt = nullptr;     // ctl: 2 counters, gate flag, spare
    int* list = nullptr;    // lmax flagged column indices
    double* part = nullptr;  // lmax x nsplit x (D + 1)
    double* g0 = nullptr;    // lmax x D: FP64 gradient at the state
    double* y = nullptr;     // lmax x D: FP64 proposal
    double* lp0 = nullptr;   // lmax: FP64 log-density at the state
};

// Size the buffers for a target, trajectory length and band (reuses buffers
// of the same shape).  lmax is at least 16x the expected number of flagged
// columns (at most 2 band T for any iterate, since u is uniform: the density
// of log u is u <= 1) and at least what 32 MiB of partials hold, capped at T.
// Flagged columns beyond lmax keep their FP32 decision.
int refine_work_ensure(RefineWork* w, const RefineTarget& tg, int64_t T, double band, std::string* msg);
void refine_work_destroy(RefineWork* w);
// clear the counters and the gate flag (once per solve, before the first iteration)
int refine_reset(RefineWork* w, hipStream_t s, std::string* msg);
// one iteration's refinement: flag pass, then FP64 gradient and
// log-density at the state and at the proposal of every flagged column
int refine_decisions(RefineWork* w, const RefineTarget& tg, const RefineArgs& a, int iter, hipStream_t s,
                     std::string* msg);

// Batched form: per-chain descriptor tables (device) indexed by the live-chain
// lists (device); `par` is the iteration parity of the launch.  glm / gauss
// list the live chains of each target kind, all lists every live chain.
struct RefineBatch {
    const RefineTarget* tgs;
    const RefineArgs* ras;  // for this parity
    const RefineWork* rws;
    const int* all;
    int n_all;
    const int* glm;
    int n_glm;
    const int* gauss;
    int n_gauss;
    int64_t max_T;       // longest live window
    int max_split_glm;   // most row splits of a live logistic chain
    int max_split_gauss;
};
void refine_decisions_batched(const RefineBatch& b, int par, hipStream_t s);
// device bytes of one chain's refinement buffers (its lmax, nsplit, D, window)
size_t refine_chain_bytes(const RefineTarget& tg, int64_t L, double band);
// carve a chain's buffers from `base` (refine_chain_bytes of them)
void refine_chain_carve(RefineWork* w, const RefineTarget& tg, int64_t L, double band, char* base);

}  // namespace pmcmc
