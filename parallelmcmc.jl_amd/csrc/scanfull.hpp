// Dense-Jacobian affine scan with materialised Jacobians (D <= 64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace pmcmc {

// device scratch a scan of T columns needs (tile maps of every level)
size_t scan_full_scratch_bytes(int64_t T, bool f64);
// S[:,t] = J_t S[:,t-1] + B[:,t] from s0; J: T column-major D x D matrices
int scan_full(bool f64, const void* J, const void* B, const void* s0, void* S, int64_t D, int64_t T, void* scratch,
              hipStream_t st, std::string* msg);

// The scan over tile maps of a structured solve: out[:,k] = M_k out[:,k-1] + r[:,k]
// from a zero start, k < nT; M_k the k-th 64 x 64 column-major map, r and out
// 64 apart (rows >= D of out are left untouched).  Up to 256 maps: one fused
// launch (8-map groups, the chain over the groups in the last workgroup to
// arrive); beyond that, or without err, the recursive reduce/apply kernels
// with 8-map tiles.  Every kernel is a no-op while *skip is nonzero (skip may
// be null).  The scratch (scan_full_spine_scratch_bytes for the largest nT
// used with it) starts with two unsigned words that must be zero when it is
// first used; the fused launch keeps its counter and epoch there.  *err
// (device) is set if a wait of the fused launch hit its spin limit.
size_t scan_full_spine_scratch_bytes(int64_t nT, bool f64);
template <typename T>
void scan_full_spine(const T* M, const T* r, T* out, int64_t D, int64_t nT, const int* skip, void* scratch,
                     hipStream_t st, unsigned* err);

}  // namespace pmcmc
