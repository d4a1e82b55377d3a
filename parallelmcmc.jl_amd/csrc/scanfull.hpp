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

}  // namespace pmcmc
