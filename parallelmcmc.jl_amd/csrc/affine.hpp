// Validation helpers of the diagonal affine scan: sequential solve, residual,
// comparison of two solutions (device kernels).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmcmc {

// 0 on success.  red: three device words (maxima as double bits, a flag).
int affine_seq(hipStream_t s, bool f64, const void* A, const void* B, const void* s0, void* S, int64_t D,
               int64_t Tn);
int affine_residual(hipStream_t s, bool f64, const void* S, const void* A, const void* B, const void* s0, int64_t D,
                    int64_t Tn, unsigned long long* red);
int affine_compare(hipStream_t s, bool f64, const void* X, const void* Y, int64_t n, double atol, double rtol,
                   unsigned long long* red);

}  // namespace pmcmc
