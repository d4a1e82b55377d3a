// Owned matmul and reductions on the device (column-major operands).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace pmcmc {

// C (m x n, ldc) = op(A) op(B), op(A) m x k, op(B) k x n; k = 0 writes zeros.
int linalg_gemm(int prec, bool ta, bool tb, int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                const void* B, int64_t ldb, void* C, int64_t ldc, hipStream_t s, std::string* msg);
// *out = sum_{i<m, j<n} A[i + j lda] B[i + j ldb] (FP64 accumulation, fixed
// order); scratch holds linalg_dotsum_scratch_bytes() of device memory.
size_t linalg_dotsum_scratch_bytes();
int linalg_dotsum(int prec, int64_t m, int64_t n, const void* A, int64_t lda, const void* B, int64_t ldb, void* out,
                  double* scratch, hipStream_t s, std::string* msg);

}  // namespace pmcmc
