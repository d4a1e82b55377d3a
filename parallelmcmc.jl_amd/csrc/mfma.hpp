// The 16x16x4 MFMA forms (FP32 and FP64) behind one interface.
#pragma once

#include <hip/hip_runtime.h>

namespace pmcmc {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// Both 16x16x4 forms take A[i][k] from lane i + 16 k and B[k][j] from lane
// j + 16 k.  They differ in the output fragment: accumulator r of lane
// li + 16 lk holds row 4 lk + r (f32) or row lk + 4 r (f64) of column li.
template <>
struct Mfma<float> {
    typedef v4f acc_t;
    static __device__ inline acc_t step(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lk, int r) { return 4 * lk + r; }
};
template <>
struct Mfma<double> {
    typedef v4d acc_t;
    static __device__ inline acc_t step(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lk, int r) { return lk + 4 * r; }
};

}  // namespace pmcmc
