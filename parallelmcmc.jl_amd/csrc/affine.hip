// Validation helpers of the diagonal affine scan (src/DEER/DEERScan.jl:
// solve_affine_seq! :51-68, affine_scan_residual :203-224,
// check_affine_scan :240-258): the sequential recurrence, its residual and
// the comparison of two solutions, on the device.
//
// The sequential solver and the residual round the product and the sum
// separately (no fused multiply-add), as the reference's broadcast
// A .* s .+ B does: floating-point contraction is off in this file.
#include "affine.hpp"

#include "common.hpp"

#pragma clang fp contract(off)

namespace pmcmc {

template <typename T>
__device__ __forceinline__ T mul_rn(T a, T b) {
    return a * b;
}
template <typename T>
__device__ __forceinline__ T add_rn(T a, T b) {
    return a + b;
}

// one thread per row: s_t = A[:, t] .* s_{t-1} + B[:, t]
template <typename T>
__global__ __launch_bounds__(256) void affine_seq_kernel(const T* A, const T* B, const T* s0, T* S, int64_t D,
                                                         int64_t Tn) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= D) return;
    T s = s0[d];
    for (int64_t t = 0; t < Tn; ++t) {
        const size_t i = (size_t)t * D + d;
        s = add_rn(mul_rn(A[i], s), B[i]);
        S[i] = s;
    }
}

// max over (d, t) of |S[d, t] - (A[d, t] S[d, t-1] + B[d, t])|, S[:, 0] from
// s0; out[0] holds the maximum's bits (non-negative doubles order like
// their bits; NaN above +Inf, so a NaN anywhere is reported)
template <typename T>
__global__ __launch_bounds__(256) void affine_residual_kernel(const T* S, const T* A, const T* B, const T* s0,
                                                              int64_t D, int64_t n, unsigned long long* out) {
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T prev = i < D ? s0[i] : S[i - D];
        const T r = S[i] - add_rn(mul_rn(A[i], prev), B[i]);
        m = nan_max(m, fabs((double)r));
    }
    for (int k = 32; k >= 1; k >>= 1) m = nan_max(m, __shfl_xor(m, k, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// max |X - Y| and max |X - Y| / (atol + rtol |X|) into out[0], out[1] (the
// ratio NaN where both are 0, as in the reference); out[2] = 1 where some
// |X - Y| <= atol + rtol |X| fails (NaN included)
template <typename T>
__global__ __launch_bounds__(256) void affine_compare_kernel(const T* X, const T* Y, int64_t n, double atol,
                                                             double rtol, unsigned long long* out) {
    double ma = 0.0, mr = 0.0;
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double x = (double)X[i], y = (double)Y[i];
        const double d = fabs(x - y), sc = atol + rtol * fabs(x);
        ma = nan_max(ma, d);
        mr = nan_max(mr, d / sc);
        bad = bad || !(d <= sc);
    }
    for (int k = 32; k >= 1; k >>= 1) {
        ma = nan_max(ma, __shfl_xor(ma, k, 64));
        mr = nan_max(mr, __shfl_xor(mr, k, 64));
    }
    bad = __any(bad);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(out, (unsigned long long)__double_as_longlong(ma));
        atomicMax(out + 1, (unsigned long long)__double_as_longlong(mr));
        if (bad) atomicMax(out + 2, 1ull);
    }
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

static unsigned grid_n(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (unsigned)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

int affine_seq(hipStream_t s, bool f64, const void* A, const void* B, const void* s0, void* S, int64_t D,
               int64_t Tn) {
    if (f64)
        hipLaunchKernelGGL((affine_seq_kernel<double>), dim3(grid_n(D)), dim3(256), 0, s, (const double*)A,
                           (const double*)B, (const double*)s0, (double*)S, D, Tn);
    else
        hipLaunchKernelGGL((affine_seq_kernel<float>), dim3(grid_n(D)), dim3(256), 0, s, (const float*)A,
                           (const float*)B, (const float*)s0, (float*)S, D, Tn);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int affine_residual(hipStream_t s, bool f64, const void* S, const void* A, const void* B, const void* s0, int64_t D,
                    int64_t Tn, unsigned long long* red) {
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, red, 2);
    const int64_t n = D * Tn;
    if (f64)
        hipLaunchKernelGGL((affine_residual_kernel<double>), dim3(grid_n(n)), dim3(256), 0, s, (const double*)S,
                           (const double*)A, (const double*)B, (const double*)s0, D, n, red);
    else
        hipLaunchKernelGGL((affine_residual_kernel<float>), dim3(grid_n(n)), dim3(256), 0, s, (const float*)S,
                           (const float*)A, (const float*)B, (const float*)s0, D, n, red);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int affine_compare(hipStream_t s, bool f64, const void* X, const void* Y, int64_t n, double atol, double rtol,
                   unsigned long long* red) {
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, red, 3);
    if (f64)
        hipLaunchKernelGGL((affine_compare_kernel<double>), dim3(grid_n(n)), dim3(256), 0, s, (const double*)X,
                           (const double*)Y, n, atol, rtol, red);
    else
        hipLaunchKernelGGL((affine_compare_kernel<float>), dim3(grid_n(n)), dim3(256), 0, s, (const float*)X,
                           (const float*)Y, n, atol, rtol, red);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace pmcmc
