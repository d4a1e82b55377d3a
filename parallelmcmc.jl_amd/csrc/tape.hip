// On-device MALA tape (opt-in): xi_t ~ N(0, I) and u_t ~ U(0, 1) for one
// block drawn on the GPU from a counter-based generator, instead of on the
// host and uploaded (the reference draws them from the sampler's RNG,
// src/interface.jl:334-358).
//
// Generator: Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC 2011), key (seed, stream), counter (j, 0,
// 0, 0): call j yields four 64-bit words.  Stream 0 feeds the noise, stream 1
// the uniforms.
//   u_t  = ((w >> 11) + 1/2) 2^-53 of word t & 3 of call t >> 2 (stream 1)
//   xi_i = Box-Muller of the word pair (i >> 1) & 1 of call i >> 2 (stream 0),
//          i = t D + d, the cosine branch for even i and the sine branch for
//          odd i: r = sqrt(-2 log u1), theta = 2 pi u2, u1 from the pair's
//          first word as above, u2 = (w >> 11) 2^-53 from its second.
// FP32 tapes round these FP64 values (u clamped below 1).
#include "tape.hpp"

namespace pmcmc {

__device__ __forceinline__ void philox4x64_10(uint64_t x[4], uint64_t k0, uint64_t k1) {
    constexpr uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
    constexpr uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t lo0 = M0 * x[0], hi0 = __umul64hi(M0, x[0]);
        const uint64_t lo1 = M1 * x[2], hi1 = __umul64hi(M1, x[2]);
        const uint64_t y0 = hi1 ^ x[1] ^ k0, y2 = hi0 ^ x[3] ^ k1;
        x[0] = y0;
        x[1] = lo1;
        x[2] = y2;
        x[3] = lo0;
        k0 += W0;
        k1 += W1;
    }
}

__device__ __forceinline__ double open01(uint64_t w) { return ((double)(w >> 11) + 0.5) * 0x1.0p-53; }

template <typename T>
__global__ __launch_bounds__(256) void tape_kernel(uint64_t seed, int64_t n_xi, int64_t n_u, T* Xi, T* U) {
    const int64_t calls_xi = (n_xi + 3) / 4, calls_u = (n_u + 3) / 4;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < calls_xi + calls_u;
         j += (int64_t)gridDim.x * blockDim.x) {
        const bool noise = j < calls_xi;
        const uint64_t c = (uint64_t)(noise ? j : j - calls_xi);
        uint64_t w[4] = {c, 0ull, 0ull, 0ull};
        philox4x64_10(w, seed, noise ? 0ull : 1ull);
        if (noise) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const double r = sqrt(-2.0 * log(open01(w[2 * q])));
                const double th = 6.283185307179586 * ((double)(w[2 * q + 1] >> 11) * 0x1.0p-53);
                double s, co;
                sincos(th, &s, &co);
                const int64_t i = 4 * (int64_t)c + 2 * q;
                if (i < n_xi) Xi[i] = (T)(r * co);
                if (i + 1 < n_xi) Xi[i + 1] = (T)(r * s);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t t = 4 * (int64_t)c + q;
                if (t < n_u) {
                    const double u = open01(w[q]);
                    U[t] = sizeof(T) == 4 ? (T)fmin(u, 0x1.fffffep-1) : (T)u;
                }
            }
        }
    }
}

int tape_generate(hipStream_t s, uint64_t seed, bool f64, void* Xi, void* U, int64_t D, int64_t T) {
    const int64_t n = D * T, calls = (n + 3) / 4 + (T + 3) / 4;
    int64_t g = (calls + 255) / 256;
    if (g > 8192) g = 8192;
    if (f64)
        hipLaunchKernelGGL((tape_kernel<double>), dim3((unsigned)g), dim3(256), 0, s, seed, n, T, (double*)Xi,
                           (double*)U);
    else
        hipLaunchKernelGGL((tape_kernel<float>), dim3((unsigned)g), dim3(256), 0, s, seed, n, T, (float*)Xi,
                           (float*)U);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace pmcmc
