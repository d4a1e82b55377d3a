// The owned matmul and reductions (src/ParallelMCMC.jl:23-25):
//   pmcmc_matmul(A, B) = A * B, pmcmc_dot(a, b) = dot(a, b),
//   pmcmc_dotsum(A, B) = sum(A .* B)
// as standalone device entry points on column-major operands.  The DEER hot
// path does not call them (its GEMMs are fused into the target passes); they
// serve callers that used the reference's wrappers directly, including the
// pullbacks of its AD rules (ext/EnzymeExt.jl), which are products with one
// operand transposed: dA = dY B^T, dB = A^T dY.
//
// GEMM: 64 x 64 output tile per workgroup, K staged through LDS 16 at a time,
// four waves of 2 x 2 MFMA 16x16x4 tiles (f32, or f64 for FP64 operands).
// Either operand may be transposed; staging maps consecutive threads to the
// operand's contiguous index, so both orientations load coalesced.
// Reductions: per-workgroup FP64 partials in a fixed order, then one
// workgroup sums them in a fixed order (deterministic for a given size).
#include <hip/hip_runtime.h>

#include "../../include/pmcmc_amd.h"
#include "linalg.hpp"
#include "mfma.hpp"

namespace pmcmc {

static constexpr int GM = 64, GN = 64, GK = 16, GPAD = 1;

// C[i + j ldc] = sum_p opA(i, p) opB(p, j); opA(i, p) = TA ? A[p + i lda] : A[i + p lda],
// opB(p, j) = TB ? B[j + p ldb] : B[p + j ldb]
template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_kernel(int64_t m, int64_t n, int64_t k, const T* __restrict__ A,
                                                   int64_t lda, const T* __restrict__ B, int64_t ldb,
                                                   T* __restrict__ C, int64_t ldc) {
    __shared__ T As[GK][GM + GPAD];  // As[p][i] = opA(i0 + i, p0 + p)
    __shared__ T Bs[GK][GN + GPAD];  // Bs[p][j] = opB(p0 + p, j0 + j)
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int wm = wv & 1, wn = wv >> 1;
    const int64_t i0 = (int64_t)blockIdx.x * GM, j0 = (int64_t)blockIdx.y * GN;
    typename Mfma<T>::acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (typename Mfma<T>::acc_t){0, 0, 0, 0};
    for (int64_t p0 = 0; p0 < k; p0 += GK) {
        for (int e = tid; e < GK * GM; e += 256) {
            const int ii = TA ? e / GK : e % GM;
            const int pp = TA ? e % GK : e / GM;
            const int64_t gi = i0 + ii, gp = p0 + pp;
            T v = 0;
            if (gi < m && gp < k) v = TA ? A[gp + gi * lda] : A[gi + gp * lda];
            As[pp][ii] = v;
        }
        for (int e = tid; e < GK * GN; e += 256) {
            const int jj = TB ? e % GN : e / GK;
            const int pp = TB ? e / GN : e % GK;
            const int64_t gj = j0 + jj, gp = p0 + pp;
            T v = 0;
            if (gj < n && gp < k) v = TB ? B[gj + gp * ldb] : B[gp + gj * ldb];
            Bs[pp][jj] = v;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GK / 4; ++ks) {
            T av[2], bv[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                av[t] = As[4 * ks + lk][wm * 32 + t * 16 + li];
                bv[t] = Bs[4 * ks + lk][wn * 32 + t * 16 + li];
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = Mfma<T>::step(av[a], bv[b], acc[a][b]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int64_t col = j0 + wn * 32 + b * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = i0 + wm * 32 + a * 16 + Mfma<T>::row(lk, r);
                if (row < m && col < n) C[row + col * ldc] = acc[a][b][r];
            }
        }
}

static constexpr int RED_BLOCKS = 1024;

__device__ inline double block_sum256(double v, double* red) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// part[block] = sum over this block's elements of A[i + j lda] * B[i + j ldb],
// elements e = i + j m taken grid-stride
template <typename T>
__global__ __launch_bounds__(256) void dotsum_partial_kernel(int64_t m, int64_t n, const T* A, int64_t lda,
                                                             const T* B, int64_t ldb, double* part) {
    __shared__ double red[4];
    const int64_t total = m * n;
    double acc = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t i = e % m, j = e / m;
        acc += (double)A[i + j * lda] * (double)B[i + j * ldb];
    }
    const double s = block_sum256(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <typename T>
__global__ __launch_bounds__(256) void dotsum_final_kernel(const double* part, int nb, T* out) {
    __shared__ double red[4];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) acc += part[b];
    const double s = block_sum256(acc, red);
    if (threadIdx.x == 0) *out = (T)s;
}

template <typename T>
static void launch_gemm(bool ta, bool tb, int64_t m, int64_t n, int64_t k, const T* A, int64_t lda, const T* B,
                        int64_t ldb, T* C, int64_t ldc, hipStream_t s) {
    const dim3 grid((unsigned)((m + GM - 1) / GM), (unsigned)((n + GN - 1) / GN));
    if (!ta && !tb)
        hipLaunchKernelGGL((gemm_kernel<T, false, false>), grid, dim3(256), 0, s, m, n, k, A, lda, B, ldb, C, ldc);
    else if (ta && !tb)
        hipLaunchKernelGGL((gemm_kernel<T, true, false>), grid, dim3(256), 0, s, m, n, k, A, lda, B, ldb, C, ldc);
    else if (!ta && tb)
        hipLaunchKernelGGL((gemm_kernel<T, false, true>), grid, dim3(256), 0, s, m, n, k, A, lda, B, ldb, C, ldc);
    else
        hipLaunchKernelGGL((gemm_kernel<T, true, true>), grid, dim3(256), 0, s, m, n, k, A, lda, B, ldb, C, ldc);
}

int linalg_gemm(int prec, bool ta, bool tb, int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                const void* B, int64_t ldb, void* C, int64_t ldc, hipStream_t s, std::string* msg) {
    if (m == 0 || n == 0) return PMCMC_OK;
    if ((m + GM - 1) / GM > 0x7fffffff || (n + GN - 1) / GN > 65535) {
        *msg = "matmul: output too large for one launch";
        return PMCMC_ERR_UNSUPPORTED;
    }
    if (prec == PMCMC_PREC_F64)
        launch_gemm<double>(ta, tb, m, n, k, (const double*)A, lda, (const double*)B, ldb, (double*)C, ldc, s);
    else
        launch_gemm<float>(ta, tb, m, n, k, (const float*)A, lda, (const float*)B, ldb, (float*)C, ldc, s);
    if (hipGetLastError() != hipSuccess) {
        *msg = "matmul: kernel launch failed";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}

size_t linalg_dotsum_scratch_bytes() { return RED_BLOCKS * sizeof(double); }

int linalg_dotsum(int prec, int64_t m, int64_t n, const void* A, int64_t lda, const void* B, int64_t ldb, void* out,
                  double* scratch, hipStream_t s, std::string* msg) {
    const int64_t total = m * n;
    int nb = (int)((total + 255) / 256);
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;
    if (nb < 1) nb = 1;  // an empty sum still writes 0
    if (prec == PMCMC_PREC_F64) {
        hipLaunchKernelGGL(dotsum_partial_kernel<double>, dim3(nb), dim3(256), 0, s, m, n, (const double*)A, lda,
                           (const double*)B, ldb, scratch);
        hipLaunchKernelGGL(dotsum_final_kernel<double>, dim3(1), dim3(256), 0, s, scratch, nb, (double*)out);
    } else {
        hipLaunchKernelGGL(dotsum_partial_kernel<float>, dim3(nb), dim3(256), 0, s, m, n, (const float*)A, lda,
                           (const float*)B, ldb, scratch);
        hipLaunchKernelGGL(dotsum_final_kernel<float>, dim3(1), dim3(256), 0, s, scratch, nb, (float*)out);
    }
    if (hipGetLastError() != hipSuccess) {
        *msg = "dotsum: kernel launch failed";
        return PMCMC_ERR_HIP;
    }
    return PMCMC_OK;
}

}  // namespace pmcmc
