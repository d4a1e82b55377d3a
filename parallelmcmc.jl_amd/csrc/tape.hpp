// On-device MALA tape generation (opt-in; Philox4x64-10 + Box-Muller).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmcmc {

// Xi: D x T column-major, U: T, in FP64 (f64) or FP32; 0 on success.
int tape_generate(hipStream_t s, uint64_t seed, bool f64, void* Xi, void* U, int64_t D, int64_t T);

}  // namespace pmcmc
