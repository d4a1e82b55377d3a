// C ABI of the MI355X DEER-over-MALA library (declared in include/pmcmc_amd.h).
//
// Host orchestration of the device-resident Newton loop: the iterate, the
// tape and every temporary stay in HBM; the host enqueues iterations in
// growing chunks and reads one convergence word per chunk (the reference
// syncs two maxima to the host every iteration, src/DEER/DEER.jl:567-571).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>

#include "../../include/pmcmc_amd.h"
#include "adjoint.hpp"
#include "affine.hpp"
#include "common.hpp"
#include "diagjac.hpp"
#include "fullscan.hpp"
#include "linalg.hpp"
#include "refine.hpp"
#include "scanfull.hpp"
#include "scanreg.hpp"
#include "tape.hpp"
#include "gauss.hpp"
#include "glm.hpp"
#include "tile_kernel.hpp"

using namespace pmcmc;

// Levels of the recursive reduce-then-scan: level l scans a D x L_l problem
// in tiles of TC_l columns; when it has more than one tile, its D x nT_l
// tile aggregates (agg_a, agg_b) form the level-(l+1) problem, whose output
// (states) holds the end state of every level-l tile.
struct ScanLevel {
    int64_t len = 0;
    int TC = 0, P = 0, nT = 0;
    size_t off_a = 0, off_b = 0, off_s = 0;  // byte offsets into the level blob
};

struct ScanPlan {
    int64_t D = 0;
    int nlev = 0;
    ScanLevel lev[24];
    size_t bytes = 0;
};

struct pmcmc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    // scratch of the standalone scan entry points (grown on demand)
    void* scan_blob = nullptr;
    size_t scan_blob_cap = 0;
    SyncState* host_sync = nullptr;  // pinned
    double* red_scratch = nullptr;   // partial sums of pmcmc_dot / pmcmc_dotsum
    unsigned long long* affine_red = nullptr;  // maxima and flag of the affine-scan checks
    Prof prof;
    // the creator's reference plus one per live target, workspace and batch
    // created on this context: pmcmc_ctx_destroy drops the creator's, the
    // stream and buffers go with the last one, so objects may be destroyed
    // in any order
    std::atomic<int> refs{1};
};

static void ctx_retain(pmcmc_ctx* c) { c->refs.fetch_add(1, std::memory_order_relaxed); }

// Scratch of the allocation-free scan entry point (pmcmc_scan_diag_f32_noalloc):
// one array per device, resident for the life of the process.  Host code
// never passes the array's name to a kernel or a copy (on the host it is a
// stand-in, not a device address): pmcmc_ctx_create looks up the array's
// device address once per device with hipGetSymbolAddress, and only that
// address is used.  Every call on a device shares its array, so calls that
// overlap on one device race.
static __device__ __align__(256) unsigned char scratch[PMCMC_NOALLOC_SCRATCH_BYTES];
static constexpr int PMCMC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[PMCMC_MAX_DEVICES] = {};

// device address of this device's scratch array (the device is current);
// looked up on the first context of the device
static hipError_t scratch_lookup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= PMCMC_MAX_DEVICES) return hipErrorInvalidDevice;
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[device] = ptr;
    return e;
}
static void* scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < PMCMC_MAX_DEVICES) ? g_scratch_ptr[device] : nullptr;
}
static void ctx_release(pmcmc_ctx* c);

enum { TGT_DIAG = 0, TGT_LOGISTIC = 1, TGT_GAUSS = 2 };

struct pmcmc_target {
    pmcmc_ctx* ctx = nullptr;
    bool ctx_ref = false;  // holds a context reference (set once created)
    int kind = TGT_DIAG;
    int64_t D = 0, N = 0;
    // diagonal Gaussian (both precisions; null = standard normal)
    float* prec_f = nullptr;
    float* mean_f = nullptr;
    double* prec_d = nullptr;
    double* mean_d = nullptr;
    // logistic regression
    GlmData glm;
    double prior_scale = 1.0;
    // dense Gaussian
    GaussData gauss;
};

struct pmcmc_deer_ws {
    pmcmc_ctx* ctx = nullptr;
    bool ctx_ref = false;  // holds a context reference (set once created)
    int64_t D = 0, T = 0;
    int prec = PMCMC_PREC_F32;
    size_t esz = 4;
    void* buf[2] = {nullptr, nullptr};
    SyncState* sync = nullptr;
    ScanPlan plan;
    void* blob = nullptr;  // scan levels
    size_t blob_bytes = 0;
    // GEMM-target buffers (allocated on first use)
    GlmWork glmw;
    FullWork fullw;  // jacobian = :full
    RefineWork refw;  // FP64 decision overrides (GEMM targets, FP32)
    DiagWork diagw;   // jacobian = :diag on a GEMM target
    // the last solve's final Newton iteration on a GEMM target: its iterate
    // S_k (state before each step), decisions and the x0 copy, from which
    // pmcmc_deer_last_trajectory assembles FT(S_k) and its log-densities
    bool last_gemm = false;
    const void* last_prev = nullptr;
    uint8_t* last_acc = nullptr;
    uint8_t* acc_ws = nullptr;  // decisions when the caller passes none
    void* x0_ws = nullptr;
    int8_t* prev_dec = nullptr;  // previous iteration's decisions (FP32 hysteresis)
    double* tile_smax = nullptr;  // max |S| per level-0 tile when it last ran (final prefix)
    int tile_smax_n = 0;
    // adjoint buffers (allocated on first use)
    void* adj_lam = nullptr;   // D x (T+1), last column zero
    void* adj_A = nullptr;     // reversed-time scan coefficients, D x T
    void* adj_B = nullptr;     // reversed-time residuals, D x T
    void* adj_D = nullptr;     // reversed-time corrections, D x T
    void* adj_zero = nullptr;  // D zeros (scan start)
    double* adj_part = nullptr;  // 2 T column partials + 2 sums
    ScanPlan adj_plan;
    void* adj_blob = nullptr;
};

static int fail(pmcmc_ctx* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

#define HIPCHK(ctx, call)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail((ctx), e_ == hipErrorOutOfMemory ? PMCMC_ERR_ALLOC : PMCMC_ERR_HIP, \
                        "%s: %s", #call, hipGetErrorString(e_));                            \
    } while (0)

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
__global__ void reset_sync_kernel(SyncState* s) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        s->done = 0u;
        s->err = 0u;
        s->dmax_bits = 0ull;
        s->smax_bits = 0ull;
        s->converged = 0;
        s->conv_iter = -1;
        s->last_metric = INFINITY;
        s->last_iter = -1;
        s->live_from = INT_MAX;
        s->skip_cols = 0;
        s->smax_prefix = 0.0;
    }
}

template <typename T>
__global__ void init_iterate_kernel(T* S, const T* s0, const T* init, int64_t D, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        S[i] = init ? init[i] : s0[i % D];
}

__global__ void fill_i8_kernel(int8_t* p, int8_t v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

__global__ void zero_ints_kernel(int* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

template <typename Tin, typename Tout>
__global__ void convert_kernel(Tout* out, const Tin* in, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (Tout)in[i];
}

// Standalone batched MALA step + surrogate JVP for diagonal-Hessian targets
// (src/MALA/MALA.jl:556-628); one group of G lanes per column.
template <typename T>
__global__ __launch_bounds__(256) void fwdjvp_diag_kernel(const T* X, const T* Xi, const T* U,
                                                          const T* Z, const T* prec, const T* mean,
                                                          T* FT, T* Jt, uint8_t* acc_out,
                                                          double* la_out, double eps, int64_t D,
                                                          int64_t Tn, int G) {
    const int lane = threadIdx.x & (G - 1);
    const int64_t col0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int64_t ncol_stride = ((int64_t)gridDim.x * blockDim.x) / G;
    const double sq2e = sqrt(2.0 * eps);
    const double c_norm = ((double)D / 2.0) * log(4.0 * 3.14159265358979323846 * eps);
    for (int64_t t = col0; t < Tn; t += ncol_stride) {
        double s_lpx = 0, s_lpy = 0, s_qyx = 0, s_qxy = 0, d1 = 0, d2 = 0, d3 = 0;
        for (int64_t d = lane; d < D; d += G) {
            const size_t gi = (size_t)t * D + d;
            const double p = prec ? (double)prec[d] : 1.0;
            const double m = mean ? (double)mean[d] : 0.0;
            const double x = (double)X[gi];
            const double gx = -p * (x - m);
            const double y = x + eps * gx + sq2e * (double)Xi[gi];
            const double gy = -p * (y - m);
            const double z = (double)Z[gi];
            const double w = z + eps * (-p * z);
            const double r = x - y - eps * gy;
            const double dr = z - w - eps * (-p * w);
            s_lpx += -0.5 * p * (x - m) * (x - m);
            s_lpy += -0.5 * p * (y - m) * (y - m);
            const double e1 = y - x - eps * gx;
            s_qyx += e1 * e1;
            s_qxy += r * r;
            d1 += gy * w;
            d2 += gx * z;
            d3 += r * dr;
        }
        s_lpx = group_sum(s_lpx, G);
        s_lpy = group_sum(s_lpy, G);
        s_qyx = group_sum(s_qyx, G);
        s_qxy = group_sum(s_qxy, G);
        d1 = group_sum(d1, G);
        d2 = group_sum(d2, G);
        d3 = group_sum(d3, G);
        const double la = (s_lpy + (-0.5 * s_qxy / (2.0 * eps) - c_norm)) -
                          (s_lpx + (-0.5 * s_qyx / (2.0 * eps) - c_norm));
        const double lu = log((double)U[t]);
        const double g = 1.0 / (1.0 + exp(-(la - lu)));
        const bool acc = lu < la;
        const double dg = g * (1.0 - g) * (d1 - d2 - d3 / (2.0 * eps));
        if (lane == 0) {
            if (acc_out) acc_out[t] = acc ? 1 : 0;
            if (la_out) la_out[t] = la;
        }
        for (int64_t d = lane; d < D; d += G) {
            const size_t gi = (size_t)t * D + d;
            const double p = prec ? (double)prec[d] : 1.0;
            const double m = mean ? (double)mean[d] : 0.0;
            const double x = (double)X[gi];
            const double y = x + eps * (-p * (x - m)) + sq2e * (double)Xi[gi];
            const double z = (double)Z[gi];
            const double w = z + eps * (-p * z);
            FT[gi] = (T)(acc ? y : x);
            Jt[gi] = (T)((acc ? w : z) + (y - x) * dg);
        }
    }
}

// Column log-densities of a diagonal Gaussian.
template <typename T>
__global__ __launch_bounds__(256) void logp_diag_kernel(const T* S, const T* prec, const T* mean,
                                                        double* out, int64_t D, int64_t Tn, int G) {
    const int lane = threadIdx.x & (G - 1);
    const int64_t col0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int64_t ncol_stride = ((int64_t)gridDim.x * blockDim.x) / G;
    for (int64_t t = col0; t < Tn; t += ncol_stride) {
        double s = 0;
        for (int64_t d = lane; d < D; d += G) {
            const double p = prec ? (double)prec[d] : 1.0;
            const double m = mean ? (double)mean[d] : 0.0;
            const double x = (double)S[(size_t)t * D + d];
            s += -0.5 * p * (x - m) * (x - m);
        }
        s = group_sum(s, G);
        if (lane == 0) out[t] = s;
    }
}

static int group_size(int64_t D) {
    int G = 1;
    while (G < D && G < 64) G <<= 1;
    return G;
}

static int grid_for(int64_t n, int block = 256, int cap = 8192) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// Tile shape of one level: TC columns so that TC*D is about the per-type
// element budget, shrunk until the LDS footprint fits 63 KB.
// ---------------------------------------------------------------------------
// Tile-shape table of the autotuner.  Level-0 tile shapes (TC columns, P
// parts, and for the standalone diagonal scan the single-pass look-back
// form) keyed by (kind, element size, D, ceil(log2 T)).  Filled by
// pmcmc_scan_tune / pmcmc_deer_tune (measured) or pmcmc_tune_set (a host's
// saved table); consulted by every plan, which falls back to tile_shape's
// rule when there is no entry or the entry does not fit.  Process-wide,
// guarded by a mutex.
// ---------------------------------------------------------------------------
struct TileChoice {
    int TC = 0, P = 0, chained = 0;
    double us = 0.0;
};

static int tune_bin(int64_t T) {
    int b = 0;
    while (b < 62 && ((int64_t)1 << b) < T) ++b;
    return b;
}

struct TuneTable {
    std::mutex mu;
    std::map<std::tuple<int, int, int64_t, int>, TileChoice> m;
};
static TuneTable& tune_table() {
    static TuneTable t;
    return t;
}
static bool tune_get(int kind, int esz, int64_t D, int64_t T, TileChoice* out) {
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    auto it = t.m.find(std::make_tuple(kind, esz, D, tune_bin(T)));
    if (it == t.m.end()) return false;
    *out = it->second;
    return true;
}
static void tune_put(int kind, int esz, int64_t D, int64_t T, const TileChoice& v) {
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    t.m[std::make_tuple(kind, esz, D, tune_bin(T))] = v;
}

// a level-0 shape the tile kernels can run: P parts fit the LDS part arrays
// (P * D <= max(D, PMCMC_NT)), the tile fits the LDS budget
// the register-tile single-pass forms: D = 64, tiles of 4 CPW columns; one
// tile per workgroup (chained == 2): CPW in {16, 32, 64} (FP32) or {16, 32}
// (FP64); persistent workgroups with the next tile in registers too
// (chained == 3): CPW in {16, 32}
// Row blocks of the register-tile scans for D rows (0: not offered): D is
// split into ceil(D / 64) blocks of at most 64 rows, each scanned as its own
// chain of tiles of tc columns, one row per lane.
static int reg_blocks(int64_t D) { return D >= 16 && D <= 256 ? (int)((D + 63) / 64) : 0; }

template <typename T>
static bool reg_tile_fits(int64_t D, const TileChoice& c) {
    if (!reg_blocks(D)) return false;
    if (c.chained == 3) return c.TC == 64 || c.TC == 128;
    return c.TC == 64 || c.TC == 128 || (c.TC == 256 && sizeof(T) == 4);
}

template <typename T>
using RegScanKernel = void (*)(TileArgs<T>);

template <typename T>
static RegScanKernel<T> reg_scan_kernel(int tc, bool persist) {
    if (persist) return tc == 64 ? scan_reg_persist_kernel<T, 16> : scan_reg_persist_kernel<T, 32>;
    if (tc == 64) return scan_reg_kernel<T, 16>;
    if (tc == 128) return scan_reg_kernel<T, 32>;
    return scan_reg_kernel<T, sizeof(T) == 4 ? 64 : 32>;
}

template <typename T>
static bool shape_fits(int64_t D, bool with_iterate, bool chained, const TileChoice& c) {
    if (c.chained >= 2) return !with_iterate && reg_tile_fits<T>(D, c);
    if (c.TC < 1 || c.P < 1 || c.P > c.TC) return false;
    const int64_t pmax = D >= PMCMC_NT ? 1 : PMCMC_NT / D;
    if (c.P > pmax) return false;
    return tile_lds_bytes<T>(D, c.TC, with_iterate, chained) <= 64512;
}

// a tuned DEER entry for (D, T), else the entry of D rounded up to a multiple
// of 16 (one tuning run covers a bucket of nearby widths: the shape's LDS
// only shrinks with D), when it fits this D
static bool deer_tuned(int esz, int64_t D, int64_t T, TileChoice* ch) {
    if (tune_get(PMCMC_TUNE_DEER, esz, D, T, ch)) return true;
    const int64_t D16 = (D + 15) / 16 * 16;
    return D16 != D && tune_get(PMCMC_TUNE_DEER, esz, D16, T, ch);
}

// level-0 tile of one window (L columns) of a batched chain: about 2048
// elements (half the single-chain rule's), not the single-chain tuner's
// entries.  Once only a few causally converging chains are left, each
// iteration is a launch over two or four tiles per chain, and a tile's
// reduce is one workgroup's chain of dependent FP64 work over its columns:
// half the elements per tile halve each thread's share of it.  On 256
// uneven chains: 6.48 s per solve at 2048 elements, 7.35 s at 4096, 6.74 s at
// 3072, 6.99 s at 1536; 1024 (tiles of 4 columns at D = 256) took 15.8 s.
static void batch_tile_shape(int64_t D, int64_t L, int* TC, int* P, int* nT) {
    tile_shape<float>(D, L, true, TC, P, nT);
    const int64_t tc = std::max<int64_t>(1, std::min<int64_t>(*TC, 2048 / D));
    *TC = (int)tc;
    *P = D >= PMCMC_NT ? 1 : (int)(PMCMC_NT / D);
    if (*P > *TC) *P = *TC;
    if (*P < 1) *P = 1;
    *nT = (int)((L + tc - 1) / tc);
}

template <typename T>
static bool plan_scan(ScanPlan* p, int64_t D, int64_t Tn, bool deer_top, const TileChoice* ov = nullptr) {
    p->D = D;
    p->nlev = 0;
    p->bytes = 0;
    int64_t len = Tn;
    for (int l = 0; l < 24; ++l) {
        ScanLevel& L = p->lev[l];
        L.len = len;
        if (l == 0 && ov && ov->TC > 0 && shape_fits<T>(D, deer_top, false, *ov)) {
            L.TC = (int)std::min<int64_t>(ov->TC, std::max<int64_t>(len, 1));
            L.P = std::min(ov->P, L.TC);
            L.nT = (int)((len + L.TC - 1) / L.TC);
        } else {
            tile_shape<T>(D, len, deer_top && l == 0, &L.TC, &L.P, &L.nT);
        }
        p->nlev = l + 1;
        if (L.nT <= 1) return tile_lds_bytes<T>(D, L.TC, deer_top && l == 0) <= 64512;
        const size_t m = (size_t)D * L.nT * sizeof(T);
        L.off_a = p->bytes;
        L.off_b = p->bytes + m;
        L.off_s = p->bytes + 2 * m;
        p->bytes += 3 * m;
        len = L.nT;
    }
    return false;
}

template <typename T, int MODE, int PASS>
static void launch_tile(const TileArgs<T>& a, int grid, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((tile_kernel<T, MODE, PASS>), dim3(grid), dim3(PMCMC_NT), lds, s, a);
}

// Scan levels l.. of a MODE_SCAN problem: S = scan(A, B) from s0.
template <typename T>
static void run_scan_levels(const ScanPlan& p, int l, const T* A, const T* B, T* S, const T* s0,
                            char* blob, const int* skip, hipStream_t s) {
    const ScanLevel& L = p.lev[l];
    TileArgs<T> a{};
    a.A = A;
    a.B = B;
    a.s0 = s0;
    a.S = S;
    a.D = p.D;
    a.T_len = L.len;
    a.TC = L.TC;
    a.numTiles = L.nT;
    a.P = L.P;
    a.skip = skip;
    a.damping = 1.0;
    const size_t lds = tile_lds_bytes<T>(p.D, L.TC, false);
    if (L.nT <= 1) {
        launch_tile<T, MODE_SCAN, PASS_SINGLE>(a, 1, lds, s);
        return;
    }
    a.agg_a = (T*)(blob + L.off_a);
    a.agg_b = (T*)(blob + L.off_b);
    launch_tile<T, MODE_SCAN, PASS_REDUCE>(a, L.nT, lds, s);
    run_scan_levels<T>(p, l + 1, a.agg_a, a.agg_b, (T*)(blob + L.off_s), s0, blob, skip, s);
    a.prev = (const T*)(blob + L.off_s);
    launch_tile<T, MODE_SCAN, PASS_APPLY>(a, L.nT, lds, s);
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" const char* pmcmc_version(void) { return "pmcmc_amd 0.1.0 (gfx950)"; }

extern "C" int pmcmc_ctx_create(int device, void* stream, pmcmc_ctx** out) {
    if (!out) return PMCMC_ERR_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return PMCMC_ERR_HIP;
    pmcmc_ctx* c = new (std::nothrow) pmcmc_ctx();
    if (!c) return PMCMC_ERR_ALLOC;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return PMCMC_ERR_HIP;
    }
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return PMCMC_ERR_HIP;
        }
        c->own_stream = true;
    }
    if (hipHostMalloc(&c->host_sync, sizeof(SyncState), hipHostMallocDefault) != hipSuccess) {
        pmcmc_ctx_destroy(c);
        return PMCMC_ERR_ALLOC;
    }
    if (scratch_lookup(device) != hipSuccess) {
        pmcmc_ctx_destroy(c);
        return PMCMC_ERR_HIP;
    }
    *out = c;
    return PMCMC_OK;
}

static void ctx_release(pmcmc_ctx* c) {
    if (c->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->scan_blob) hipFree(c->scan_blob);
    if (c->red_scratch) hipFree(c->red_scratch);
    if (c->affine_red) hipFree(c->affine_red);
    if (c->host_sync) hipHostFree(c->host_sync);
    if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int pmcmc_ctx_destroy(pmcmc_ctx* c) {
    if (!c) return PMCMC_OK;
    ctx_release(c);
    return PMCMC_OK;
}

extern "C" const char* pmcmc_last_error(const pmcmc_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" int pmcmc_ctx_synchronize(pmcmc_ctx* c) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PMCMC_OK;
}

extern "C" int pmcmc_ctx_set_profiling(pmcmc_ctx* c, int on) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    c->prof.on = on != 0;
    c->prof.kinds = ~0u;
    return PMCMC_OK;
}

extern "C" int pmcmc_ctx_set_profiling_kinds(pmcmc_ctx* c, unsigned kinds) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    if (kinds >> KT_COUNT) return fail(c, PMCMC_ERR_ARGUMENT, "bad kernel-time kind mask 0x%x", kinds);
    c->prof.on = kinds != 0;
    c->prof.kinds = kinds;
    return PMCMC_OK;
}

extern "C" int pmcmc_ctx_kernel_time(pmcmc_ctx* c, int kind, int64_t* launches, double* total_ms) {
    if (!c || kind < 0 || kind >= KT_COUNT) return fail(c, PMCMC_ERR_ARGUMENT, "bad kernel-time kind %d", kind);
    if (launches) *launches = c->prof.n[kind];
    if (total_ms) *total_ms = c->prof.ms[kind];
    return PMCMC_OK;
}

extern "C" int pmcmc_ctx_reset_kernel_times(pmcmc_ctx* c) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    c->prof.reset();
    return PMCMC_OK;
}

extern "C" int pmcmc_buffer_alloc(pmcmc_ctx* c, int64_t nbytes, void** out) {
    if (!c || !out || nbytes < 0) return fail(c, PMCMC_ERR_ARGUMENT, "bad buffer request");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMalloc(out, nbytes > 0 ? (size_t)nbytes : 1));
    return PMCMC_OK;
}
extern "C" int pmcmc_buffer_free(pmcmc_ctx* c, void* p) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    if (p) HIPCHK(c, hipFree(p));
    return PMCMC_OK;
}
extern "C" int pmcmc_buffer_upload(pmcmc_ctx* c, void* dst, const void* host, int64_t nbytes) {
    if (!c || nbytes < 0) return fail(c, PMCMC_ERR_ARGUMENT, "bad upload");
    if (nbytes == 0) return PMCMC_OK;
    HIPCHK(c, hipMemcpyAsync(dst, host, (size_t)nbytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PMCMC_OK;
}
extern "C" int pmcmc_buffer_download(pmcmc_ctx* c, void* host, const void* src, int64_t nbytes) {
    if (!c || nbytes < 0) return fail(c, PMCMC_ERR_ARGUMENT, "bad download");
    if (nbytes == 0) return PMCMC_OK;
    HIPCHK(c, hipMemcpyAsync(host, src, (size_t)nbytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PMCMC_OK;
}

// ---------------------------------------------------------------------------
// standalone diagonal scan
// ---------------------------------------------------------------------------
// Scratch of one standalone scan: the device's fixed array (noalloc; too
// small a problem's need is an argument error) or the context's blob, grown
// on demand.
static int scan_scratch(pmcmc_ctx* c, size_t need, bool noalloc, char** out) {
    if (noalloc) {
        char* p = (char*)scratch_of(c->device);
        if (!p) return fail(c, PMCMC_ERR_HIP, "scan: no scratch address for device %d", c->device);
        if (need > PMCMC_NOALLOC_SCRATCH_BYTES)
            return fail(c, PMCMC_ERR_ARGUMENT,
                        "scan: this D x T needs %zu bytes of scratch, above the allocation-free entry point's "
                        "%zu (use pmcmc_scan_diag_f32)",
                        need, (size_t)PMCMC_NOALLOC_SCRATCH_BYTES);
        *out = p;
        return PMCMC_OK;
    }
    if (need > c->scan_blob_cap) {
        if (c->scan_blob) HIPCHK(c, hipFree(c->scan_blob));
        c->scan_blob = nullptr;
        c->scan_blob_cap = 0;
        HIPCHK(c, hipMalloc(&c->scan_blob, need));
        c->scan_blob_cap = need;
    }
    *out = (char*)c->scan_blob;
    return PMCMC_OK;
}

// Look-back error word of a single-pass scan, read into the context's pinned
// word after the scan (an error means a bounded wait ran out).
static int scan_err_check(pmcmc_ctx* c, const unsigned* err_dev) {
    unsigned* hw = reinterpret_cast<unsigned*>(c->host_sync);
    HIPCHK(c, hipMemcpyAsync(hw, err_dev, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->prof.flush(0);
    if (*hw) return fail(c, PMCMC_ERR_DEVICE, "scan: look-back wait exceeded its bound");
    return PMCMC_OK;
}
// noalloc: the scratch is the device's fixed array (no allocation; a problem
// that needs more fails with PMCMC_ERR_ARGUMENT) instead of the context's
// blob, grown on demand; no kernel-time events are recorded.
template <typename T>
static int scan_diag(pmcmc_ctx* c, const T* A, const T* B, const T* s0, T* S, int64_t D,
                     int64_t Tn, const TileChoice* force = nullptr, bool noalloc = false) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    if (D < 1 || Tn < 0) return fail(c, PMCMC_ERR_DIMENSION, "scan: need D >= 1 and T >= 0 (D=%lld T=%lld)", (long long)D, (long long)Tn);
    if (Tn == 0) return PMCMC_OK;
    if (!A || !B || !s0 || !S) return fail(c, PMCMC_ERR_ARGUMENT, "scan: null pointer");
    const size_t bytes = (size_t)D * Tn * sizeof(T);
    auto overlap = [&](const void* p) {
        const char* a0 = (const char*)p;
        const char* b0 = (const char*)S;
        return a0 < b0 + bytes && b0 < a0 + bytes;
    };
    if (overlap(A)) return fail(c, PMCMC_ERR_ARGUMENT, "S must not alias A");
    if (overlap(B)) return fail(c, PMCMC_ERR_ARGUMENT, "S must not alias B");
    HIPCHK(c, hipSetDevice(c->device));
    // Shape: the caller's (the tuner's candidates), else the tuned table entry,
    // else the rule.  The single-pass form (decoupled look-back, one read of A
    // and B) is used when the table says so, or by default with
    // PMCMC_SCAN_CHAINED=1: on this GPU its look-back front (one window of
    // predecessors per coherent round trip) is slower than the two-pass
    // reduce/apply scan at D >= 64 and about even at D = 1.
    TileChoice ch;
    bool have = force ? (ch = *force, true) : tune_get(PMCMC_TUNE_SCAN_DIAG, (int)sizeof(T), D, Tn, &ch);
    bool chained;
    if (have) {
        chained = ch.chained != 0;
    } else {
        const char* tp = getenv("PMCMC_SCAN_CHAINED");
        chained = tp && tp[0] == '1';
    }
    if (have && ch.chained >= 2 && reg_tile_fits<T>(D, ch) && Tn > ch.TC) {
        const int TC = ch.TC;
        const int nb = reg_blocks(D);                 // row blocks, each its own chain of tiles
        const int nT = (int)((Tn + TC - 1) / TC);     // tiles per block
        const int nS = (nT + SCAN_SUPER - 1) / SCAN_SUPER;
        const int64_t ntk = (int64_t)nb * nT;         // tickets
        const size_t lbm = align256((size_t)64 * ntk * sizeof(T));
        const size_t sbm = align256((size_t)64 * nb * nS * sizeof(T));
        const size_t nint = (size_t)ntk + (size_t)nb * nS + 64;
        const size_t need = 2 * lbm + 3 * sbm + nint * sizeof(int);
        char* blob = nullptr;
        const int rb = scan_scratch(c, need, noalloc, &blob);
        if (rb != PMCMC_OK) return rb;
        TileArgs<T> a{};
        a.A = A;
        a.B = B;
        a.s0 = s0;
        a.S = S;
        a.D = D;
        a.T_len = Tn;
        a.TC = TC;
        a.numTiles = nT;
        a.lb_agg_a = (T*)blob;
        a.lb_agg_b = (T*)(blob + lbm);
        a.lb_incl = (T*)(blob + 2 * lbm);  // end states of the super-tiles
        a.lb_sagg_a = (T*)(blob + 2 * lbm + sbm);
        a.lb_sagg_b = (T*)(blob + 2 * lbm + 2 * sbm);
        int* ints = (int*)(blob + 2 * lbm + 3 * sbm);
        a.lb_status = ints;
        a.lb_sstatus = ints + ntk;
        a.lb_ticket = (unsigned*)(ints + ntk + (size_t)nb * nS);
        a.lb_err = (unsigned*)(ints + ntk + (size_t)nb * nS + 16);
        const int e = noalloc ? -1 : c->prof.begin(KT_ITERATION, 0, c->stream);
        hipLaunchKernelGGL(zero_ints_kernel, dim3(grid_for((int64_t)nint)), dim3(256), 0, c->stream, ints,
                           (int64_t)nint);
        const auto kern = reg_scan_kernel<T>(TC, ch.chained == 3);
        if (ch.chained == 3) {
            // persistent: no more workgroups than are resident at once
            int per_cu = 0, n_cu = 0;
            HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0));
            HIPCHK(c, hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, c->device));
            const int64_t resident = (int64_t)std::max(per_cu, 1) * std::max(n_cu, 1);
            const unsigned grid = (unsigned)std::min<int64_t>(ntk, resident);
            hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, c->stream, a);
        } else {
            hipLaunchKernelGGL(kern, dim3((unsigned)ntk), dim3(256), 0, c->stream, a);
        }
        c->prof.end(e, c->stream);
        HIPCHK(c, hipGetLastError());
        return scan_err_check(c, a.lb_err);
    }
    if (have && ch.chained >= 2) {  // register forms not applicable here: the rule
        have = false;
        chained = false;
    }
    ScanLevel L0;
    if (chained) {
        if (have && ch.TC > 0 && shape_fits<T>(D, false, true, ch)) {
            L0.TC = (int)std::min<int64_t>(ch.TC, Tn);
            L0.P = std::min(ch.P, L0.TC);
        } else {
            // bound by the look-back front: the widest tile that fits the LDS budget
            int tc = 1;
            while (tc * 2 <= Tn && tile_lds_bytes<T>(D, tc * 2, false, true) <= 64512) tc *= 2;
            while (tc + tc / 8 <= Tn && tile_lds_bytes<T>(D, tc + tc / 8, false, true) <= 64512 && tc >= 8)
                tc += tc / 8;
            L0.TC = tc;
            L0.P = D >= PMCMC_NT ? 1 : (int)(PMCMC_NT / D);
            if (L0.P > tc) L0.P = tc;
            if (L0.P < 1) L0.P = 1;
        }
        L0.nT = (int)((Tn + L0.TC - 1) / L0.TC);
        chained = L0.nT > 1 && tile_lds_bytes<T>(D, L0.TC, false, true) <= 64512;
    }
    ScanPlan plan;
    if (!chained && !plan_scan<T>(&plan, D, Tn, false, have && !ch.chained ? &ch : nullptr))
        return fail(c, PMCMC_ERR_UNSUPPORTED, "scan: D=%lld too large for one tile", (long long)D);
    const size_t lbm = (size_t)D * L0.nT * sizeof(T);
    const size_t lb_bytes = 3 * lbm + ((size_t)L0.nT + 64) * sizeof(int);
    const size_t need = chained ? lb_bytes : plan.bytes;
    char* blob = nullptr;
    const int rb = scan_scratch(c, need, noalloc, &blob);
    if (rb != PMCMC_OK) return rb;
    const int e = noalloc ? -1 : c->prof.begin(KT_ITERATION, 0, c->stream);
    if (chained) {
        TileArgs<T> a{};
        a.A = A;
        a.B = B;
        a.s0 = s0;
        a.S = S;
        a.D = D;
        a.T_len = Tn;
        a.TC = L0.TC;
        a.P = L0.P;
        a.numTiles = L0.nT;
        a.damping = 1.0;
        a.lb_agg_a = (T*)blob;
        a.lb_agg_b = (T*)(blob + lbm);
        a.lb_incl = (T*)(blob + 2 * lbm);
        int* ints = (int*)(blob + 3 * lbm);
        a.lb_status = ints;
        a.lb_ticket = (unsigned*)(ints + L0.nT);
        a.lb_err = (unsigned*)(ints + L0.nT + 16);
        hipLaunchKernelGGL(zero_ints_kernel, dim3(grid_for(L0.nT + 64)), dim3(256), 0, c->stream, ints,
                           (int64_t)L0.nT + 64);
        hipLaunchKernelGGL((tile_kernel_chained<T, MODE_SCAN>), dim3((unsigned)L0.nT), dim3(PMCMC_NT),
                           tile_lds_bytes<T>(D, L0.TC, false, true), c->stream, a);
        c->prof.end(e, c->stream);
        HIPCHK(c, hipGetLastError());
        return scan_err_check(c, a.lb_err);
    }
    run_scan_levels<T>(plan, 0, A, B, S, s0, blob, nullptr, c->stream);
    c->prof.end(e, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->prof.flush(0);
    return PMCMC_OK;
}

extern "C" int pmcmc_scan_diag_f32(pmcmc_ctx* c, const float* A, const float* B, const float* s0,
                                   float* S, int64_t D, int64_t T) {
    return scan_diag<float>(c, A, B, s0, S, D, T);
}
extern "C" int pmcmc_scan_diag_f32_noalloc(pmcmc_ctx* c, const float* A, const float* B, const float* s0,
                                           float* S, int64_t D, int64_t T) {
    return scan_diag<float>(c, A, B, s0, S, D, T, nullptr, true);
}
extern "C" int pmcmc_scan_diag_f64(pmcmc_ctx* c, const double* A, const double* B,
                                   const double* s0, double* S, int64_t D, int64_t T) {
    return scan_diag<double>(c, A, B, s0, S, D, T);
}

// sequential solve, residual and comparison of diagonal affine scans
// (src/DEER/DEERScan.jl:51-68, 203-258)
static int affine_args(pmcmc_ctx* c, const char* what, int64_t D, int64_t Tn) {
    if (D < 1 || Tn < 0)
        return fail(c, PMCMC_ERR_DIMENSION, "%s: need D >= 1 and T >= 0 (D=%lld T=%lld)", what, (long long)D,
                    (long long)Tn);
    return PMCMC_OK;
}

static int affine_red_get(pmcmc_ctx* c, const unsigned long long* red, double* out, int n) {
    unsigned long long h[3] = {0ull, 0ull, 0ull};
    HIPCHK(c, hipMemcpyAsync(h, red, (size_t)n * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < n && i < 2; ++i) memcpy(&out[i], &h[i], sizeof(double));
    if (n > 2) out[2] = h[2] ? 0.0 : 1.0;  // 1: every entry within atol + rtol |X|
    return PMCMC_OK;
}

static int affine_seq_entry(pmcmc_ctx* c, bool f64, const void* A, const void* B, const void* s0, void* S,
                            int64_t D, int64_t Tn) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    int rc = affine_args(c, "sequential affine solve", D, Tn);
    if (rc != PMCMC_OK) return rc;
    if (Tn == 0) return PMCMC_OK;
    if (!A || !B || !s0 || !S) return fail(c, PMCMC_ERR_ARGUMENT, "sequential affine solve: null buffer");
    HIPCHK(c, hipSetDevice(c->device));
    if (affine_seq(c->stream, f64, A, B, s0, S, D, Tn) != 0)
        return fail(c, PMCMC_ERR_HIP, "sequential affine solve: launch failed");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PMCMC_OK;
}
extern "C" int pmcmc_affine_seq_f32(pmcmc_ctx* c, const float* A, const float* B, const float* s0, float* S,
                                    int64_t D, int64_t T) {
    return affine_seq_entry(c, false, A, B, s0, S, D, T);
}
extern "C" int pmcmc_affine_seq_f64(pmcmc_ctx* c, const double* A, const double* B, const double* s0, double* S,
                                    int64_t D, int64_t T) {
    return affine_seq_entry(c, true, A, B, s0, S, D, T);
}

static int affine_residual_entry(pmcmc_ctx* c, bool f64, const void* S, const void* A, const void* B,
                                 const void* s0, int64_t D, int64_t Tn, double* out) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    int rc = affine_args(c, "affine scan residual", D, Tn);
    if (rc != PMCMC_OK) return rc;
    if (!out) return fail(c, PMCMC_ERR_ARGUMENT, "affine scan residual: null output");
    if (Tn == 0) {
        *out = 0.0;
        return PMCMC_OK;
    }
    if (!A || !B || !s0 || !S) return fail(c, PMCMC_ERR_ARGUMENT, "affine scan residual: null buffer");
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->affine_red) HIPCHK(c, hipMalloc(&c->affine_red, 3 * sizeof(unsigned long long)));
    if (affine_residual(c->stream, f64, S, A, B, s0, D, Tn, c->affine_red) != 0)
        return fail(c, PMCMC_ERR_HIP, "affine scan residual: launch failed");
    return affine_red_get(c, c->affine_red, out, 1);
}
extern "C" int pmcmc_affine_residual_f32(pmcmc_ctx* c, const float* S, const float* A, const float* B,
                                         const float* s0, int64_t D, int64_t T, double* out) {
    return affine_residual_entry(c, false, S, A, B, s0, D, T, out);
}
extern "C" int pmcmc_affine_residual_f64(pmcmc_ctx* c, const double* S, const double* A, const double* B,
                                         const double* s0, int64_t D, int64_t T, double* out) {
    return affine_residual_entry(c, true, S, A, B, s0, D, T, out);
}

static int affine_compare_entry(pmcmc_ctx* c, bool f64, const void* X, const void* Y, int64_t n, double atol,
                                double rtol, double* out) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    if (n < 0) return fail(c, PMCMC_ERR_DIMENSION, "affine scan compare: negative length");
    if (!out) return fail(c, PMCMC_ERR_ARGUMENT, "affine scan compare: null output");
    if (n == 0) {
        out[0] = out[1] = 0.0;
        out[2] = 1.0;
        return PMCMC_OK;
    }
    if (!X || !Y) return fail(c, PMCMC_ERR_ARGUMENT, "affine scan compare: null buffer");
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->affine_red) HIPCHK(c, hipMalloc(&c->affine_red, 3 * sizeof(unsigned long long)));
    if (affine_compare(c->stream, f64, X, Y, n, atol, rtol, c->affine_red) != 0)
        return fail(c, PMCMC_ERR_HIP, "affine scan compare: launch failed");
    return affine_red_get(c, c->affine_red, out, 3);
}
extern "C" int pmcmc_affine_compare_f32(pmcmc_ctx* c, const float* X, const float* Y, int64_t n, double atol,
                                        double rtol, double* out) {
    return affine_compare_entry(c, false, X, Y, n, atol, rtol, out);
}
extern "C" int pmcmc_affine_compare_f64(pmcmc_ctx* c, const double* X, const double* Y, int64_t n, double atol,
                                        double rtol, double* out) {
    return affine_compare_entry(c, true, X, Y, n, atol, rtol, out);
}

// dense-Jacobian scan with materialised Jacobians (no reference counterpart)
static int scan_full_entry(pmcmc_ctx* c, bool f64, const void* J, const void* B, const void* s0, void* S, int64_t D,
                           int64_t Tn) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    if (D < 1 || Tn < 0)
        return fail(c, PMCMC_ERR_DIMENSION, "dense scan: need D >= 1 and T >= 0 (D=%lld T=%lld)", (long long)D,
                    (long long)Tn);
    if (D > 64) return fail(c, PMCMC_ERR_UNSUPPORTED, "dense scan: D=%lld > 64", (long long)D);
    if (Tn == 0) return PMCMC_OK;
    if (!J || !B || !s0 || !S) return fail(c, PMCMC_ERR_ARGUMENT, "dense scan: null pointer");
    const size_t esz = f64 ? 8 : 4;
    const size_t sb = (size_t)D * Tn * esz;
    auto overlap = [&](const void* p, size_t pb) {
        const char* a0 = (const char*)p;
        const char* b0 = (const char*)S;
        return a0 < b0 + sb && b0 < a0 + pb;
    };
    if (overlap(J, sb * (size_t)D)) return fail(c, PMCMC_ERR_ARGUMENT, "S must not alias J");
    if (overlap(B, sb)) return fail(c, PMCMC_ERR_ARGUMENT, "S must not alias B");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t need = scan_full_scratch_bytes(Tn, f64);
    if (need > c->scan_blob_cap) {
        if (c->scan_blob) HIPCHK(c, hipFree(c->scan_blob));
        c->scan_blob = nullptr;
        c->scan_blob_cap = 0;
        HIPCHK(c, hipMalloc(&c->scan_blob, need));
        c->scan_blob_cap = need;
    }
    std::string msg;
    const int e = c->prof.begin(KT_ITERATION, 0, c->stream);
    const int rc = scan_full(f64, J, B, s0, S, D, Tn, c->scan_blob, c->stream, &msg);
    c->prof.end(e, c->stream);
    if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->prof.flush(0);
    return PMCMC_OK;
}

extern "C" int pmcmc_scan_full_f32(pmcmc_ctx* c, const float* J, const float* B, const float* s0, float* S,
                                   int64_t D, int64_t T) {
    return scan_full_entry(c, false, J, B, s0, S, D, T);
}

extern "C" int pmcmc_scan_full_f64(pmcmc_ctx* c, const double* J, const double* B, const double* s0, double* S,
                                   int64_t D, int64_t T) {
    return scan_full_entry(c, true, J, B, s0, S, D, T);
}

// ---------------------------------------------------------------------------
// targets
// ---------------------------------------------------------------------------
static int copy_vec_both(pmcmc_ctx* c, const void* src, int src_prec, int64_t n, float** f,
                         double** d) {
    if (!src) return PMCMC_OK;
    HIPCHK(c, hipMalloc(f, n * sizeof(float)));
    HIPCHK(c, hipMalloc(d, n * sizeof(double)));
    if (src_prec == PMCMC_PREC_F64) {
        HIPCHK(c, hipMemcpyAsync(*d, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(n)), dim3(256), 0, c->stream, *f, (const double*)src, n);
    } else {
        HIPCHK(c, hipMemcpyAsync(*f, src, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid_for(n)), dim3(256), 0, c->stream, *d, (const float*)src, n);
    }
    HIPCHK(c, hipGetLastError());
    return PMCMC_OK;
}

static bool valid_prec(int p) { return p == PMCMC_PREC_F64 || p == PMCMC_PREC_F32 || p == PMCMC_PREC_BF16_JAC; }

extern "C" int pmcmc_target_diag_gaussian(pmcmc_ctx* c, int64_t D, int dtype_prec, const void* prec,
                                          const void* mean, pmcmc_target** out) {
    if (!c || !out) return PMCMC_ERR_ARGUMENT;
    *out = nullptr;
    if (D < 1) return fail(c, PMCMC_ERR_DIMENSION, "target dim must be >= 1");
    if (!valid_prec(dtype_prec)) return fail(c, PMCMC_ERR_ARGUMENT, "bad precision %d", dtype_prec);
    HIPCHK(c, hipSetDevice(c->device));
    pmcmc_target* t = new (std::nothrow) pmcmc_target();
    if (!t) return fail(c, PMCMC_ERR_ALLOC, "host allocation failed");
    t->ctx = c;
    t->kind = TGT_DIAG;
    t->D = D;
    int rc = copy_vec_both(c, prec, dtype_prec, D, &t->prec_f, &t->prec_d);
    if (rc == PMCMC_OK) rc = copy_vec_both(c, mean, dtype_prec, D, &t->mean_f, &t->mean_d);
    if (rc == PMCMC_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, PMCMC_ERR_HIP, "sync failed");
    if (rc != PMCMC_OK) {
        pmcmc_target_destroy(t);
        return rc;
    }
    ctx_retain(c);
    t->ctx_ref = true;
    *out = t;
    return PMCMC_OK;
}

extern "C" int pmcmc_target_logistic(pmcmc_ctx* c, int64_t N, int64_t D, int dtype_prec,
                                     const void* Xd, const void* y, double prior_scale,
                                     pmcmc_target** out) {
    if (!c || !out) return PMCMC_ERR_ARGUMENT;
    *out = nullptr;
    if (N < 1 || D < 1) return fail(c, PMCMC_ERR_DIMENSION, "logistic target needs N >= 1 and D >= 1");
    if (!Xd || !y) return fail(c, PMCMC_ERR_ARGUMENT, "logistic target: null data");
    if (!(prior_scale > 0)) return fail(c, PMCMC_ERR_ARGUMENT, "prior_scale must be > 0");
    if (!valid_prec(dtype_prec)) return fail(c, PMCMC_ERR_ARGUMENT, "bad precision %d", dtype_prec);
    HIPCHK(c, hipSetDevice(c->device));
    pmcmc_target* t = new (std::nothrow) pmcmc_target();
    if (!t) return fail(c, PMCMC_ERR_ALLOC, "host allocation failed");
    t->ctx = c;
    t->kind = TGT_LOGISTIC;
    t->D = D;
    t->N = N;
    t->prior_scale = prior_scale;
    std::string msg;
    int rc = glm_data_create(&t->glm, c->stream, N, D, dtype_prec == PMCMC_PREC_F64, Xd, y, &msg);
    if (rc != PMCMC_OK) {
        fail(c, rc, "%s", msg.c_str());
        pmcmc_target_destroy(t);
        return rc;
    }
    ctx_retain(c);
    t->ctx_ref = true;
    *out = t;
    return PMCMC_OK;
}

extern "C" int pmcmc_target_logistic_set_prior(pmcmc_target* t, const void* P, int dtype_prec) {
    if (!t || !t->ctx) return PMCMC_ERR_ARGUMENT;
    pmcmc_ctx* c = t->ctx;
    if (t->kind != TGT_LOGISTIC) return fail(c, PMCMC_ERR_ARGUMENT, "prior precision: not a logistic target");
    if (!P) return fail(c, PMCMC_ERR_ARGUMENT, "prior precision: null matrix");
    if (!valid_prec(dtype_prec)) return fail(c, PMCMC_ERR_ARGUMENT, "bad precision %d", dtype_prec);
    HIPCHK(c, hipSetDevice(c->device));
    std::string msg;
    const int rc = glm_data_set_prior(&t->glm, c->stream, dtype_prec == PMCMC_PREC_F64, P, &msg);
    if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    return PMCMC_OK;
}

extern "C" int pmcmc_target_gaussian(pmcmc_ctx* c, int64_t D, int dtype_prec, const void* Lambda,
                                     const void* mean, pmcmc_target** out) {
    if (!c || !out) return PMCMC_ERR_ARGUMENT;
    *out = nullptr;
    if (D < 1) return fail(c, PMCMC_ERR_DIMENSION, "target dim must be >= 1");
    if (!Lambda) return fail(c, PMCMC_ERR_ARGUMENT, "gaussian target: null precision matrix");
    if (!valid_prec(dtype_prec)) return fail(c, PMCMC_ERR_ARGUMENT, "bad precision %d", dtype_prec);
    HIPCHK(c, hipSetDevice(c->device));
    pmcmc_target* t = new (std::nothrow) pmcmc_target();
    if (!t) return fail(c, PMCMC_ERR_ALLOC, "host allocation failed");
    t->ctx = c;
    t->kind = TGT_GAUSS;
    t->D = D;
    std::string msg;
    int rc = gauss_data_create(&t->gauss, c->stream, D, dtype_prec == PMCMC_PREC_F64, Lambda, mean, &msg);
    if (rc != PMCMC_OK) {
        fail(c, rc, "%s", msg.c_str());
        pmcmc_target_destroy(t);
        return rc;
    }
    ctx_retain(c);
    t->ctx_ref = true;
    *out = t;
    return PMCMC_OK;
}

extern "C" int pmcmc_target_destroy(pmcmc_target* t) {
    if (!t) return PMCMC_OK;
    if (t->ctx) hipSetDevice(t->ctx->device);
    if (t->ctx && t->ctx->stream) hipStreamSynchronize(t->ctx->stream);
    hipFree(t->prec_f);
    hipFree(t->mean_f);
    hipFree(t->prec_d);
    hipFree(t->mean_d);
    glm_data_destroy(&t->glm);
    gauss_data_destroy(&t->gauss);
    pmcmc_ctx* c = t->ctx_ref ? t->ctx : nullptr;
    delete t;
    if (c) ctx_release(c);
    return PMCMC_OK;
}

extern "C" int64_t pmcmc_target_dim(const pmcmc_target* t) { return t ? t->D : -1; }

// ---------------------------------------------------------------------------
// batched MALA step + JVP
// ---------------------------------------------------------------------------
extern "C" int pmcmc_mala_fwd_and_jvp(pmcmc_ctx* c, const pmcmc_target* tgt, int prec, double eps,
                                      const void* X, const void* Xi, const void* U, const void* Z,
                                      void* FT, void* Jt, uint8_t* accepted, double* logalpha,
                                      int64_t D, int64_t Tn) {
    if (!c || !tgt) return fail(c, PMCMC_ERR_ARGUMENT, "null context/target");
    if (D != tgt->D) return fail(c, PMCMC_ERR_DIMENSION, "X has %lld rows but the target has dim %lld", (long long)D, (long long)tgt->D);
    if (Tn < 0) return fail(c, PMCMC_ERR_DIMENSION, "T must be >= 0");
    if (!valid_prec(prec)) return fail(c, PMCMC_ERR_ARGUMENT, "bad precision %d", prec);
    if (!(eps > 0)) return fail(c, PMCMC_ERR_ARGUMENT, "eps must be > 0");
    if (Tn == 0) return PMCMC_OK;
    if (!X || !Xi || !U || !Z || !FT || !Jt) return fail(c, PMCMC_ERR_ARGUMENT, "null buffer");
    HIPCHK(c, hipSetDevice(c->device));
    const int G = group_size(D);
    const int grid = grid_for(Tn * G);
    if (tgt->kind == TGT_DIAG) {
        if (prec == PMCMC_PREC_F64)
            hipLaunchKernelGGL((fwdjvp_diag_kernel<double>), dim3(grid), dim3(256), 0, c->stream,
                               (const double*)X, (const double*)Xi, (const double*)U, (const double*)Z,
                               tgt->prec_d, tgt->mean_d, (double*)FT, (double*)Jt, accepted, logalpha,
                               eps, D, Tn, G);
        else
            hipLaunchKernelGGL((fwdjvp_diag_kernel<float>), dim3(grid), dim3(256), 0, c->stream,
                               (const float*)X, (const float*)Xi, (const float*)U, (const float*)Z,
                               tgt->prec_f, tgt->mean_f, (float*)FT, (float*)Jt, accepted, logalpha,
                               eps, D, Tn, G);
        HIPCHK(c, hipGetLastError());
    } else if (tgt->kind == TGT_GAUSS) {
        std::string msg;
        int rc = gauss_fwd_and_jvp(&tgt->gauss, c->stream, prec, eps, X, Xi, U, Z, FT, Jt, accepted, logalpha,
                                   D, Tn, &msg);
        if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    } else {
        std::string msg;
        int rc = glm_fwd_and_jvp(&tgt->glm, c->stream, prec, eps, tgt->prior_scale, X, Xi, U, Z, FT,
                                 Jt, accepted, logalpha, D, Tn, &msg);
        if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PMCMC_OK;
}

extern "C" int pmcmc_mala_run_batched(pmcmc_ctx* c, const pmcmc_target* tgt, int prec, double eps,
                                      const void* X0, const void* Xi, const void* U, void* S_out,
                                      uint8_t* accepted, int64_t D, int64_t N, int64_t K) {
    if (!c || !tgt) return fail(c, PMCMC_ERR_ARGUMENT, "null context/target");
    if (D != tgt->D) return fail(c, PMCMC_ERR_DIMENSION, "X0 has %lld rows but the target has dim %lld", (long long)D, (long long)tgt->D);
    if (N < 0 || K < 0) return fail(c, PMCMC_ERR_DIMENSION, "N and K must be >= 0");
    if (!valid_prec(prec)) return fail(c, PMCMC_ERR_ARGUMENT, "bad precision %d", prec);
    if (!(eps > 0)) return fail(c, PMCMC_ERR_ARGUMENT, "eps must be > 0");
    if (N == 0 || K == 0) return PMCMC_OK;
    if (!X0 || !Xi || !U || !S_out || !accepted) return fail(c, PMCMC_ERR_ARGUMENT, "null buffer");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t esz = prec == PMCMC_PREC_F64 ? 8 : 4;
    void* Z = nullptr;
    void* Jw = nullptr;
    if (hipMalloc(&Z, (size_t)D * N * esz) != hipSuccess || hipMalloc(&Jw, (size_t)D * N * esz) != hipSuccess) {
        (void)hipFree(Z);
        return fail(c, PMCMC_ERR_ALLOC, "MALA run: scratch allocation failed");
    }
    int rc = PMCMC_OK;
    std::string msg;
    if (hipMemsetAsync(Z, 0, (size_t)D * N * esz, c->stream) != hipSuccess) {
        rc = PMCMC_ERR_HIP;
        msg = "MALA run: zero fill failed";
    } else if (tgt->kind == TGT_DIAG) {
        const int G = group_size(D);
        const int grid = grid_for(N * G);
        for (int64_t k = 0; k < K; ++k) {
            const size_t o = (size_t)k * D * N;
            if (prec == PMCMC_PREC_F64) {
                const double* X = k == 0 ? (const double*)X0 : (const double*)S_out + o - (size_t)D * N;
                hipLaunchKernelGGL((fwdjvp_diag_kernel<double>), dim3(grid), dim3(256), 0, c->stream, X,
                                   (const double*)Xi + o, (const double*)U + (size_t)k * N, (const double*)Z,
                                   tgt->prec_d, tgt->mean_d, (double*)S_out + o, (double*)Jw, accepted + (size_t)k * N,
                                   nullptr, eps, D, N, G);
            } else {
                const float* X = k == 0 ? (const float*)X0 : (const float*)S_out + o - (size_t)D * N;
                hipLaunchKernelGGL((fwdjvp_diag_kernel<float>), dim3(grid), dim3(256), 0, c->stream, X,
                                   (const float*)Xi + o, (const float*)U + (size_t)k * N, (const float*)Z,
                                   tgt->prec_f, tgt->mean_f, (float*)S_out + o, (float*)Jw, accepted + (size_t)k * N,
                                   nullptr, eps, D, N, G);
            }
        }
        if (hipGetLastError() != hipSuccess) {
            rc = PMCMC_ERR_HIP;
            msg = "MALA run: kernel launch failed";
        }
    } else if (tgt->kind == TGT_GAUSS) {
        rc = gauss_mala_run(&tgt->gauss, c->stream, prec, eps, X0, Xi, U, K, D, N, S_out, accepted, Z, Jw, &msg);
    } else {
        rc = glm_mala_run(&tgt->glm, c->stream, prec, eps, tgt->prior_scale, X0, Xi, U, K, D, N, S_out, accepted, Z,
                          Jw, &msg);
    }
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(Z);
    (void)hipFree(Jw);
    if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    return PMCMC_OK;
}

// ---------------------------------------------------------------------------
// owned matmul and reductions (src/ParallelMCMC.jl:23-25)
// ---------------------------------------------------------------------------
static bool linalg_prec(int prec) { return prec == PMCMC_PREC_F32 || prec == PMCMC_PREC_F64; }

extern "C" int pmcmc_matmul(pmcmc_ctx* c, int prec, int trans_a, int trans_b, int64_t m, int64_t n, int64_t k,
                            const void* A, int64_t lda, const void* B, int64_t ldb, void* C, int64_t ldc) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    if (!linalg_prec(prec)) return fail(c, PMCMC_ERR_ARGUMENT, "matmul: precision must be F32 or F64");
    if (m < 0 || n < 0 || k < 0) return fail(c, PMCMC_ERR_DIMENSION, "matmul: negative size");
    const int64_t rows_a = trans_a ? k : m, rows_b = trans_b ? n : k;
    if (lda < (rows_a > 1 ? rows_a : 1) || ldb < (rows_b > 1 ? rows_b : 1) || ldc < (m > 1 ? m : 1))
        return fail(c, PMCMC_ERR_DIMENSION, "matmul: leading dimension smaller than the stored rows");
    if (m > 0 && n > 0 && (!C || (k > 0 && (!A || !B))))
        return fail(c, PMCMC_ERR_ARGUMENT, "matmul: null operand");
    HIPCHK(c, hipSetDevice(c->device));
    std::string msg;
    const int rc = linalg_gemm(prec, trans_a != 0, trans_b != 0, m, n, k, A, lda, B, ldb, C, ldc, c->stream, &msg);
    if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PMCMC_OK;
}

extern "C" int pmcmc_dotsum(pmcmc_ctx* c, int prec, int64_t m, int64_t n, const void* A, int64_t lda,
                            const void* B, int64_t ldb, void* out) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    if (!linalg_prec(prec)) return fail(c, PMCMC_ERR_ARGUMENT, "dotsum: precision must be F32 or F64");
    if (m < 0 || n < 0) return fail(c, PMCMC_ERR_DIMENSION, "dotsum: negative size");
    if (lda < (m > 1 ? m : 1) || ldb < (m > 1 ? m : 1))
        return fail(c, PMCMC_ERR_DIMENSION, "dotsum: leading dimension smaller than m");
    if (!out || (m > 0 && n > 0 && (!A || !B))) return fail(c, PMCMC_ERR_ARGUMENT, "dotsum: null operand");
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->red_scratch) HIPCHK(c, hipMalloc(&c->red_scratch, linalg_dotsum_scratch_bytes()));
    std::string msg;
    const int rc = linalg_dotsum(prec, m, n, A, lda, B, ldb, out, c->red_scratch, c->stream, &msg);
    if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PMCMC_OK;
}

extern "C" int pmcmc_dot(pmcmc_ctx* c, int prec, int64_t n, const void* a, const void* b, void* out) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    if (n < 0) return fail(c, PMCMC_ERR_DIMENSION, "dot: negative length");
    return pmcmc_dotsum(c, prec, n, 1, a, n > 1 ? n : 1, b, n > 1 ? n : 1, out);
}

extern "C" int pmcmc_target_logdensity_batch(pmcmc_ctx* c, const pmcmc_target* tgt, int prec,
                                             const void* S, double* out, int64_t Tn) {
    if (!c || !tgt) return fail(c, PMCMC_ERR_ARGUMENT, "null context/target");
    if (!valid_prec(prec)) return fail(c, PMCMC_ERR_ARGUMENT, "bad precision %d", prec);
    if (Tn < 0) return fail(c, PMCMC_ERR_DIMENSION, "T must be >= 0");
    if (Tn == 0) return PMCMC_OK;
    if (!S || !out) return fail(c, PMCMC_ERR_ARGUMENT, "null buffer");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t D = tgt->D;
    if (tgt->kind == TGT_DIAG) {
        const int G = group_size(D);
        const int grid = grid_for(Tn * G);
        if (prec == PMCMC_PREC_F64)
            hipLaunchKernelGGL((logp_diag_kernel<double>), dim3(grid), dim3(256), 0, c->stream,
                               (const double*)S, tgt->prec_d, tgt->mean_d, out, D, Tn, G);
        else
            hipLaunchKernelGGL((logp_diag_kernel<float>), dim3(grid), dim3(256), 0, c->stream,
                               (const float*)S, tgt->prec_f, tgt->mean_f, out, D, Tn, G);
        HIPCHK(c, hipGetLastError());
    } else if (tgt->kind == TGT_GAUSS) {
        std::string msg;
        int rc = gauss_logdensity(&tgt->gauss, c->stream, prec, S, out, Tn, &msg);
        if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    } else {
        std::string msg;
        int rc = glm_logdensity(&tgt->glm, c->stream, prec, tgt->prior_scale, S, out, Tn, &msg);
        if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PMCMC_OK;
}

// ---------------------------------------------------------------------------
// DEER workspace + solve
// ---------------------------------------------------------------------------
extern "C" int pmcmc_deer_ws_create(pmcmc_ctx* c, int64_t D, int64_t Tn, int prec,
                                    pmcmc_deer_ws** out) {
    if (!c || !out) return PMCMC_ERR_ARGUMENT;
    *out = nullptr;
    if (D < 1 || Tn < 1) return fail(c, PMCMC_ERR_DIMENSION, "workspace needs D >= 1 and T >= 1");
    if (!valid_prec(prec)) return fail(c, PMCMC_ERR_ARGUMENT, "bad precision %d", prec);
    HIPCHK(c, hipSetDevice(c->device));
    pmcmc_deer_ws* w = new (std::nothrow) pmcmc_deer_ws();
    if (!w) return fail(c, PMCMC_ERR_ALLOC, "host allocation failed");
    w->ctx = c;
    w->D = D;
    w->T = Tn;
    w->prec = prec;
    w->esz = prec == PMCMC_PREC_F64 ? 8 : 4;
    TileChoice tuned;
    const bool have = deer_tuned((int)w->esz, D, Tn, &tuned);
    const bool ok_plan = w->esz == 8 ? plan_scan<double>(&w->plan, D, Tn, true, have ? &tuned : nullptr)
                                     : plan_scan<float>(&w->plan, D, Tn, true, have ? &tuned : nullptr);
    if (!ok_plan) {
        delete w;
        return fail(c, PMCMC_ERR_UNSUPPORTED, "D=%lld too large for the fused tile kernel", (long long)D);
    }
    const size_t mat = (size_t)D * Tn * w->esz;
    bool ok = hipMalloc(&w->buf[0], mat) == hipSuccess && hipMalloc(&w->buf[1], mat) == hipSuccess &&
              hipMalloc(&w->sync, sizeof(SyncState)) == hipSuccess &&
              hipMemset(w->sync, 0, sizeof(SyncState)) == hipSuccess &&
              hipMalloc(&w->blob, w->plan.bytes > 0 ? w->plan.bytes : 16) == hipSuccess;
    w->blob_bytes = w->plan.bytes > 0 ? w->plan.bytes : 16;
    if (!ok) {
        pmcmc_deer_ws_destroy(w);
        return fail(c, PMCMC_ERR_ALLOC, "workspace allocation failed (D=%lld T=%lld)", (long long)D, (long long)Tn);
    }
    ctx_retain(c);
    w->ctx_ref = true;
    *out = w;
    return PMCMC_OK;
}

extern "C" int pmcmc_deer_ws_destroy(pmcmc_deer_ws* w) {
    if (!w) return PMCMC_OK;
    if (w->ctx) {
        hipSetDevice(w->ctx->device);
        hipStreamSynchronize(w->ctx->stream);
    }
    hipFree(w->buf[0]);
    hipFree(w->buf[1]);
    hipFree(w->sync);
    hipFree(w->blob);
    glm_work_destroy(&w->glmw);
    full_work_destroy(&w->fullw);
    refine_work_destroy(&w->refw);
    diag_work_destroy(&w->diagw);
    hipFree(w->acc_ws);
    hipFree(w->x0_ws);
    hipFree(w->prev_dec);
    hipFree(w->tile_smax);
    hipFree(w->adj_lam);
    hipFree(w->adj_A);
    hipFree(w->adj_B);
    hipFree(w->adj_D);
    hipFree(w->adj_zero);
    hipFree(w->adj_part);
    hipFree(w->adj_blob);
    pmcmc_ctx* c = w->ctx_ref ? w->ctx : nullptr;
    delete w;
    if (c) ctx_release(c);
    return PMCMC_OK;
}

// One Newton iteration on the device: [GEMM products] -> level-0 reduce ->
// scan of the tile aggregates -> level-0 apply (damping, maxima, metric).
template <typename T, int MODE>
static void launch_deer_iteration(pmcmc_deer_ws* w, TileArgs<T> a, hipStream_t s, Prof* prof) {
    const ScanPlan& p = w->plan;
    const ScanLevel& L = p.lev[0];
    char* blob = (char*)w->blob;
    a.TC = L.TC;
    a.P = L.P;
    a.numTiles = L.nT;
    a.skip = &w->sync->converged;
    const size_t lds = tile_lds_bytes<T>(p.D, L.TC, true);
    if (L.nT <= 1) {
        const int e = prof->begin(KT_TILE_APPLY, a.iter, s);
        launch_tile<T, MODE, PASS_SINGLE>(a, 1, lds, s);
        prof->end(e, s);
        return;
    }
    a.agg_a = (T*)(blob + L.off_a);
    a.agg_b = (T*)(blob + L.off_b);
    if (w->tile_smax && w->tile_smax_n >= L.nT) {
        a.final_cols = &w->sync->skip_cols;
        a.tile_smax = w->tile_smax;
    }
    if (MODE == MODE_GEMM) {
        // the reduce pass keeps A and B in the probe-product buffers it has
        // consumed (rebuilt by the target kernels every iteration)
        a.a_keep = const_cast<T*>(a.W);
        a.b_keep = const_cast<T*>(a.hw);
    }
    int e = prof->begin(KT_TILE_REDUCE, a.iter, s);
    launch_tile<T, MODE, PASS_REDUCE>(a, L.nT, lds, s);
    prof->end(e, s);
    e = prof->begin(KT_SCAN_SPINE, a.iter, s);
    // the corrections start from zero (x0 is fixed)
    run_scan_levels<T>(p, 1, a.agg_a, a.agg_b, (T*)(blob + L.off_s), nullptr, blob, a.skip, s);
    prof->end(e, s);
    a.prev = (const T*)(blob + L.off_s);
    e = prof->begin(KT_TILE_APPLY, a.iter, s);
    launch_tile<T, MODE, PASS_APPLY>(a, L.nT, lds, s);
    prof->end(e, s);
}

template <typename T>
static int deer_solve_t(pmcmc_ctx* c, pmcmc_deer_ws* w, const pmcmc_target* tgt,
                        const pmcmc_deer_params* p, const void* x0, const void* Xi, const void* U,
                        const void* init, void* S_out, uint8_t* acc_out, pmcmc_deer_info* info) {
    const int64_t D = w->D, Tn = w->T;
    const int mode = tgt->kind == TGT_DIAG ? MODE_ELEM : MODE_GEMM;
    const bool full = p->jacobian == PMCMC_JAC_FULL || p->jacobian == PMCMC_JAC_FULL_SIGMOID;
    if (full && tgt->kind != TGT_GAUSS)
        return fail(c, PMCMC_ERR_UNSUPPORTED, "jacobian=:full needs a dense Gaussian target");
    // :diag on a GEMM target: the products of one probe set, then c and
    // diag H_x in its W / hw slots (gemm_diag_products)
    const bool gdiag = mode == MODE_GEMM && p->jacobian == PMCMC_JAC_DIAG;
    const int gprobes = gdiag ? 1 : p->probes;
    if (gdiag && D > 256) return fail(c, PMCMC_ERR_UNSUPPORTED, "jacobian=:diag on a GEMM target needs D <= 256");
    if (mode == MODE_GEMM) {
        std::string msg;
        GlmData gdummy{};
        gdummy.N = 1;
        gdummy.Dp = tgt->gauss.Dp;
        const GlmData* gd = tgt->kind == TGT_GAUSS ? &gdummy : &tgt->glm;
        int rc = glm_work_ensure(&w->glmw, gd, c->stream, D, Tn, full ? 0 : gprobes, sizeof(T) == 8, &msg);
        if (rc == PMCMC_OK && gdiag) rc = diag_work_ensure(&w->diagw, D, Tn, sizeof(T), &msg);
        if (rc == PMCMC_OK && full) rc = full_work_ensure(&w->fullw, D, Tn, sizeof(T) == 8, &msg);
        if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    }
    // FP64 recompute of near-threshold decisions for the GEMM targets in FP32,
    // from the first iteration whose predecessor's update is within 100
    // tolerance units, or from iteration 16.  PMCMC_REFINE=0 turns it off,
    // PMCMC_REFINE=all runs it from the first iteration.
    const char* rv = getenv("PMCMC_REFINE");
    const bool refine = sizeof(T) == 4 && mode == MODE_GEMM && !full && tgt->D <= 256 && !(rv && !strcmp(rv, "0"));
    // band: above the FP32 products' log-alpha error, which grows about like
    // N for the logistic target (max 4.6e-5 at N = 10000, D = 100 and 2e-5 at
    // N = 2000, D = 250 over 4096 columns each); 1e-3 for the Gaussian
    const double refine_band = tgt->kind == TGT_LOGISTIC ? std::max(1e-4, 2.5e-8 * (double)tgt->glm.N) : 1e-3;
    const double refine_gate = rv && !strcmp(rv, "all") ? INFINITY : 100.0;
    RefineTarget rtg{};
    if (refine) {
        rtg.D = D;
        if (tgt->kind == TGT_LOGISTIC) {
            rtg.kind = 0;
            rtg.N = tgt->glm.N;
            rtg.X64 = tgt->glm.Xd64;
            rtg.y64 = tgt->glm.y64;
            rtg.inv_s2 = 1.0 / (tgt->prior_scale * tgt->prior_scale);
            rtg.P64 = tgt->glm.P64;
        } else {
            rtg.kind = 1;
            rtg.L64 = tgt->gauss.L64;
            rtg.mu64 = tgt->gauss.mu64;
        }
        std::string msg;
        int rc = refine_work_ensure(&w->refw, rtg, Tn, refine_band, &msg);
        if (rc == PMCMC_OK) rc = refine_reset(&w->refw, c->stream, &msg);
        if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    }
    // decisions of FP32 iterates within 1e-6 of the threshold are held (the
    // tile kernel's hysteresis); a fill kernel, not a memset, starts them unset
    const bool hold = sizeof(T) == 4 && !full;
    if (hold) {
        if (!w->prev_dec) HIPCHK(c, hipMalloc(&w->prev_dec, (size_t)Tn));
        hipLaunchKernelGGL(fill_i8_kernel, dim3(grid_for(Tn)), dim3(256), 0, c->stream, w->prev_dec, (int8_t)-1, Tn);
    }
    if (!full && w->plan.lev[0].nT > w->tile_smax_n) {
        hipFree(w->tile_smax);
        w->tile_smax = nullptr;
        w->tile_smax_n = 0;
        HIPCHK(c, hipMalloc(&w->tile_smax, (size_t)w->plan.lev[0].nT * sizeof(double)));
        w->tile_smax_n = w->plan.lev[0].nT;
    }
    T* bufs[2] = {(T*)w->buf[0], (T*)w->buf[1]};
    // GEMM targets keep the final iteration's decisions and x0 for the
    // trajectory assembly (pmcmc_deer_last_trajectory)
    w->last_gemm = false;
    if (mode == MODE_GEMM) {
        if (!acc_out) {
            if (!w->acc_ws) HIPCHK(c, hipMalloc(&w->acc_ws, (size_t)Tn));
            acc_out = w->acc_ws;
        }
        if (!w->x0_ws) HIPCHK(c, hipMalloc(&w->x0_ws, (size_t)D * sizeof(double)));
        HIPCHK(c, hipMemcpyAsync(w->x0_ws, x0, (size_t)D * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    }
    hipLaunchKernelGGL(reset_sync_kernel, dim3(1), dim3(64), 0, c->stream, w->sync);
    if (full) full_geometry_reset(&w->fullw);
    hipLaunchKernelGGL((init_iterate_kernel<T>), dim3(grid_for(D * Tn)), dim3(256), 0, c->stream,
                       bufs[0], (const T*)x0, (const T*)init, D, D * Tn);
    HIPCHK(c, hipGetLastError());

    TileArgs<T> a{};
    a.s0 = (const T*)x0;
    a.D = D;
    a.T_len = Tn;
    a.sync = w->sync;
    a.Xi = (const T*)Xi;
    a.U = (const T*)U;
    a.acc_out = acc_out;
    a.la_out = nullptr;
    a.eps = p->eps;
    a.damping = p->damping;
    a.tol_abs = p->tol_abs;
    a.tol_rel = p->tol_rel;
    a.probes = gprobes;
    a.jacobian = p->jacobian;
    a.seed = p->seed;
    a.prev_dec = hold ? w->prev_dec : nullptr;
    a.hyst = hold ? 1e-6 : 0.0;
    if (tgt->kind == TGT_DIAG) {
        a.prec = sizeof(T) == 8 ? (const T*)tgt->prec_d : (const T*)tgt->prec_f;
        a.mean = sizeof(T) == 8 ? (const T*)tgt->mean_d : (const T*)tgt->mean_f;
    }

    int k = 0, chunk = 1;
    bool done = false;
    while (k < p->maxiter && !done) {
        const int n = (p->maxiter - k) < chunk ? (p->maxiter - k) : chunk;
        for (int i = 0; i < n; ++i) {
            const int it = k + i;
            a.iter = it;
            a.probe_iter = p->fixed_probes ? 0xFFFFFFFFu : (uint32_t)it;
            a.S_in = bufs[it & 1];
            a.S = bufs[(it + 1) & 1];
            const int eit = c->prof.begin(KT_ITERATION, it, c->stream);
            if (full) {
                std::string msg;
                // columns below sync->skip_cols are final (see full_newton_step)
                int rc = gauss_newton_products(&w->glmw, &tgt->gauss, c->stream, w->sync, p->precision, p->eps,
                                               (const T*)x0, a.S_in, (const T*)Xi, p->seed, it, a.probe_iter, 0,
                                               D, Tn, &c->prof, &msg, &w->sync->skip_cols);
                if (rc == PMCMC_OK)
                    rc = full_newton_step<T>(&w->fullw, &w->glmw, &tgt->gauss, c->stream, w->sync, p->precision,
                                             p->jacobian == PMCMC_JAC_FULL_SIGMOID, p->eps, p->damping, p->tol_abs, p->tol_rel, (const T*)x0, a.S_in, a.S,
                                             (const T*)U, acc_out, it, &c->prof, &msg);
                if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
            } else if (mode == MODE_GEMM) {
                std::string msg;
                int rc = tgt->kind == TGT_GAUSS
                             ? gauss_newton_products(&w->glmw, &tgt->gauss, c->stream, w->sync, p->precision,
                                                     p->eps, (const T*)x0, a.S_in, (const T*)Xi, p->seed, it,
                                                     a.probe_iter, gprobes, D, Tn, &c->prof, &msg,
                                                     &w->sync->skip_cols)
                             : glm_newton_products(&w->glmw, &tgt->glm, c->stream, w->sync, p->precision,
                                                   p->eps, tgt->prior_scale, (const T*)x0, a.S_in,
                                                   (const T*)Xi, p->seed, it, a.probe_iter, gprobes, D, Tn,
                                                   &c->prof, &msg);
                if (rc == PMCMC_OK && gdiag)
                    rc = gemm_diag_products<T>(&w->diagw, &w->glmw, tgt->kind == TGT_LOGISTIC, &tgt->glm,
                                               &tgt->gauss, tgt->prior_scale, c->stream, p->eps, (const T*)x0,
                                               a.S_in, D, Tn, &w->sync->converged, &c->prof, it,
                                               p->precision == PMCMC_PREC_BF16_JAC, &msg);
                if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
                a.Y = (const T*)w->glmw.Y;
                a.gx = (const T*)w->glmw.gx;
                a.gy = (const T*)w->glmw.gy;
                a.W = (const T*)w->glmw.W;
                a.hw = (const T*)w->glmw.hw;
                a.lpx = w->glmw.lpx;
                a.lpy = w->glmw.lpy;
                a.dec = nullptr;
                a.refine_on = nullptr;
                if (refine) {
                    RefineArgs ra{};
                    ra.x0 = (const float*)x0;
                    ra.S_in = (const float*)a.S_in;
                    ra.Xi = (const float*)Xi;
                    ra.U = (const float*)U;
                    ra.Y = (const float*)a.Y;
                    ra.gx = (const float*)a.gx;
                    ra.gy = (const float*)a.gy;
                    ra.lpx = a.lpx;
                    ra.lpy = a.lpy;
                    ra.eps = p->eps;
                    ra.band = refine_band;
                    ra.skip = &w->sync->converged;
                    ra.last_metric = &w->sync->last_metric;
                    ra.gate = refine_gate;
                    ra.iter = it;
                    ra.gate_iters = 16;
                    ra.hyst = a.hyst;
                    ra.final_cols = &w->sync->skip_cols;
                    rc = refine_decisions(&w->refw, rtg, ra, it, c->stream, &msg);
                    if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
                    a.dec = w->refw.dec;
                    a.refine_on = w->refw.cnt + 2;
                }
                launch_deer_iteration<T, MODE_GEMM>(w, a, c->stream, &c->prof);
            } else {
                launch_deer_iteration<T, MODE_ELEM>(w, a, c->stream, &c->prof);
            }
            c->prof.end(eit, c->stream);
            HIPCHK(c, hipGetLastError());
        }
        k += n;
        HIPCHK(c, hipMemcpyAsync(c->host_sync, w->sync, sizeof(SyncState), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        // a bounded device-side wait that gave up leaves stale end states in
        // the iterate: stop at once rather than iterate on them
        if (c->host_sync->err)
            return fail(c, PMCMC_ERR_DEVICE, "DEER solve: a device-side wait hit its spin limit (iteration %d)",
                        c->host_sync->last_iter);
        done = c->host_sync->converged != 0;
        if (chunk < 16) chunk <<= 1;
        // full Jacobian: tile only the columns that are not final yet
        if (full && !done) full_retile(&w->fullw, w->sync, c->host_sync->skip_cols, c->stream);
    }
    const SyncState hs = *c->host_sync;
    if (hs.err) return fail(c, PMCMC_ERR_DEVICE, "DEER solve: a device-side wait hit its spin limit");
    const int iters = hs.converged ? hs.conv_iter + 1 : (hs.last_iter + 1);
    T* result = bufs[iters & 1];
    if ((void*)result != S_out)
        HIPCHK(c, hipMemcpyAsync(S_out, result, (size_t)D * Tn * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->prof.flush(iters - 1);
    if (mode == MODE_GEMM) {
        w->last_gemm = true;
        w->last_prev = bufs[(iters - 1) & 1];
        w->last_acc = acc_out;
    }
    if (info) {
        info->converged = hs.converged;
        info->iters = iters;
        info->metric = hs.last_metric;
    }
    return PMCMC_OK;
}

// FT_t of the final Newton iteration (Y_t where step t accepted, else the
// state before t) and its log-density (lpy or lpx of that iteration).
template <typename T>
__global__ __launch_bounds__(256) void last_trajectory_kernel(const T* x0, const T* S_prev, const T* Y,
                                                              const uint8_t* acc, const double* lpx,
                                                              const double* lpy, int64_t D, int64_t n, T* S_out,
                                                              double* lp_out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = i / D, d = i - t * D;
        const bool a = acc[t] != 0;
        S_out[i] = a ? Y[i] : (t == 0 ? x0[d] : S_prev[i - D]);
        if (d == 0 && lp_out) lp_out[t] = a ? lpy[t] : lpx[t];
    }
}

extern "C" int pmcmc_deer_solve(pmcmc_ctx* c, pmcmc_deer_ws* w, const pmcmc_target* tgt,
                                const pmcmc_deer_params* p, const void* x0, const void* Xi,
                                const void* U, const void* init, void* S_out, uint8_t* acc_out,
                                pmcmc_deer_info* info) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    if (!w || !tgt || !p) return fail(c, PMCMC_ERR_ARGUMENT, "null workspace/target/params");
    // argument validation of src/DEER/DEER.jl:525-528 and :391-392
    if (p->maxiter < 1) return fail(c, PMCMC_ERR_ARGUMENT, "maxiter must be >= 1");
    if (!(p->tol_abs >= 0)) return fail(c, PMCMC_ERR_ARGUMENT, "tol_abs must be >= 0");
    if (!(p->tol_rel >= 0)) return fail(c, PMCMC_ERR_ARGUMENT, "tol_rel must be >= 0");
    if (!(p->damping > 0 && p->damping <= 1)) return fail(c, PMCMC_ERR_ARGUMENT, "damping must be in (0,1]");
    if (p->probes < 1) return fail(c, PMCMC_ERR_ARGUMENT, "probes must be >= 1");
    if (p->probes > PMCMC_MAXP) return fail(c, PMCMC_ERR_UNSUPPORTED, "probes must be <= %d", PMCMC_MAXP);
    if (p->jacobian < PMCMC_JAC_STOCH_DIAG || p->jacobian > PMCMC_JAC_FULL_SIGMOID)
        return fail(c, PMCMC_ERR_ARGUMENT, "jacobian must be :diag, :stoch_diag, :full or :full_sigmoid");
    if (!(p->eps > 0)) return fail(c, PMCMC_ERR_ARGUMENT, "epsilon must be > 0");
    if (p->precision != w->prec) return fail(c, PMCMC_ERR_ARGUMENT, "params precision %d does not match the workspace (%d)", p->precision, w->prec);
    if (tgt->D != w->D) return fail(c, PMCMC_ERR_DIMENSION, "target dim %lld != workspace D %lld", (long long)tgt->D, (long long)w->D);
    if (!x0 || !Xi || !U || !S_out) return fail(c, PMCMC_ERR_ARGUMENT, "null buffer");
    HIPCHK(c, hipSetDevice(c->device));
    if (w->prec == PMCMC_PREC_F64)
        return deer_solve_t<double>(c, w, tgt, p, x0, Xi, U, init, S_out, acc_out, info);
    return deer_solve_t<float>(c, w, tgt, p, x0, Xi, U, init, S_out, acc_out, info);
}

extern "C" int pmcmc_tape_generate(pmcmc_ctx* c, uint64_t seed, int precision, void* Xi, void* U, int64_t D,
                                   int64_t T) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    if (!Xi || !U) return fail(c, PMCMC_ERR_ARGUMENT, "tape: null buffer");
    if (D < 1 || T < 1) return fail(c, PMCMC_ERR_DIMENSION, "tape: D and T must be >= 1");
    if (precision != PMCMC_PREC_F64 && precision != PMCMC_PREC_F32 && precision != PMCMC_PREC_BF16_JAC)
        return fail(c, PMCMC_ERR_ARGUMENT, "tape: bad precision %d", precision);
    HIPCHK(c, hipSetDevice(c->device));
    if (tape_generate(c->stream, seed, precision == PMCMC_PREC_F64, Xi, U, D, T) != 0)
        return fail(c, PMCMC_ERR_HIP, "tape: kernel launch failed");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PMCMC_OK;
}

extern "C" int pmcmc_deer_last_trajectory(pmcmc_deer_ws* w, void* S_out, double* logp_out) {
    if (!w || !w->ctx) return PMCMC_ERR_ARGUMENT;
    pmcmc_ctx* c = w->ctx;
    if (!w->last_gemm || !w->last_prev || !w->last_acc)
        return fail(c, PMCMC_ERR_ARGUMENT, "last trajectory: the workspace's last solve was not a GEMM-target solve");
    if (!S_out) return fail(c, PMCMC_ERR_ARGUMENT, "last trajectory: null S_out");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = w->D * w->T;
    if (w->prec == PMCMC_PREC_F64)
        hipLaunchKernelGGL((last_trajectory_kernel<double>), dim3(grid_for(n)), dim3(256), 0, c->stream,
                           (const double*)w->x0_ws, (const double*)w->last_prev, (const double*)w->glmw.Y,
                           w->last_acc, w->glmw.lpx, w->glmw.lpy, w->D, n, (double*)S_out, logp_out);
    else
        hipLaunchKernelGGL((last_trajectory_kernel<float>), dim3(grid_for(n)), dim3(256), 0, c->stream,
                           (const float*)w->x0_ws, (const float*)w->last_prev, (const float*)w->glmw.Y,
                           w->last_acc, w->glmw.lpx, w->glmw.lpy, w->D, n, (float*)S_out, logp_out);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PMCMC_OK;
}

// ---------------------------------------------------------------------------
// autotuner: measure candidate level-0 tile shapes, keep the fastest
// ---------------------------------------------------------------------------
template <typename T>
__global__ void fill_value_kernel(T* p, T v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// the profiler's accumulated times are the caller's: tuning runs on a copy
struct ProfSaved {
    bool on;
    double ms[KT_COUNT];
    long long n[KT_COUNT];
    explicit ProfSaved(Prof& p) : on(p.on) {
        for (int k = 0; k < KT_COUNT; ++k) {
            ms[k] = p.ms[k];
            n[k] = p.n[k];
        }
    }
    void restore(Prof& p) const {
        p.on = on;
        for (int k = 0; k < KT_COUNT; ++k) {
            p.ms[k] = ms[k];
            p.n[k] = n[k];
        }
    }
};

// Level-0 candidates around the rule's shape: TC x {1/4 .. 8}, P from the
// most parts that fit down to 1; for the standalone scan also the single-pass
// look-back form at its widest tile and two narrower ones.
template <typename T>
static std::vector<TileChoice> tile_candidates(int64_t D, int64_t Tn, bool with_iterate, bool chained_too) {
    std::vector<TileChoice> out;
    int tc0, p0, nt0;
    tile_shape<T>(D, Tn, with_iterate, &tc0, &p0, &nt0);
    const int pmax = D >= PMCMC_NT ? 1 : (int)(PMCMC_NT / D);
    const int mults[6][2] = {{1, 4}, {1, 2}, {1, 1}, {2, 1}, {4, 1}, {8, 1}};
    auto add = [&](TileChoice c) {
        for (const auto& o : out)
            if (o.TC == c.TC && o.P == c.P && o.chained == c.chained) return;
        if (shape_fits<T>(D, with_iterate, c.chained != 0, c)) out.push_back(c);
    };
    for (const auto& m : mults) {
        const int64_t tc = (int64_t)tc0 * m[0] / m[1];
        if (tc < 1 || tc > Tn) continue;
        for (int pp = pmax; pp >= 1; pp = pp > 1 ? pp / 2 : 0) {
            TileChoice c;
            c.TC = (int)tc;
            c.P = std::min<int64_t>(pp, tc);
            add(c);
            if (pp == 1) break;
        }
    }
    if (chained_too && reg_blocks(D)) {
        for (int form : {2, 3})
            for (int tc : {64, 128, 256}) {
                TileChoice c;
                c.TC = tc;
                c.P = 1;
                c.chained = form;
                if (Tn > tc && reg_tile_fits<T>(D, c)) add(c);
            }
    }
    if (chained_too) {
        int tc = 1;
        while (tc * 2 <= Tn && tile_lds_bytes<T>(D, tc * 2, false, true) <= 64512) tc *= 2;
        for (int div = 1; div <= 4; div *= 2) {
            TileChoice c;
            c.TC = std::max(1, tc / div);
            c.P = std::min(pmax, c.TC);
            c.chained = 1;
            if ((Tn + c.TC - 1) / c.TC > 1) add(c);
        }
    }
    return out;
}

template <typename T>
static int scan_tune_t(pmcmc_ctx* c, int64_t D, int64_t Tn, int reps, TileChoice* best) {
    const size_t n = (size_t)D * Tn;
    T *A = nullptr, *B = nullptr, *s0 = nullptr, *S = nullptr;
    bool ok = hipMalloc(&A, n * sizeof(T)) == hipSuccess && hipMalloc(&B, n * sizeof(T)) == hipSuccess &&
              hipMalloc(&s0, (size_t)D * sizeof(T)) == hipSuccess && hipMalloc(&S, n * sizeof(T)) == hipSuccess;
    auto release = [&]() {
        hipFree(A);
        hipFree(B);
        hipFree(s0);
        hipFree(S);
    };
    if (!ok) {
        release();
        return fail(c, PMCMC_ERR_ALLOC, "scan tune: cannot allocate the synthetic operands");
    }
    hipLaunchKernelGGL(fill_value_kernel<T>, dim3(grid_for((int64_t)n)), dim3(256), 0, c->stream, A, (T)0.999,
                       (int64_t)n);
    hipLaunchKernelGGL(fill_value_kernel<T>, dim3(grid_for((int64_t)n)), dim3(256), 0, c->stream, B, (T)1e-3,
                       (int64_t)n);
    hipLaunchKernelGGL(fill_value_kernel<T>, dim3(grid_for(D)), dim3(256), 0, c->stream, s0, (T)0, D);
    const ProfSaved saved(c->prof);
    best->TC = 0;
    best->us = 1e300;
    int rc = PMCMC_OK;
    for (const TileChoice& cand : tile_candidates<T>(D, Tn, false, true)) {
        c->prof.on = false;
        rc = scan_diag<T>(c, A, B, s0, S, D, Tn, &cand);  // warm-up (and allocation)
        if (rc != PMCMC_OK) break;
        c->prof.reset();
        c->prof.on = true;
        for (int r = 0; r < reps && rc == PMCMC_OK; ++r) rc = scan_diag<T>(c, A, B, s0, S, D, Tn, &cand);
        if (rc != PMCMC_OK) break;
        const double us = c->prof.n[KT_ITERATION] ? 1e3 * c->prof.ms[KT_ITERATION] / c->prof.n[KT_ITERATION] : 1e300;
        if (us < best->us) {
            *best = cand;
            best->us = us;
        }
    }
    saved.restore(c->prof);
    release();
    if (rc != PMCMC_OK) return rc;
    if (best->TC > 0) tune_put(PMCMC_TUNE_SCAN_DIAG, (int)sizeof(T), D, Tn, *best);
    return PMCMC_OK;
}

static void shape_out(const TileChoice& c, pmcmc_tile_shape* out) {
    if (!out) return;
    out->tc = c.TC;
    out->parts = c.P;
    out->single_pass = c.chained;
    out->pad = 0;
    out->us = c.us;
}

extern "C" int pmcmc_scan_tune(pmcmc_ctx* c, int prec, int64_t D, int64_t T, int32_t reps, pmcmc_tile_shape* best) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    if (prec != PMCMC_PREC_F32 && prec != PMCMC_PREC_F64)
        return fail(c, PMCMC_ERR_ARGUMENT, "scan tune: precision must be F32 or F64");
    if (D < 1 || T < 1) return fail(c, PMCMC_ERR_DIMENSION, "scan tune: need D >= 1 and T >= 1");
    if (reps < 1) return fail(c, PMCMC_ERR_ARGUMENT, "scan tune: reps must be >= 1");
    HIPCHK(c, hipSetDevice(c->device));
    TileChoice b;
    const int rc = prec == PMCMC_PREC_F64 ? scan_tune_t<double>(c, D, T, reps, &b) : scan_tune_t<float>(c, D, T, reps, &b);
    if (rc != PMCMC_OK) return rc;
    shape_out(b, best);
    return PMCMC_OK;
}

// re-plan a workspace's level-0 tile (the blob grows when the plan needs more)
template <typename T>
static int ws_replan(pmcmc_ctx* c, pmcmc_deer_ws* w, const TileChoice* ch) {
    ScanPlan plan;
    if (!plan_scan<T>(&plan, w->D, w->T, true, ch))
        return fail(c, PMCMC_ERR_UNSUPPORTED, "tile shape does not fit the fused tile kernel");
    if (plan.bytes > w->blob_bytes) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        void* nb = nullptr;
        HIPCHK(c, hipMalloc(&nb, plan.bytes));
        hipFree(w->blob);
        w->blob = nb;
        w->blob_bytes = plan.bytes;
    }
    w->plan = plan;
    return PMCMC_OK;
}

extern "C" int pmcmc_deer_tune(pmcmc_ctx* c, pmcmc_deer_ws* w, const pmcmc_target* tgt, const pmcmc_deer_params* p,
                               const void* x0, const void* Xi, const void* U, int32_t iters,
                               pmcmc_tile_shape* best) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    if (!w || !tgt || !p || !x0 || !Xi || !U) return fail(c, PMCMC_ERR_ARGUMENT, "deer tune: null argument");
    if (iters < 1) return fail(c, PMCMC_ERR_ARGUMENT, "deer tune: iters must be >= 1");
    if (p->jacobian == PMCMC_JAC_FULL || p->jacobian == PMCMC_JAC_FULL_SIGMOID)
        return fail(c, PMCMC_ERR_UNSUPPORTED, "deer tune: the dense-Jacobian solve has no diagonal tile to tune");
    HIPCHK(c, hipSetDevice(c->device));
    const bool f64 = w->prec == PMCMC_PREC_F64;
    void* S = nullptr;
    HIPCHK(c, hipMalloc(&S, (size_t)w->D * w->T * w->esz));
    pmcmc_deer_params q = *p;
    q.tol_abs = 0.0;  // never converges: every candidate runs `iters` iterations
    q.tol_rel = 0.0;
    q.maxiter = iters;
    const ProfSaved saved(c->prof);
    TileChoice b;
    b.us = 1e300;
    const ScanPlan orig = w->plan;
    const std::vector<TileChoice> cands =
        f64 ? tile_candidates<double>(w->D, w->T, true, false) : tile_candidates<float>(w->D, w->T, true, false);
    int rc = PMCMC_OK;
    for (const TileChoice& cand : cands) {
        rc = f64 ? ws_replan<double>(c, w, &cand) : ws_replan<float>(c, w, &cand);
        if (rc != PMCMC_OK) break;
        c->prof.on = true;
        c->prof.reset();
        rc = pmcmc_deer_solve(c, w, tgt, &q, x0, Xi, U, nullptr, S, nullptr, nullptr);
        if (rc != PMCMC_OK) break;
        // the shape-dependent kernels: level-0 reduce and apply, the spine
        const long long nit = c->prof.n[KT_TILE_APPLY];
        const double ms = c->prof.ms[KT_TILE_REDUCE] + c->prof.ms[KT_TILE_APPLY] + c->prof.ms[KT_SCAN_SPINE];
        const double us = nit ? 1e3 * ms / nit : 1e300;
        if (us < b.us) {
            b = cand;
            b.us = us;
        }
    }
    saved.restore(c->prof);
    hipFree(S);
    if (rc != PMCMC_OK) {
        w->plan = orig;
        return rc;
    }
    rc = f64 ? ws_replan<double>(c, w, &b) : ws_replan<float>(c, w, &b);
    if (rc != PMCMC_OK) return rc;
    tune_put(PMCMC_TUNE_DEER, (int)w->esz, w->D, w->T, b);
    shape_out(b, best);
    return PMCMC_OK;
}

extern "C" int pmcmc_tune_get(int kind, int prec, int64_t D, int64_t T, pmcmc_tile_shape* out) {
    if (!out || (kind != PMCMC_TUNE_SCAN_DIAG && kind != PMCMC_TUNE_DEER)) return PMCMC_ERR_ARGUMENT;
    TileChoice ch;
    if (!tune_get(kind, prec == PMCMC_PREC_F64 ? 8 : 4, D, T, &ch)) ch = TileChoice{};
    shape_out(ch, out);
    return PMCMC_OK;
}

extern "C" int pmcmc_tune_set(int kind, int prec, int64_t D, int64_t T, const pmcmc_tile_shape* shape) {
    if (!shape || (kind != PMCMC_TUNE_SCAN_DIAG && kind != PMCMC_TUNE_DEER) || D < 1 || T < 1)
        return PMCMC_ERR_ARGUMENT;
    TileChoice ch;
    ch.TC = shape->tc;
    ch.P = shape->parts;
    ch.chained = kind == PMCMC_TUNE_SCAN_DIAG ? std::min(std::max(shape->single_pass, 0), 3) : 0;
    ch.us = shape->us;
    const bool with_iterate = kind == PMCMC_TUNE_DEER;
    const bool fits = prec == PMCMC_PREC_F64 ? shape_fits<double>(D, with_iterate, ch.chained != 0, ch)
                                             : shape_fits<float>(D, with_iterate, ch.chained != 0, ch);
    if (!fits) return PMCMC_ERR_UNSUPPORTED;
    tune_put(kind, prec == PMCMC_PREC_F64 ? 8 : 4, D, T, ch);
    return PMCMC_OK;
}

extern "C" int pmcmc_tune_clear(void) {
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    t.m.clear();
    return PMCMC_OK;
}

extern "C" int64_t pmcmc_tune_count(void) {
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    return (int64_t)t.m.size();
}

extern "C" int pmcmc_tune_entry(int64_t i, int* kind, int* prec, int64_t* D, int32_t* log2T, pmcmc_tile_shape* out) {
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    if (i < 0 || i >= (int64_t)t.m.size()) return PMCMC_ERR_ARGUMENT;
    auto it = t.m.begin();
    std::advance(it, i);
    if (kind) *kind = std::get<0>(it->first);
    if (prec) *prec = std::get<1>(it->first) == 8 ? PMCMC_PREC_F64 : PMCMC_PREC_F32;
    if (D) *D = std::get<2>(it->first);
    if (log2T) *log2T = std::get<3>(it->first);
    shape_out(it->second, out);
    return PMCMC_OK;
}

// ---------------------------------------------------------------------------
// adjoint of a converged trajectory (decisions held fixed)
// ---------------------------------------------------------------------------
template <typename T>
static int adj_products(pmcmc_ctx* c, pmcmc_deer_ws* w, const pmcmc_target* tgt, double eps, const T* V,
                        const T* x0, const T* Z, uint32_t key, const T* Xi, const int* skip, int iter,
                        std::string* msg) {
    const int64_t D = w->D, Tn = w->T;
    if (tgt->kind == TGT_DIAG) {
        const T* prec = sizeof(T) == 8 ? (const T*)tgt->prec_d : (const T*)tgt->prec_f;
        const T* mean = sizeof(T) == 8 ? (const T*)tgt->mean_d : (const T*)tgt->mean_f;
        hipLaunchKernelGGL((adj_diag_products_kernel<T>), dim3(grid_for(D * Tn)), dim3(256), 0, c->stream, V, x0, Z,
                           key, prec, mean, eps, D, Tn, (T*)w->glmw.gx, (T*)w->glmw.W, skip);
        if (hipGetLastError() != hipSuccess) {
            *msg = "adjoint products kernel failed";
            return PMCMC_ERR_HIP;
        }
        return PMCMC_OK;
    }
    if (tgt->kind == TGT_GAUSS)
        return gauss_dir_products<T>(&w->glmw, &tgt->gauss, c->stream, eps, V, x0, Z, key, Xi, D, Tn, skip,
                                     &c->prof, iter, msg);
    return glm_dir_products<T>(&w->glmw, &tgt->glm, c->stream, eps, tgt->prior_scale, V, x0, Z, key, Xi, D, Tn,
                               skip, &c->prof, iter, msg);
}

template <typename T>
static int deer_adjoint_t(pmcmc_ctx* c, pmcmc_deer_ws* w, const pmcmc_target* tgt, const pmcmc_deer_params* p,
                          const T* x0, const T* Xi, const T* S, const uint8_t* acc, const T* G, double* grad_out,
                          T* dx0_out, T* lam_out, pmcmc_deer_info* info) {
    const int64_t D = w->D, Tn = w->T;
    std::string msg;
    GlmData gdummy{};
    gdummy.N = 1;
    gdummy.Dp = tgt->kind == TGT_GAUSS ? tgt->gauss.Dp : (D + 15) / 16 * 16;
    const GlmData* gd = tgt->kind == TGT_LOGISTIC ? &tgt->glm : &gdummy;
    int rc = glm_work_ensure(&w->glmw, gd, c->stream, D, Tn, 1, sizeof(T) == 8, &msg);
    if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    if (!w->adj_lam) {
        const size_t mat = (size_t)D * Tn * sizeof(T);
        if (!plan_scan<T>(&w->adj_plan, D, Tn, false))
            return fail(c, PMCMC_ERR_UNSUPPORTED, "adjoint: D=%lld too large for the scan tile", (long long)D);
        bool ok = hipMalloc(&w->adj_lam, mat + D * sizeof(T)) == hipSuccess &&
                  hipMalloc(&w->adj_A, mat) == hipSuccess && hipMalloc(&w->adj_B, mat) == hipSuccess &&
                  hipMalloc(&w->adj_D, mat) == hipSuccess && hipMalloc(&w->adj_zero, D * sizeof(T)) == hipSuccess &&
                  hipMemset(w->adj_zero, 0, D * sizeof(T)) == hipSuccess &&
                  hipMalloc(&w->adj_part, (2 * Tn + 2) * sizeof(double)) == hipSuccess &&
                  hipMalloc(&w->adj_blob, w->adj_plan.bytes > 0 ? w->adj_plan.bytes : 16) == hipSuccess;
        if (!ok) return fail(c, PMCMC_ERR_ALLOC, "adjoint workspace allocation failed");
    }
    T* lam = (T*)w->adj_lam;
    T* Ar = (T*)w->adj_A;
    T* Br = (T*)w->adj_B;
    T* Dr = (T*)w->adj_D;
    const int* skip = &w->sync->converged;
    const uint32_t key = probe_key(p->seed, 0x41444A31u, 0u);
    const bool exact = tgt->kind == TGT_DIAG;
    const T* prec = sizeof(T) == 8 ? (const T*)tgt->prec_d : (const T*)tgt->prec_f;

    hipLaunchKernelGGL(reset_sync_kernel, dim3(1), dim3(64), 0, c->stream, w->sync);
    if (!exact) {
        rc = adj_products<T>(c, w, tgt, p->eps, S, nullptr, nullptr, key, Xi, nullptr, 0, &msg);
        if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    }
    hipLaunchKernelGGL((adj_coef_kernel<T>), dim3(grid_for(D * Tn + D)), dim3(256), 0, c->stream,
                       (const T*)w->glmw.W, key, exact ? prec : nullptr, exact, p->eps, acc, D, Tn, Ar, lam);
    HIPCHK(c, hipGetLastError());
    const int ublocks = grid_for(D * Tn, 256, 2048);
    int k = 0, chunk = 1;
    bool done = false;
    while (k < p->maxiter && !done) {
        const int n = (p->maxiter - k) < chunk ? (p->maxiter - k) : chunk;
        for (int i = 0; i < n; ++i) {
            const int it = k + i;
            const int e = c->prof.begin(KT_ITERATION, it, c->stream);
            rc = adj_products<T>(c, w, tgt, p->eps, S, nullptr, lam + D, 0u, Xi, skip, it, &msg);
            if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
            hipLaunchKernelGGL((adj_resid_kernel<T>), dim3(grid_for(D * Tn)), dim3(256), 0, c->stream, G,
                               (const T*)w->glmw.W, (const T*)lam, acc, D, Tn, Br, skip);
            run_scan_levels<T>(w->adj_plan, 0, Ar, Br, Dr, (const T*)w->adj_zero, (char*)w->adj_blob, skip,
                               c->stream);
            hipLaunchKernelGGL((adj_update_kernel<T>), dim3(ublocks), dim3(256), 0, c->stream, (const T*)Dr, lam, D,
                               Tn, w->sync, it, p->tol_abs, p->tol_rel);
            c->prof.end(e, c->stream);
            HIPCHK(c, hipGetLastError());
        }
        k += n;
        HIPCHK(c, hipMemcpyAsync(c->host_sync, w->sync, sizeof(SyncState), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        done = c->host_sync->converged != 0;
        if (chunk < 16) chunk <<= 1;
    }
    const SyncState hs = *c->host_sync;
    const int iters = hs.converged ? hs.conv_iter + 1 : (hs.last_iter + 1);
    c->prof.flush(iters - 1);
    // gradient at Xbar: grad and J^T lambda (for dx0), then the column sums
    rc = adj_products<T>(c, w, tgt, p->eps, S, x0, (const T*)lam, 0u, Xi, nullptr, iters, &msg);
    if (rc != PMCMC_OK) return fail(c, rc, "%s", msg.c_str());
    hipLaunchKernelGGL((adj_grad_kernel<T>), dim3((unsigned)((Tn + 3) / 4)), dim3(256), 0, c->stream, S, x0, Xi,
                       (const T*)w->glmw.gx, (const T*)w->glmw.W, (const T*)lam, acc, p->eps,
                       tgt->kind == TGT_LOGISTIC ? 1 : 0, tgt->prior_scale, D, Tn, w->adj_part, dx0_out);
    hipLaunchKernelGGL(adj_sum_kernel, dim3(1), dim3(256), 0, c->stream, (const double*)w->adj_part, Tn,
                       w->adj_part + 2 * Tn);
    HIPCHK(c, hipGetLastError());
    if (lam_out)
        HIPCHK(c, hipMemcpyAsync(lam_out, lam, (size_t)D * Tn * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    double host[2];
    HIPCHK(c, hipMemcpyAsync(host, w->adj_part + 2 * Tn, sizeof host, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    grad_out[0] = host[0];
    grad_out[1] = host[1];
    if (info) {
        info->converged = hs.converged;
        info->iters = iters;
        info->metric = hs.last_metric;
    }
    return PMCMC_OK;
}

extern "C" int pmcmc_deer_adjoint(pmcmc_ctx* c, pmcmc_deer_ws* w, const pmcmc_target* tgt,
                                  const pmcmc_deer_params* p, const void* x0, const void* Xi, const void* S,
                                  const uint8_t* accepted, const void* G, double* grad_out, void* dx0_out,
                                  void* lam_out, pmcmc_deer_info* info) {
    if (!c) return PMCMC_ERR_ARGUMENT;
    if (!w || !tgt || !p) return fail(c, PMCMC_ERR_ARGUMENT, "null workspace/target/params");
    if (p->maxiter < 1) return fail(c, PMCMC_ERR_ARGUMENT, "maxiter must be >= 1");
    if (!(p->tol_abs >= 0)) return fail(c, PMCMC_ERR_ARGUMENT, "tol_abs must be >= 0");
    if (!(p->tol_rel >= 0)) return fail(c, PMCMC_ERR_ARGUMENT, "tol_rel must be >= 0");
    if (!(p->eps > 0)) return fail(c, PMCMC_ERR_ARGUMENT, "epsilon must be > 0");
    if (p->precision != w->prec) return fail(c, PMCMC_ERR_ARGUMENT, "params precision %d does not match the workspace (%d)", p->precision, w->prec);
    if (tgt->D != w->D) return fail(c, PMCMC_ERR_DIMENSION, "target dim %lld != workspace D %lld", (long long)tgt->D, (long long)w->D);
    if (!x0 || !Xi || !S || !accepted || !G || !grad_out) return fail(c, PMCMC_ERR_ARGUMENT, "null buffer");
    HIPCHK(c, hipSetDevice(c->device));
    if (w->prec == PMCMC_PREC_F64)
        return deer_adjoint_t<double>(c, w, tgt, p, (const double*)x0, (const double*)Xi, (const double*)S,
                                      accepted, (const double*)G, grad_out, (double*)dx0_out, (double*)lam_out,
                                      info);
    return deer_adjoint_t<float>(c, w, tgt, p, (const float*)x0, (const float*)Xi, (const float*)S, accepted,
                                 (const float*)G, grad_out, (float*)dx0_out, (float*)lam_out, info);
}

// ---------------------------------------------------------------------------
// batched solve of many chains (uneven D, T, targets) in sliding windows
// ---------------------------------------------------------------------------
// Every kernel of an iteration is one launch over all live chains: the work
// items (chain, tile) / (chain, column block, N split) index per-chain
// descriptor tables in device memory, GEMM passes grouped by padded width.
// Each chain iterates over its own window [front, hi) of its whole
// trajectory (SlideState).  A window kernel at the start of every iteration
// places each chain's window: it rewrites the window-dependent fields of the
// chain's descriptors, gives columns entering the window for the first time
// the state just before them, and copies the columns the last iteration
// accepted into the other iterate buffer.  The last workgroup of the apply
// pass moves the front (tile_kernel.hpp).  No chain waits for another.  The
// host reads the chains' states after every chunk of iterations, picks each
// chain's window length and rebuilds the work items when a length changes or
// chains have finished; the iterations in between are replayed from a
// captured hipGraph (an iteration pair, since the iterate buffers alternate).
namespace pmcmc {

// One chain's window placement: its buffers and the addresses of its entries
// in the device descriptor tables.
struct WinDesc {
    int64_t D, T, L;
    int TC, kind;  // kind: 0 logistic, 1 dense Gaussian
    float* buf[2];
    const float* x0;
    const float* Xi;
    const float* U;
    uint8_t* acc_out;
    int8_t* prev_dec;  // T entries (absolute columns)
    SlideState* win;
    SyncState* sy;
    const int* span;
    TileArgs<float>* ta[2];
    GlmPassArgs* gp[3];
    GlmReduceArgs* gr[3];
    GaussPassArgs* qp[3];
    RefineArgs* ra[2];
    RefineWork* rw;
};

struct SlideInit {
    SlideState* win;
    SyncState* sy;
    int* rcnt;  // refinement counters and gate flag (4), cleared
    int8_t* prev_dec;
    int64_t T;
};

struct BCopy {
    const float* b0;
    const float* b1;
    float* dst;
    int64_t n;
    const SyncState* sy;
};

__global__ __launch_bounds__(256) void batch_slide_init_kernel(const SlideInit* t) {
    const SlideInit& a = t[blockIdx.y];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        SyncState* s = a.sy;
        s->done = 0u;
        s->err = 0u;
        s->dmax_bits = 0ull;
        s->smax_bits = 0ull;
        s->converged = 0;
        s->conv_iter = -1;
        s->last_metric = INFINITY;
        s->last_iter = -1;
        s->live_from = INT_MAX;
        s->skip_cols = 0;
        s->smax_prefix = 0.0;
        SlideState* w = a.win;
        w->front = 0;
        w->hi = 0;
        w->T = a.T;
        w->hw = 0;
        w->pend_lo = w->pend_hi = 0;
        w->seg_start = 0;
        w->accepts = 0;
        w->last_adv_iter = -1;
    }
    if (blockIdx.x == 0 && threadIdx.x < 4) a.rcnt[threadIdx.x] = 0;
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < a.T; i += (int64_t)gridDim.x * 256) a.prev_dec[i] = -1;
}

// Places the windows of iteration parity `par` (its input buffer is buf[par]).
// Reads the chain's SlideState without changing it: every workgroup of the
// chain sees the same state, and the apply pass's last workgroup, which owns
// the updates, runs after this kernel.
__global__ __launch_bounds__(256) void batch_window_kernel(const WinDesc* __restrict__ wds, const int* __restrict__ ids, int par) {
    const WinDesc& w = wds[ids[blockIdx.y]];
    // the flag, the slide state and the span come from earlier launches:
    // plain loads, all in flight together
    const int conv = *gp(&w.sy->converged);
    struct {
        long long front, hw, pend_lo, pend_hi;
    } st;
    st.front = *gp(&w.win->front);
    st.hw = *gp(&w.win->hw);
    st.pend_lo = *gp(&w.win->pend_lo);
    st.pend_hi = *gp(&w.win->pend_hi);
    int64_t span = *gp(w.span);
    if (conv) return;
    const int64_t D = w.D;
    if (span < 1) span = 1;
    if (span > w.L) span = w.L;
    const int64_t front = st.front;
    const int64_t hi = front + span < w.T ? front + span : w.T;
    float* sin = w.buf[par];
    float* sout = w.buf[par ^ 1];
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x, stride = (int64_t)gridDim.x * 256;
    // columns the last iteration accepted: into the other buffer too, so that
    // both hold them (the next iteration reads its state just before the front)
    if (st.pend_hi > st.pend_lo) {
        const int64_t lo = st.pend_lo * D, n = (st.pend_hi - st.pend_lo) * D;
        for (int64_t i = gid; i < n; i += stride) sout[lo + i] = sin[lo + i];
    }
    // columns entering the window for the first time start from the state
    // just before them
    if (hi > st.hw) {
        const float* src = st.hw == 0 ? w.x0 : sin + (st.hw - 1) * D;
        const int64_t lo = st.hw * D;
        const int n = (int)((hi - st.hw) * D), Di = (int)D;  // < 2^31: one window
        for (int i = (int)gid; i < n; i += (int)stride) sin[lo + i] = src[i % Di];
    }
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int64_t n = hi - front;
    const int64_t off = front * D;
    const float* s0 = front == 0 ? w.x0 : sin + off - D;
    TileArgs<float>& a = *w.ta[par];
    a.s0 = s0;
    a.T_len = n;
    a.numTiles = (int)((n + w.TC - 1) / w.TC);
    a.agg_inline = a.numTiles <= PMCMC_AGG_INLINE && D <= PMCMC_NT;
    a.S_in = sin + off;
    a.S = sout + off;
    a.Xi = w.Xi + off;
    a.U = w.U + front;
    a.acc_out = w.acc_out ? w.acc_out + front : nullptr;
    a.prev_dec = w.prev_dec + front;
    a.zoff = off;
    const int vs[2] = {par, 2};
    for (int q = 0; q < 2; ++q) {
        const int v = vs[q];
        ColSrc<float>* src = w.kind == 0 ? &w.gp[v]->src : &w.qp[v]->src;
        src->Tn = n;
        src->zoff = off;
        if (v < 2) {
            src->V = sin + off;
            src->x0 = s0;
        }
        if (w.kind == 0) {
            w.gr[v]->src = *src;
            w.gr[v]->Xi = w.Xi + off;
            w.gr[v]->llstride = n;  // the pass lays its log-likelihood partials out per split over Tn columns
        } else {
            w.qp[v]->Xi = w.Xi + off;
        }
    }
    RefineArgs& r = *w.ra[par];
    r.x0 = s0;
    r.S_in = sin + off;
    r.Xi = w.Xi + off;
    r.U = w.U + front;
    w.rw->T = n;
    w.win->hi = hi;
}

// the final iterate of a chain that finished at iteration k sits in buffer (k + 1) & 1
__global__ __launch_bounds__(256) void batch_copy_kernel(const BCopy* t) {
    const BCopy& a = t[blockIdx.y];
    const SyncState* s = a.sy;
    const int k = s->converged ? s->conv_iter : s->last_iter;
    const float* src = (k >= 0 && ((k + 1) & 1)) ? a.b1 : a.b0;
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * 256) a.dst[i] = src[i];
}

// Scan of the level-0 tile aggregates of each chain (one workgroup per chain,
// rows over threads): prev[:, k] = Newton correction after tile k, from zero
// (the window's start state is fixed).
__global__ __launch_bounds__(256) void batch_agg_scan_kernel(const TileArgs<float>* __restrict__ args, const int* __restrict__ chains) {
    const TileArgs<float>& a = args[chains[blockIdx.x]];
    if (a.agg_inline) return;  // the apply pass folds this chain's few aggregates itself
    if (a.skip && *gp(a.skip)) return;
    const int64_t D = a.D;
    const int nT = a.numTiles;
    float* prev = const_cast<float*>(a.prev);
    for (int64_t d = threadIdx.x; d < D; d += 256) {
        float s = 0.0f;
        int k = 0;
        // the chain through the tiles is cheap; the loads are not: 16 tiles'
        // aggregates in flight per round trip
        for (; k + 16 <= nT; k += 16) {
            float aa[16], bb[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                aa[q] = a.agg_a[(size_t)(k + q) * D + d];
                bb[q] = a.agg_b[(size_t)(k + q) * D + d];
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                s = aa[q] * s + bb[q];
                prev[(size_t)(k + q) * D + d] = s;
            }
        }
        for (; k < nT; ++k) {
            s = a.agg_a[(size_t)k * D + d] * s + a.agg_b[(size_t)k * D + d];
            prev[(size_t)k * D + d] = s;
        }
    }
}

}  // namespace pmcmc

struct BatchChain {
    const pmcmc_target* tgt = nullptr;
    int64_t D = 0, T = 0, Dp = 0, L = 0;
    double eps = 0;
    const float* x0 = nullptr;
    const float* Xi = nullptr;
    const float* U = nullptr;
    float* S_out = nullptr;
    uint8_t* acc_out = nullptr;
    // device state (carved from the arena)
    float* buf[2] = {nullptr, nullptr};  // D x T each: the whole trajectory
    float *Y = nullptr, *gx = nullptr, *gy = nullptr, *W = nullptr, *hw = nullptr;  // D x L (window)
    double *lpx = nullptr, *lpy = nullptr;
    float* part = nullptr;  // logistic only
    float *agg_a = nullptr, *agg_b = nullptr, *prev = nullptr;
    double* tile_smax = nullptr;
    int8_t* prev_dec = nullptr;  // T
    int ns = 1;  // the most N splits of the GEMM passes (sizes the partial sums)
    int TC = 0, P = 0, nTmax = 0;
    int nslab = 1;  // 64-row slabs of the Lambda pass (dense Gaussian), else 1
    RefineTarget rtg{};
    double band = 0;
    RefineWork rw;
    // window length policy (host): the chain's narrow length, and what the
    // previous synchronisation saw
    int span = 0, narrow = 0;
    int64_t prev_front = 0;
    int prev_iter = -1;
    double prev_metric = INFINITY;
    bool stopped = false;
    // per-solve results
    int iters = 0;
    bool converged = true;
    double metric = 0;
};

struct pmcmc_batch {
    pmcmc_ctx* ctx = nullptr;
    bool ctx_ref = false;  // holds a context reference (set once created)
    std::vector<BatchChain> ch;
    pmcmc_deer_params p{};
    int64_t W = 0;
    bool jbf = false;
    bool use_graph = true;
    bool refine = true;  // FP64 decisions near the threshold (PMCMC_REFINE=0: off)
    void* arena = nullptr;
    SyncState* syncs = nullptr;    // n, device
    SlideState* slides = nullptr;  // n, device
    int* spans = nullptr;          // n, device (host-written)
    SyncState* hsync = nullptr;    // n, pinned host
    SlideState* hslide = nullptr;  // n, pinned host
    int* hspans = nullptr;         // n, pinned host
    void* tables = nullptr;        // device descriptor tables
    size_t tables_cap = 0;
    void* items = nullptr;         // device work-item lists
    size_t items_cap = 0;
    // profiling of the Lambda passes (pmcmc_batch_set_profiling): event pairs
    // around every launch (direct launches, no graph) and the passes'
    // algorithmic flops counted on the device
    bool prof_on = false;
    Prof prof;
    unsigned long long* flop_ctr = nullptr;  // device: [0] Lambda passes, [1] GLM passes
    double prof_flops[2] = {0, 0};
};

namespace {

// Work-item lists over the live chains.  Tile items are split into LDS-size
// classes so that small tiles are not launched with the largest chain's LDS
// footprint (and its occupancy).
static constexpr int NLDS = 3;
static int lds_class(size_t lds) { return lds <= 16384 ? 0 : (lds <= 32768 ? 1 : 2); }
struct Items {
    std::vector<int2> red[NLDS];
    std::vector<int4> glm_red;
    std::vector<int> multi, wins;
    std::vector<int4> glm[17];
    std::vector<int4> gslab;  // (chain, column block, slab) of the Lambda passes
    size_t lds_multi[NLDS] = {0, 0, 0};
    // FP64 decision recompute: live chains (all, logistic, Gaussian)
    std::vector<int> r_all, r_glm, r_gauss;
    int64_t r_maxT = 0;
    int r_split_glm = 0, r_split_gauss = 0;
};

// Host staging of a set of device tables, each 256-byte aligned.
struct Pack {
    std::vector<char> host;
    size_t put(const void* src, size_t bytes) {
        const size_t o = align256(host.size());
        host.resize(o + bytes);
        if (bytes) memcpy(host.data() + o, src, bytes);
        return o;
    }
};

}  // namespace

static void batch_free(pmcmc_batch* b) {
    hipFree(b->arena);
    hipFree(b->syncs);
    hipFree(b->slides);
    hipFree(b->spans);
    hipHostFree(b->hsync);
    hipHostFree(b->hslide);
    hipHostFree(b->hspans);
    hipFree(b->tables);
    hipFree(b->items);
    hipFree(b->flop_ctr);
    b->flop_ctr = nullptr;
    b->items = nullptr;
    b->arena = nullptr;
    b->syncs = nullptr;
    b->slides = nullptr;
    b->spans = nullptr;
    b->hsync = nullptr;
    b->hslide = nullptr;
    b->hspans = nullptr;
    b->tables = nullptr;
}

extern "C" int pmcmc_batch_destroy(pmcmc_batch* b) {
    if (!b) return PMCMC_OK;
    if (b->ctx) {
        hipSetDevice(b->ctx->device);
        hipStreamSynchronize(b->ctx->stream);
    }
    batch_free(b);
    pmcmc_ctx* c = b->ctx_ref ? b->ctx : nullptr;
    delete b;
    if (c) ctx_release(c);
    return PMCMC_OK;
}

// Window length of a chain whose front only moves about one column per
// iteration (the causal progress of a Newton iteration whose Jacobian
// estimate misses the dynamics): at least two tiles and 32 columns past the
// front, since the columns further ahead cannot settle before the front
// reaches them.  On 256 uneven chains (128 of them causal), with 4096-element
// tiles a 32-column floor measured 8.08 s per solve, against 8.35 s at 64,
// 10.6 s at 128 and 12.5 s at 256: a wider window only adds work to every
// iteration; with 2048-element tiles, 6.48 s at 32 against 6.83 s at 48,
// and 12.9 s and 14.6 s at 16 and 24.
static int narrow_span(int64_t L, int TC) {
    static const int base = [] {
        const char* e = getenv("PMCMC_SLIDE_NARROW");
        const int v = e ? atoi(e) : 0;
        return v > 0 ? v : 32;
    }();
    int64_t n = std::max<int64_t>(base, 2 * (int64_t)TC);
    n = (n + TC - 1) / TC * TC;
    return (int)std::min<int64_t>(n, L);
}

extern "C" int pmcmc_batch_create(pmcmc_ctx* c, int32_t n, const pmcmc_chain_desc* chains,
                                  const pmcmc_deer_params* p, int64_t window, pmcmc_batch** out) {
    if (!c || !out) return PMCMC_ERR_ARGUMENT;
    *out = nullptr;
    if (n < 1 || !chains || !p) return fail(c, PMCMC_ERR_ARGUMENT, "batch: need n >= 1 chains and params");
    if (window < 1) return fail(c, PMCMC_ERR_ARGUMENT, "batch: window must be >= 1");
    if (p->maxiter < 1) return fail(c, PMCMC_ERR_ARGUMENT, "maxiter must be >= 1");
    if (!(p->tol_abs >= 0) || !(p->tol_rel >= 0)) return fail(c, PMCMC_ERR_ARGUMENT, "tolerances must be >= 0");
    if (!(p->damping > 0 && p->damping <= 1)) return fail(c, PMCMC_ERR_ARGUMENT, "damping must be in (0,1]");
    if (p->jacobian != PMCMC_JAC_STOCH_DIAG || p->probes != 1)
        return fail(c, PMCMC_ERR_UNSUPPORTED, "batch: jacobian=:stoch_diag with one probe");
    if (p->precision != PMCMC_PREC_F32 && p->precision != PMCMC_PREC_BF16_JAC)
        return fail(c, PMCMC_ERR_UNSUPPORTED, "batch: FP32 or BF16-Jacobian precision");
    HIPCHK(c, hipSetDevice(c->device));
    pmcmc_batch* b = new (std::nothrow) pmcmc_batch();
    if (!b) return fail(c, PMCMC_ERR_ALLOC, "host allocation failed");
    b->ctx = c;
    b->p = *p;
    b->W = window;
    b->jbf = p->precision == PMCMC_PREC_BF16_JAC;
    b->ch.resize(n);
    size_t bytes = 0;
    std::vector<size_t> off(n);
    for (int i = 0; i < n; ++i) {
        const pmcmc_chain_desc& d = chains[i];
        BatchChain& h = b->ch[i];
        if (!d.target || d.T < 1 || !(d.eps > 0) || !d.x0 || !d.Xi || !d.U || !d.S_out) {
            delete b;
            return fail(c, PMCMC_ERR_ARGUMENT, "batch: chain %d: bad descriptor", i);
        }
        if (d.target->kind != TGT_LOGISTIC && d.target->kind != TGT_GAUSS) {
            delete b;
            return fail(c, PMCMC_ERR_UNSUPPORTED, "batch: chain %d: logistic or dense Gaussian targets only", i);
        }
        h.tgt = d.target;
        h.D = d.target->D;
        h.T = d.T;
        h.eps = d.eps;
        h.x0 = (const float*)d.x0;
        h.Xi = (const float*)d.Xi;
        h.U = (const float*)d.U;
        h.S_out = (float*)d.S_out;
        h.acc_out = d.accepted_out;
        h.Dp = d.target->kind == TGT_GAUSS ? d.target->gauss.Dp : d.target->glm.Dp;
        h.nslab = d.target->kind == TGT_GAUSS ? (int)((h.Dp + 63) / 64) : 1;
        if (h.Dp > 256) {
            delete b;
            return fail(c, PMCMC_ERR_UNSUPPORTED, "batch: chain %d: D > 256", i);
        }
        const int64_t L = window < d.T ? window : d.T;
        h.L = L;
        int TC, P, nT;
        batch_tile_shape(h.D, L, &TC, &P, &nT);
        h.TC = TC;
        h.P = P;
        h.nTmax = nT;
        h.span = (int)L;
        h.narrow = narrow_span(L, TC);
        if (tile_lds_bytes<float>(h.D, TC, true) > 64512) {
            delete b;
            return fail(c, PMCMC_ERR_UNSUPPORTED, "batch: chain %d: D too large for the tile kernel", i);
        }
        const size_t traj = align256((size_t)h.D * h.T * sizeof(float));
        const size_t mat = align256((size_t)h.D * L * sizeof(float));
        // log-densities: up to four 64-row slab partials per column (dense Gaussian)
        size_t sz = 2 * traj + 5 * mat + 2 * align256((size_t)4 * L * sizeof(double)) +
                    3 * align256((size_t)h.D * nT * sizeof(float)) + align256((size_t)nT * sizeof(double)) +
                    align256((size_t)h.T);
        h.rtg = RefineTarget{};
        h.rtg.D = h.D;
        if (d.target->kind == TGT_LOGISTIC) {
            h.rtg.kind = 0;
            h.rtg.N = d.target->glm.N;
            h.rtg.X64 = d.target->glm.Xd64;
            h.rtg.y64 = d.target->glm.y64;
            h.rtg.inv_s2 = 1.0 / (d.target->prior_scale * d.target->prior_scale);
            h.rtg.P64 = d.target->glm.P64;
            h.band = std::max(1e-4, 2.5e-8 * (double)d.target->glm.N);
        } else {
            h.rtg.kind = 1;
            h.rtg.L64 = d.target->gauss.L64;
            h.rtg.mu64 = d.target->gauss.mu64;
            h.band = 1e-3;
        }
        sz += align256(refine_chain_bytes(h.rtg, L, h.band));
        if (d.target->kind == TGT_LOGISTIC) {
            int64_t rps;
            glm_split_plan(&d.target->glm, L, &h.ns, &rps);
            sz += align256((size_t)h.ns * 2 * L * h.Dp * sizeof(float) + (size_t)h.ns * L * sizeof(double));
        }
        off[i] = bytes;
        bytes += sz;
    }
    bool ok = hipMalloc(&b->arena, bytes > 0 ? bytes : 256) == hipSuccess &&
              hipMalloc(&b->syncs, n * sizeof(SyncState)) == hipSuccess &&
              hipMemset(b->syncs, 0, n * sizeof(SyncState)) == hipSuccess &&
              hipMalloc(&b->slides, n * sizeof(SlideState)) == hipSuccess &&
              hipMalloc(&b->spans, n * sizeof(int)) == hipSuccess &&
              hipHostMalloc(&b->hsync, n * sizeof(SyncState)) == hipSuccess &&
              hipHostMalloc(&b->hslide, n * sizeof(SlideState)) == hipSuccess &&
              hipHostMalloc(&b->hspans, n * sizeof(int)) == hipSuccess;
    if (!ok) {
        batch_free(b);
        delete b;
        return fail(c, PMCMC_ERR_ALLOC, "batch: device allocation failed (%zu bytes)", bytes);
    }
    for (int i = 0; i < n; ++i) {
        BatchChain& h = b->ch[i];
        const int64_t L = h.L;
        const int nT = h.nTmax;
        char* q = (char*)b->arena + off[i];
        const size_t traj = align256((size_t)h.D * h.T * sizeof(float));
        const size_t mat = align256((size_t)h.D * L * sizeof(float));
        auto take = [&](size_t s) {
            char* r = q;
            q += s;
            return r;
        };
        h.buf[0] = (float*)take(traj);
        h.buf[1] = (float*)take(traj);
        h.Y = (float*)take(mat);
        h.gx = (float*)take(mat);
        h.gy = (float*)take(mat);
        h.W = (float*)take(mat);
        h.hw = (float*)take(mat);
        h.lpx = (double*)take(align256((size_t)4 * L * sizeof(double)));
        h.lpy = (double*)take(align256((size_t)4 * L * sizeof(double)));
        const size_t agg = align256((size_t)h.D * nT * sizeof(float));
        h.agg_a = (float*)take(agg);
        h.agg_b = (float*)take(agg);
        h.prev = (float*)take(agg);
        h.tile_smax = (double*)take(align256((size_t)nT * sizeof(double)));
        h.prev_dec = (int8_t*)take(align256((size_t)h.T));
        refine_chain_carve(&h.rw, h.rtg, L, h.band, take(align256(refine_chain_bytes(h.rtg, L, h.band))));
        if (h.tgt->kind == TGT_LOGISTIC) h.part = (float*)q;
    }
    {
        const char* rv = getenv("PMCMC_REFINE");
        b->refine = !(rv && !strcmp(rv, "0"));
    }
    ctx_retain(c);
    b->ctx_ref = true;
    *out = b;
    return PMCMC_OK;
}

extern "C" int pmcmc_batch_set_graph(pmcmc_batch* b, int32_t on) {
    if (!b) return PMCMC_ERR_ARGUMENT;
    b->use_graph = on != 0;
    return PMCMC_OK;
}

extern "C" int pmcmc_batch_set_profiling(pmcmc_batch* b, int32_t on) {
    if (!b) return PMCMC_ERR_ARGUMENT;
    pmcmc_ctx* c = b->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if (on && !b->flop_ctr) HIPCHK(c, hipMalloc(&b->flop_ctr, 2 * sizeof(unsigned long long)));
    b->prof_on = on != 0;
    b->prof.on = b->prof_on;
    b->prof.kinds = (1u << KT_LAMBDA_PASS) | (1u << KT_GLM_PASS);
    return PMCMC_OK;
}

extern "C" int pmcmc_batch_pass_stats(const pmcmc_batch* b, int64_t* launches, double* total_ms, double* flops) {
    if (!b) return PMCMC_ERR_ARGUMENT;
    if (launches) *launches = b->prof.n[KT_LAMBDA_PASS];
    if (total_ms) *total_ms = b->prof.ms[KT_LAMBDA_PASS];
    if (flops) *flops = b->prof_flops[0];
    return PMCMC_OK;
}

extern "C" int pmcmc_batch_glm_pass_stats(const pmcmc_batch* b, int64_t* launches, double* total_ms,
                                          double* flops) {
    if (!b) return PMCMC_ERR_ARGUMENT;
    if (launches) *launches = b->prof.n[KT_GLM_PASS];
    if (total_ms) *total_ms = b->prof.ms[KT_GLM_PASS];
    if (flops) *flops = b->prof_flops[1];
    return PMCMC_OK;
}

namespace {
// host mirrors of the per-chain descriptor tables (indexed by chain)
struct BatchTables {
    std::vector<TileArgs<float>> targs[2];
    std::vector<GlmPassArgs> gpass[3];  // pass 0 parity 0, pass 0 parity 1, pass 1
    std::vector<GlmReduceArgs> gred[3];
    std::vector<GaussPassArgs> qpass[3];
    std::vector<RefineTarget> rtgs;
    std::vector<RefineArgs> rargs[2];
    std::vector<RefineWork> rws;
    std::vector<SlideInit> inits;
    std::vector<BCopy> copies;
    std::vector<WinDesc> wins;
};
}  // namespace

// Window-independent fields of every descriptor; the window kernel fills in
// the rest at the start of each iteration.
static void batch_tables(pmcmc_batch* b, BatchTables* bt) {
    const pmcmc_deer_params& p = b->p;
    const int n = (int)b->ch.size();
    for (int i = 0; i < n; ++i) {
        BatchChain& h = b->ch[i];
        const int64_t D = h.D, L = h.L;
        SyncState* sy = b->syncs + i;
        const int* skip = &sy->converged;
        // one probe set per chain, attached to its absolute columns
        const uint64_t seed = p.seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(i + 1));
        const uint32_t key = probe_key(seed, 0xFFFFFFFFu, 0u);
        const int TC = (int)std::min<int64_t>(h.TC, L);
        const int P = std::min(h.P, TC);
        for (int par = 0; par < 2; ++par) {
            TileArgs<float> a{};
            a.D = D;
            a.TC = TC;
            a.P = P;
            a.skip = skip;
            a.sync = sy;
            a.eps = h.eps;
            a.damping = p.damping;
            a.tol_abs = p.tol_abs;
            a.tol_rel = p.tol_rel;
            a.iter = -1;
            a.probe_iter = 0xFFFFFFFFu;
            a.probes = 1;
            a.jacobian = JAC_STOCH;
            a.seed = seed;
            a.Y = h.Y;
            a.gx = h.gx;
            a.gy = h.gy;
            a.W = h.W;
            a.hw = h.hw;
            a.a_keep = h.W;  // the reduce pass keeps A, B over the probe products it has consumed
            a.b_keep = h.hw;
            a.lpx = h.lpx;
            a.lpy = h.lpy;
            a.lp_nslab = h.nslab;
            a.lp_stride = L;
            a.agg_a = h.agg_a;
            a.agg_b = h.agg_b;
            a.prev = h.prev;
            a.final_cols = &sy->skip_cols;  // stays 0: the window starts at the front
            a.tile_smax = h.tile_smax;
            a.hyst = 1e-6;
            a.win = b->slides + i;
            if (b->refine) {
                a.refine_on = h.rw.cnt + 2;
                if (h.tgt->kind == TGT_LOGISTIC) {
                    a.dec = h.rw.dec;  // decided by the refinement kernels
                } else {
                    // decided inline by the tile reduce (its target table entry is
                    // filled in once the tables have a device home)
                    a.rflag = h.rw.cnt + 2;
                    a.band = h.band;
                    a.gate = 100.0;
                    a.gate_iters = 16;
                }
            }
            bt->targs[par].push_back(a);
            RefineArgs ra{};
            ra.Y = h.Y;
            ra.gx = h.gx;
            ra.gy = h.gy;
            ra.lpx = h.lpx;
            ra.lpy = h.lpy;
            ra.lp_nslab = h.nslab;
            ra.lp_stride = L;
            ra.eps = h.eps;
            ra.band = h.band;
            ra.skip = skip;
            ra.last_metric = &sy->last_metric;
            ra.gate = 100.0;
            ra.iter = 0;
            ra.gate_iters = 16;
            ra.iter_dev = &sy->last_iter;
            ra.hyst = 1e-6;
            ra.final_cols = &sy->skip_cols;
            bt->rargs[par].push_back(ra);
        }
        bt->rws.push_back(h.rw);
        bt->rtgs.push_back(h.rtg);
        ColSrc<float> s0src{};
        s0src.keys[0] = key;
        s0src.K = 1;
        s0src.D = D;
        ColSrc<float> s1src = s0src;
        s1src.V = h.Y;
        s1src.Z = h.W;
        if (h.tgt->kind == TGT_LOGISTIC) {
            const GlmData& g = h.tgt->glm;
            const double inv_s2 = 1.0 / (h.tgt->prior_scale * h.tgt->prior_scale);
            double* llpart = (double*)((char*)h.part + (size_t)h.ns * 2 * L * h.Dp * sizeof(float));
            for (int v = 0; v < 3; ++v) {
                GlmPassArgs ga{};
                ga.src = v < 2 ? s0src : s1src;
                ga.Xrm = g.Xrm;
                ga.Xb = g.Xb;
                ga.Xtb = g.Xtb;
                ga.Np = g.Np;
                ga.y32 = g.y;
                ga.N = g.N;
                ga.part = h.part;
                ga.llpart = llpart;
                ga.skip = skip;
                ga.flops = b->prof_on ? b->flop_ctr + 1 : nullptr;
                bt->gpass[v].push_back(ga);
                GlmReduceArgs gr{};
                gr.src = ga.src;
                gr.part = h.part;
                gr.llpart = llpart;
                gr.Dp = h.Dp;
                gr.llstride = L;
                gr.pass = v < 2 ? 0 : 1;
                gr.inv_s2 = inv_s2;
                gr.P = h.tgt->glm.P64;
                gr.eps = h.eps;
                gr.gout = v < 2 ? h.gx : h.gy;
                gr.hout = h.hw;
                gr.lpout = v < 2 ? h.lpx : h.lpy;
                gr.Yout = v < 2 ? h.Y : nullptr;
                gr.Wout = v < 2 ? h.W : nullptr;
                gr.skip = skip;
                bt->gred[v].push_back(gr);
                bt->qpass[v].push_back(GaussPassArgs{});
            }
        } else {
            const GaussData& g = h.tgt->gauss;
            for (int v = 0; v < 3; ++v) {
                GaussPassArgs qa{};
                qa.src = v < 2 ? s0src : s1src;
                qa.L32 = g.L32;
                qa.ldL = (int)g.Dp;
                qa.mu = g.mu32;
                qa.pass = v < 2 ? 0 : 1;
                qa.eps = (float)h.eps;
                qa.gout = v < 2 ? h.gx : h.gy;
                qa.hout = h.hw;
                qa.lpout = v < 2 ? h.lpx : h.lpy;
                qa.Yout = v < 2 ? h.Y : nullptr;
                qa.Wout = v < 2 ? h.W : nullptr;
                qa.skip = skip;
                qa.lp_stride = L;
                qa.flops = b->prof_on ? b->flop_ctr : nullptr;
                bt->qpass[v].push_back(qa);
                bt->gpass[v].push_back(GlmPassArgs{});
                bt->gred[v].push_back(GlmReduceArgs{});
            }
        }
        SlideInit in{};
        in.win = b->slides + i;
        in.sy = sy;
        in.rcnt = h.rw.cnt;
        in.prev_dec = h.prev_dec;
        in.T = h.T;
        bt->inits.push_back(in);
        BCopy cp{};
        cp.b0 = h.buf[0];
        cp.b1 = h.buf[1];
        cp.dst = h.S_out;
        cp.n = D * h.T;
        cp.sy = sy;
        bt->copies.push_back(cp);
        WinDesc wd{};
        wd.D = D;
        wd.T = h.T;
        wd.L = L;
        wd.TC = TC;
        wd.kind = h.tgt->kind == TGT_LOGISTIC ? 0 : 1;
        wd.buf[0] = h.buf[0];
        wd.buf[1] = h.buf[1];
        wd.x0 = h.x0;
        wd.Xi = h.Xi;
        wd.U = h.U;
        wd.acc_out = h.acc_out;
        wd.prev_dec = h.prev_dec;
        wd.win = b->slides + i;
        wd.sy = sy;
        wd.span = b->spans + i;
        bt->wins.push_back(wd);  // table addresses filled in once the tables have a device home
    }
}

static void build_items(const pmcmc_batch* b, const std::vector<char>& live, Items* it) {
    const int n = (int)b->ch.size();
    // N splits of the GEMM passes: sized to fill the GPU from each chain (768
    // workgroups) while few GEMM column blocks are live, to 512 per chain while
    // at least 512 are, since every split adds a partial sum of the window
    // that the pass writes and the reduce reads (the plan's split count, the
    // most, sets the buffers).  256 uneven chains: 5.91 s at 768 throughout,
    // 5.69 s at 512 throughout; their 32-chain subset at damping 0.5, where
    // a few logistic chains stay live for long, 23.5 s and 28.5 s.
    int64_t gemm_blocks = 0;
    for (int k = 0; k < n; ++k)
        if (live[k] && b->ch[k].tgt->kind == TGT_LOGISTIC)
            gemm_blocks += (std::min<int64_t>(b->ch[k].span, b->ch[k].L) + GLM_CB - 1) / GLM_CB;
    const int split_target = gemm_blocks >= 512 ? 512 : 768;
    for (int k = 0; k < n; ++k) {
        if (!live[k]) continue;
        const BatchChain& h = b->ch[k];
        const int64_t span = std::min<int64_t>(h.span, h.L);
        const int TC = (int)std::min<int64_t>(h.TC, h.L);
        const int nT = (int)((span + TC - 1) / TC);
        const size_t lds = tile_lds_bytes<float>(h.D, TC, true);
        const int lc = lds_class(lds);
        for (int t = 0; t < nT; ++t) it->red[lc].push_back(make_int2(k, t));
        // chains whose windows never exceed PMCMC_AGG_INLINE tiles fold their
        // aggregates in the apply pass (agg_inline, set by the window kernel)
        if (nT > PMCMC_AGG_INLINE || h.D > PMCMC_NT) it->multi.push_back(k);
        it->wins.push_back(k);
        if (lds > it->lds_multi[lc]) it->lds_multi[lc] = lds;
        it->r_all.push_back(k);
        it->r_maxT = std::max(it->r_maxT, span);
        const int dpt = (int)(h.Dp / 16);
        if (h.tgt->kind == TGT_LOGISTIC) {
            it->r_glm.push_back(k);
            it->r_split_glm = std::max(it->r_split_glm, h.rw.nsplit);
            const int ncb = (int)((span + GLM_CB - 1) / GLM_CB);
            int nsk;
            int64_t rps;
            glm_split_plan(&h.tgt->glm, span, &nsk, &rps, split_target);
            nsk = std::min(nsk, h.ns);
            for (int cb = 0; cb < ncb; ++cb)
                for (int sp = 0; sp < nsk; ++sp) it->glm[dpt].push_back(make_int4(k, cb, sp, nsk));
            for (int64_t blk = 0; blk < (span + 3) / 4; ++blk) it->glm_red.push_back(make_int4(k, (int)blk, nsk, 0));
        } else {
            it->r_gauss.push_back(k);
            it->r_split_gauss = std::max(it->r_split_gauss, h.rw.nsplit);
            const int gcb = (int)((span + GAUSS_CB - 1) / GAUSS_CB);
            for (int cb = 0; cb < gcb; ++cb)
                for (int sl = 0; sl < h.nslab; ++sl) it->gslab.push_back(make_int4(k, cb, sl, 0));
        }
    }
}

static int upload(pmcmc_ctx* c, hipStream_t s, const Pack& pk, void** buf, size_t* cap) {
    if (pk.host.size() > *cap) {
        HIPCHK(c, hipStreamSynchronize(s));
        if (*buf) HIPCHK(c, hipFree(*buf));
        *buf = nullptr;
        *cap = 0;
        HIPCHK(c, hipMalloc(buf, pk.host.size()));
        *cap = pk.host.size();
    }
    HIPCHK(c, hipMemcpyAsync(*buf, pk.host.data(), pk.host.size(), hipMemcpyHostToDevice, s));
    return PMCMC_OK;
}

// The host's window-length policy at a synchronisation.  A chain starts with
// a short window (PMCMC_SLIDE_START=wide: the full window), a few tiles past
// its front.  A chain whose front moves 4 or more columns per iteration over
// the last chunk converges faster than causally: its window widens 4x (up to
// the full window).  A wide chain that after 24 iterations of a window
// segment neither reduced its metric 4x over the last chunk nor moved its
// front 2 columns per iteration is converging causally (about one column per
// iteration, the Jacobian estimate missing the coupling of its dynamics): its
// window shrinks back to the short one, where all its progress happens.
// Returns true when a length changed.
static bool slide_start_wide() {
    static const bool wide = [] {
        const char* e = getenv("PMCMC_SLIDE_START");
        return e && !strcmp(e, "wide");
    }();
    return wide;
}

static bool batch_policy(pmcmc_batch* b, int i) {
    BatchChain& h = b->ch[i];
    const SyncState& hs = b->hsync[i];
    const SlideState& sl = b->hslide[i];
    const int k = hs.last_iter;
    const int dk = k - h.prev_iter;
    const int64_t prog = sl.front - h.prev_front;
    const int old = h.span;
    if (h.span > h.narrow && dk > 0 && prog < 2 * (int64_t)dk) {
        const int seg = k + 1 - sl.seg_start;
        const bool falling = hs.last_metric < h.prev_metric * 0.25;  // false for NaN
        if (seg >= 24 && dk >= 8 && !falling) h.span = h.narrow;
    } else if (h.span < h.L && dk > 0 && prog >= 4 * (int64_t)dk) {
        const int64_t TC = std::min<int64_t>(h.TC, h.L);
        int64_t ns = std::min<int64_t>(4 * (int64_t)h.span, h.L);
        ns = std::min<int64_t>((ns + TC - 1) / TC * TC, h.L);
        h.span = (int)ns;
    }
    h.prev_front = sl.front;
    h.prev_iter = k;
    h.prev_metric = hs.last_metric;
    return h.span != old;
}

extern "C" int pmcmc_batch_solve(pmcmc_batch* b, pmcmc_deer_info* infos) {
    if (!b) return PMCMC_ERR_ARGUMENT;
    // PMCMC_BATCH_PROGRESS=1: a line per rebuild of the work items on stderr
    const char* pe = getenv("PMCMC_BATCH_PROGRESS");
    const bool progress = pe && pe[0] == '1';
    const auto t_start = std::chrono::steady_clock::now();
    pmcmc_ctx* c = b->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int n = (int)b->ch.size();
    const bool graphs = b->use_graph && !b->prof_on;  // profiled solves launch directly
    if (b->prof_on) {
        b->prof.reset();
        b->prof_flops[0] = b->prof_flops[1] = 0;
        HIPCHK(c, hipMemsetAsync(b->flop_ctr, 0, 2 * sizeof(unsigned long long), s));
    }
    for (auto& h : b->ch) {
        h.iters = 0;
        h.converged = true;
        h.metric = 0;
        h.span = slide_start_wide() ? (int)h.L : h.narrow;
        h.prev_front = 0;
        h.prev_iter = -1;
        h.prev_metric = INFINITY;
        h.stopped = false;
    }
    // descriptor tables: uploaded once; the window kernel keeps their
    // window-dependent fields current
    BatchTables bt;
    batch_tables(b, &bt);
    Pack tp;
    size_t o_t[2], o_gp[3], o_gr[3], o_qp[3], o_ra[2];
    for (int v = 0; v < 2; ++v) o_t[v] = tp.put(bt.targs[v].data(), n * sizeof(TileArgs<float>));
    for (int v = 0; v < 3; ++v) {
        o_gp[v] = tp.put(bt.gpass[v].data(), n * sizeof(GlmPassArgs));
        o_gr[v] = tp.put(bt.gred[v].data(), n * sizeof(GlmReduceArgs));
        o_qp[v] = tp.put(bt.qpass[v].data(), n * sizeof(GaussPassArgs));
    }
    const size_t o_rt = tp.put(bt.rtgs.data(), n * sizeof(RefineTarget));
    const size_t o_rw = tp.put(bt.rws.data(), n * sizeof(RefineWork));
    for (int v = 0; v < 2; ++v) o_ra[v] = tp.put(bt.rargs[v].data(), n * sizeof(RefineArgs));
    const size_t o_init = tp.put(bt.inits.data(), n * sizeof(SlideInit));
    const size_t o_copy = tp.put(bt.copies.data(), n * sizeof(BCopy));
    const size_t o_win = tp.put(bt.wins.data(), n * sizeof(WinDesc));
    if (tp.host.size() > b->tables_cap) {
        HIPCHK(c, hipStreamSynchronize(s));
        if (b->tables) HIPCHK(c, hipFree(b->tables));
        b->tables = nullptr;
        b->tables_cap = 0;
        HIPCHK(c, hipMalloc(&b->tables, tp.host.size()));
        b->tables_cap = tp.host.size();
    }
    char* T0 = (char*)b->tables;
    auto at = [&](size_t o) { return (void*)(T0 + o); };
    {
        WinDesc* wh = reinterpret_cast<WinDesc*>(tp.host.data() + o_win);
        for (int i = 0; i < n; ++i) {
            for (int v = 0; v < 2; ++v) {
                wh[i].ta[v] = (TileArgs<float>*)at(o_t[v]) + i;
                wh[i].ra[v] = (RefineArgs*)at(o_ra[v]) + i;
            }
            for (int v = 0; v < 3; ++v) {
                wh[i].gp[v] = (GlmPassArgs*)at(o_gp[v]) + i;
                wh[i].gr[v] = (GlmReduceArgs*)at(o_gr[v]) + i;
                wh[i].qp[v] = (GaussPassArgs*)at(o_qp[v]) + i;
            }
            wh[i].rw = (RefineWork*)at(o_rw) + i;
            if (b->refine && b->ch[i].tgt->kind != TGT_LOGISTIC)
                for (int v = 0; v < 2; ++v) {
                    TileArgs<float>* th = reinterpret_cast<TileArgs<float>*>(tp.host.data() + o_t[v]) + i;
                    th->rtg = (const RefineTarget*)at(o_rt) + i;
                }
        }
    }
    HIPCHK(c, hipMemcpyAsync(b->tables, tp.host.data(), tp.host.size(), hipMemcpyHostToDevice, s));
    for (int i = 0; i < n; ++i) b->hspans[i] = b->ch[i].span;
    HIPCHK(c, hipMemcpyAsync(b->spans, b->hspans, n * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(batch_slide_init_kernel, dim3(16, (unsigned)n), dim3(256), 0, s, (const SlideInit*)at(o_init));
    HIPCHK(c, hipGetLastError());

    std::vector<char> live(n, 1);
    int nlive = n, built_live = 0;
    Items it;
    Pack ip;
    size_t o_red[NLDS] = {0, 0, 0}, o_multi = 0, o_wins = 0, o_glmred = 0, o_gi[17], o_qs = 0;
    size_t o_rall = 0, o_rglm = 0, o_rgauss = 0;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    auto drop_graph = [&]() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    };
    auto iteration = [&](int par) {
        char* I0 = (char*)b->items;
        auto ia = [&](size_t o) { return (void*)(I0 + o); };
        hipLaunchKernelGGL(batch_window_kernel, dim3(4, (unsigned)it.wins.size()), dim3(256), 0, s,
                           (const WinDesc*)at(o_win), (const int*)ia(o_wins), par);
        for (int d = 1; d <= 16; ++d) {
            const int pg = it.glm[d].empty() ? -1 : b->prof.begin(KT_GLM_PASS, 0, s);
            launch_glm_pass_batched(d, b->jbf, (const GlmPassArgs*)at(o_gp[par]), (const int4*)ia(o_gi[d]),
                                    (int)it.glm[d].size(), s);
            b->prof.end(pg, s);
        }
        {
            const int pe = it.gslab.empty() ? -1 : b->prof.begin(KT_LAMBDA_PASS, 0, s);
            launch_gauss_slab_batched(b->jbf, (const GaussPassArgs*)at(o_qp[par]), (const int4*)ia(o_qs),
                                      (int)it.gslab.size(), s);
            b->prof.end(pe, s);
        }
        launch_glm_reduce_batched((const GlmReduceArgs*)at(o_gr[par]), (const int4*)ia(o_glmred),
                                  (int)it.glm_red.size(), s);
        for (int d = 1; d <= 16; ++d) {
            const int pg = it.glm[d].empty() ? -1 : b->prof.begin(KT_GLM_PASS, 0, s);
            launch_glm_pass_batched(d, b->jbf, (const GlmPassArgs*)at(o_gp[2]), (const int4*)ia(o_gi[d]),
                                    (int)it.glm[d].size(), s);
            b->prof.end(pg, s);
        }
        {
            const int pe = it.gslab.empty() ? -1 : b->prof.begin(KT_LAMBDA_PASS, 0, s);
            launch_gauss_slab_batched(b->jbf, (const GaussPassArgs*)at(o_qp[2]), (const int4*)ia(o_qs),
                                      (int)it.gslab.size(), s);
            b->prof.end(pe, s);
        }
        launch_glm_reduce_batched((const GlmReduceArgs*)at(o_gr[2]), (const int4*)ia(o_glmred),
                                  (int)it.glm_red.size(), s);
        if (b->refine) {
            RefineBatch rb{};
            rb.tgs = (const RefineTarget*)at(o_rt);
            rb.ras = (const RefineArgs*)at(o_ra[par]);
            rb.rws = (const RefineWork*)at(o_rw);
            rb.all = (const int*)ia(o_rall);
            rb.n_all = (int)it.r_all.size();
            rb.glm = (const int*)ia(o_rglm);
            rb.n_glm = (int)it.r_glm.size();
            rb.gauss = (const int*)ia(o_rgauss);
            rb.n_gauss = 0;  // dense-Gaussian chains decide inline in the tile reduce
            rb.max_T = it.r_maxT;
            rb.max_split_glm = it.r_split_glm;
            rb.max_split_gauss = it.r_split_gauss;
            refine_decisions_batched(rb, par, s);
        }
        const TileArgs<float>* ta = (const TileArgs<float>*)at(o_t[par]);
        for (int lc = 0; lc < NLDS; ++lc)
            if (!it.red[lc].empty())
                hipLaunchKernelGGL((tile_kernel_batched<float, MODE_GEMM, PASS_REDUCE>),
                                   dim3((unsigned)it.red[lc].size()), dim3(PMCMC_NT), it.lds_multi[lc], s, ta,
                                   (const int2*)ia(o_red[lc]));
        if (!it.multi.empty())
            hipLaunchKernelGGL(batch_agg_scan_kernel, dim3((unsigned)it.multi.size()), dim3(256), 0, s, ta,
                               (const int*)ia(o_multi));
        for (int lc = 0; lc < NLDS; ++lc)
            if (!it.red[lc].empty())
                hipLaunchKernelGGL((tile_kernel_batched<float, MODE_GEMM, PASS_APPLY>),
                                   dim3((unsigned)it.red[lc].size()), dim3(PMCMC_NT), it.lds_multi[lc], s, ta,
                                   (const int2*)ia(o_red[lc]));
    };
    double rebuild_s = 0.0;
    auto rebuild = [&]() -> int {
        const auto r0 = std::chrono::steady_clock::now();
        struct Acc {
            double* s;
            std::chrono::steady_clock::time_point t;
            ~Acc() { *s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count(); }
        } acc_{&rebuild_s, r0};
        drop_graph();
        it = Items();
        build_items(b, live, &it);
        ip = Pack();
        for (int lc = 0; lc < NLDS; ++lc) o_red[lc] = ip.put(it.red[lc].data(), it.red[lc].size() * sizeof(int2));
        o_multi = ip.put(it.multi.data(), it.multi.size() * sizeof(int));
        o_wins = ip.put(it.wins.data(), it.wins.size() * sizeof(int));
        o_rall = ip.put(it.r_all.data(), it.r_all.size() * sizeof(int));
        o_rglm = ip.put(it.r_glm.data(), it.r_glm.size() * sizeof(int));
        o_rgauss = ip.put(it.r_gauss.data(), it.r_gauss.size() * sizeof(int));
        o_glmred = ip.put(it.glm_red.data(), it.glm_red.size() * sizeof(int4));
        for (int d = 0; d < 17; ++d) o_gi[d] = ip.put(it.glm[d].data(), it.glm[d].size() * sizeof(int4));
        o_qs = ip.put(it.gslab.data(), it.gslab.size() * sizeof(int4));
        const int rc2 = upload(c, s, ip, &b->items, &b->items_cap);
        if (rc2 != PMCMC_OK) return rc2;
        built_live = nlive;
        if (!graphs) return PMCMC_OK;
        HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        iteration(0);
        iteration(1);
        const hipError_t ec = hipStreamEndCapture(s, &graph);
        if (ec != hipSuccess) return fail(c, PMCMC_ERR_HIP, "batch: graph capture failed: %s", hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            drop_graph();
            return fail(c, PMCMC_ERR_HIP, "batch: graph instantiate failed: %s", hipGetErrorString(ei));
        }
        return PMCMC_OK;
    };
    int rc = rebuild();
    if (rc != PMCMC_OK) return rc;
    // iteration caps: maxiter per window's worth of columns, and maxiter
    // iterations without the front moving
    int k = 0, chunk = 2, rebuilds = 1;
    while (nlive > 0) {
        const int m = chunk;  // even
        for (int j = 0; j < m / 2; ++j) {
            if (graphs) {
                const hipError_t el = hipGraphLaunch(exec, s);
                if (el != hipSuccess) {
                    drop_graph();
                    return fail(c, PMCMC_ERR_HIP, "batch: graph launch failed");
                }
            } else {
                iteration(0);
                iteration(1);
            }
        }
        k += m;
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(b->hsync, b->syncs, n * sizeof(SyncState), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipMemcpyAsync(b->hslide, b->slides, n * sizeof(SlideState), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        if (b->prof_on) b->prof.flush(INT_MAX);
        nlive = 0;
        bool changed = false, stop_any = false;
        for (int q = 0; q < n; ++q) {
            if (!live[q]) continue;
            SyncState& hs = b->hsync[q];
            BatchChain& h = b->ch[q];
            if (hs.converged) {
                live[q] = 0;
                continue;
            }
            const int64_t windows = (h.T + h.L - 1) / h.L;
            const bool over = (int64_t)hs.last_iter + 1 >= (int64_t)b->p.maxiter * windows;
            const bool stalled = hs.last_iter - std::max(b->hslide[q].last_adv_iter, b->hslide[q].seg_start - 1) >=
                                 b->p.maxiter;
            if (over || stalled) {
                // give up on the chain: mark it done on the device (the host copy is
                // current and no kernel runs now)
                h.stopped = true;
                hs.converged = 1;
                hs.conv_iter = hs.last_iter;
                HIPCHK(c, hipMemcpyAsync(b->syncs + q, &hs, sizeof(SyncState), hipMemcpyHostToDevice, s));
                live[q] = 0;
                stop_any = true;
                continue;
            }
            ++nlive;
            changed = batch_policy(b, q) || changed;
        }
        if (chunk < 32) chunk <<= 1;
        if (nlive > 0 && (changed || stop_any || 8 * nlive <= 7 * built_live)) {
            for (int q = 0; q < n; ++q) b->hspans[q] = b->ch[q].span;
            HIPCHK(c, hipMemcpyAsync(b->spans, b->hspans, n * sizeof(int), hipMemcpyHostToDevice, s));
            rc = rebuild();
            if (rc != PMCMC_OK) {
                drop_graph();
                return rc;
            }
            ++rebuilds;
            if (progress) {
                const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
                size_t nt = 0;
                for (int lc = 0; lc < NLDS; ++lc) nt += it.red[lc].size();
                fprintf(stderr, "[pmcmc batch] iteration %d: %d live chains, %zu tiles, %.3f s\n", k, nlive, nt, el);
                fflush(stderr);
            }
        }
    }
    drop_graph();
    hipLaunchKernelGGL(batch_copy_kernel, dim3(64, (unsigned)n), dim3(256), 0, s, (const BCopy*)at(o_copy));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(b->hsync, b->syncs, n * sizeof(SyncState), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(b->hslide, b->slides, n * sizeof(SlideState), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (b->prof_on) {
        unsigned long long fl[2] = {0, 0};
        HIPCHK(c, hipMemcpy(fl, b->flop_ctr, sizeof fl, hipMemcpyDeviceToHost));
        b->prof_flops[0] = (double)fl[0];
        b->prof_flops[1] = (double)fl[1];
    }
    for (int i = 0; i < n; ++i) {
        const SyncState& hs = b->hsync[i];
        BatchChain& h = b->ch[i];
        h.iters = hs.last_iter + 1;
        h.converged = hs.converged && !h.stopped;
        h.metric = hs.last_metric;
    }
    if (progress) {
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        fprintf(stderr, "[pmcmc batch] done: %d iterations, %d rebuilds (%.3f s), %.3f s\n", k, rebuilds, rebuild_s,
                el);
        fflush(stderr);
    }
    if (infos)
        for (int i = 0; i < n; ++i) {
            infos[i].converged = b->ch[i].converged;
            infos[i].iters = b->ch[i].iters;
            infos[i].metric = b->ch[i].metric;
        }
    return PMCMC_OK;
}
