// Logistic-regression target kernels (placeholder until the MFMA path lands).
#include "glm.hpp"

#include "../../include/pmcmc_amd.h"

namespace pmcmc {

int glm_data_create(GlmData* g, hipStream_t, int64_t N, int64_t D, bool, const void*, const void*,
                    std::string* msg) {
    g->N = N;
    g->D = D;
    *msg = "logistic target: not built yet";
    return PMCMC_ERR_UNSUPPORTED;
}

void glm_data_destroy(GlmData* g) {
    hipFree(g->Xrm);
    hipFree(g->Xcm);
    hipFree(g->Xd64);
    hipFree(g->y);
    hipFree(g->y64);
    *g = GlmData{};
}

int glm_work_ensure(GlmWork*, const GlmData*, hipStream_t, int64_t, int64_t, int, bool,
                    std::string* msg) {
    *msg = "logistic target: not built yet";
    return PMCMC_ERR_UNSUPPORTED;
}

void glm_work_destroy(GlmWork* w) {
    hipFree(w->Y);
    hipFree(w->gx);
    hipFree(w->gy);
    hipFree(w->W);
    hipFree(w->hw);
    hipFree(w->lpx);
    hipFree(w->lpy);
    hipFree(w->part);
    *w = GlmWork{};
}

template <typename T>
int glm_newton_products(GlmWork*, const GlmData*, hipStream_t, SyncState*, int, double, double,
                        const T*, const T*, const T*, uint64_t, int, int, int64_t, int64_t,
                        std::string* msg) {
    *msg = "logistic target: not built yet";
    return PMCMC_ERR_UNSUPPORTED;
}
template int glm_newton_products<float>(GlmWork*, const GlmData*, hipStream_t, SyncState*, int,
                                        double, double, const float*, const float*, const float*,
                                        uint64_t, int, int, int64_t, int64_t, std::string*);
template int glm_newton_products<double>(GlmWork*, const GlmData*, hipStream_t, SyncState*, int,
                                         double, double, const double*, const double*,
                                         const double*, uint64_t, int, int, int64_t, int64_t,
                                         std::string*);

int glm_fwd_and_jvp(const GlmData*, hipStream_t, int, double, double, const void*, const void*,
                    const void*, const void*, void*, void*, uint8_t*, double*, int64_t, int64_t,
                    std::string* msg) {
    *msg = "logistic target: not built yet";
    return PMCMC_ERR_UNSUPPORTED;
}

int glm_logdensity(const GlmData*, hipStream_t, int, double, const void*, double*, int64_t,
                   std::string* msg) {
    *msg = "logistic target: not built yet";
    return PMCMC_ERR_UNSUPPORTED;
}

}  // namespace pmcmc
