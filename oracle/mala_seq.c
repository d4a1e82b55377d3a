/*
 * Sequential taped MALA in plain C, FP64: a second, independent restatement of
 * the reference's scalar sampler, written to cross-check the numpy oracle
 * (oracle/deer_oracle.py).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/ (and tests/golden/'s fixture script)
 * load it, as a checker; the product path never links or calls it.
 *
 * Restates (rsenne/ParallelMCMC.jl):
 *   log q(y | x) = -|y - x - eps g(x)|^2 / (4 eps) - (D/2) log(4 pi eps)
 *                                                 src/MALA/MALA.jl:172-193
 *   proposal y = x + eps g(x) + sqrt(2 eps) xi    src/MALA/MALA.jl:207-229
 *   taped step, accept iff log(u) < log alpha     src/MALA/MALA.jl:276-305 (line 299)
 *   sequential run over the tape                  src/MALA/MALA.jl:322-346
 * Targets (the analytic ones the GPU kernels implement):
 *   0 diagonal Gaussian  logp = -1/2 sum_d prec_d (x_d - m_d)^2
 *   1 dense Gaussian     logp = -1/2 (x - m)^T Lam (x - m)
 *   2 logistic regression, prior N(0, s^2 I)
 *     (benchmarks/ParallelMCMCBenchmarks/src/models/bayes_logreg.jl:118-154):
 *     logp = sum_n [-y_n softplus(-z_n) - (1 - y_n) softplus(z_n)] - |x|^2 / (2 s^2),
 *     z = X x, grad = X^T (y - sigmoid(z)) - x / s^2
 * A zero uniform accepts, as in the batched reference path (src/MALA/MALA.jl:523-526).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int32_t kind;       /* 0 diagonal Gaussian, 1 dense Gaussian, 2 logistic */
    int32_t pad;
    int64_t D, N;
    const double* prec; /* kind 0: D */
    const double* mean; /* kinds 0, 1: D, or NULL for zero */
    const double* Lam;  /* kind 1: D x D, column-major */
    const double* X;    /* kind 2: N x D design, column-major */
    const double* y;    /* kind 2: N labels */
    double prior_scale; /* kind 2 */
} mala_target;

static double softplus(double z) { return fmax(z, 0.0) + log1p(exp(-fabs(z))); }

static double sigmoid(double z) {
    if (z >= 0.0) return 1.0 / (1.0 + exp(-z));
    const double e = exp(z);
    return e / (1.0 + e);
}

/* log density at x; its gradient into g (work: N doubles for the logistic target) */
static double logp_grad(const mala_target* t, const double* x, double* g, double* work) {
    const int64_t D = t->D;
    double lp = 0.0;
    if (t->kind == 0) {
        for (int64_t d = 0; d < D; ++d) {
            const double r = x[d] - (t->mean ? t->mean[d] : 0.0);
            lp += -0.5 * t->prec[d] * r * r;
            g[d] = -t->prec[d] * r;
        }
    } else if (t->kind == 1) {
        for (int64_t i = 0; i < D; ++i) {
            double s = 0.0;
            for (int64_t k = 0; k < D; ++k) s += t->Lam[i + k * D] * (x[k] - (t->mean ? t->mean[k] : 0.0));
            g[i] = -s;
        }
        for (int64_t i = 0; i < D; ++i) lp += -0.5 * (x[i] - (t->mean ? t->mean[i] : 0.0)) * (-g[i]);
    } else {
        const int64_t N = t->N;
        const double is2 = 1.0 / (t->prior_scale * t->prior_scale);
        for (int64_t n = 0; n < N; ++n) {
            double z = 0.0;
            for (int64_t d = 0; d < D; ++d) z += t->X[n + d * N] * x[d];
            work[n] = z;
        }
        double ll = 0.0, xx = 0.0;
        for (int64_t n = 0; n < N; ++n) ll += -t->y[n] * softplus(-work[n]) - (1.0 - t->y[n]) * softplus(work[n]);
        for (int64_t d = 0; d < D; ++d) xx += x[d] * x[d];
        lp = ll - 0.5 * xx * is2;
        for (int64_t n = 0; n < N; ++n) work[n] = t->y[n] - sigmoid(work[n]);
        for (int64_t d = 0; d < D; ++d) {
            double s = 0.0;
            for (int64_t n = 0; n < N; ++n) s += t->X[n + d * N] * work[n];
            g[d] = s - x[d] * is2;
        }
    }
    return lp;
}

/* log q(a | b) with b's gradient gb */
static double logq(const double* a, const double* b, const double* gb, double eps, int64_t D) {
    double q = 0.0;
    for (int64_t d = 0; d < D; ++d) {
        const double r = a[d] - (b[d] + eps * gb[d]);
        q += r * r;
    }
    return -0.5 * q / (2.0 * eps) - ((double)D / 2.0) * log(4.0 * M_PI * eps);
}

/*
 * Runs T taped steps from x0; S (D x T, column-major) receives the states,
 * acc the decisions (1 accepted), la the log acceptance ratios.  Xi is D x T
 * column-major, U has T entries.  Returns 0, or 1 when a buffer cannot be
 * allocated or an argument is out of range.
 */
int mala_seq_taped(const mala_target* t, double eps, const double* x0, const double* Xi, const double* U,
                   int64_t T, double* S, uint8_t* acc, double* la) {
    if (!t || t->D < 1 || T < 0 || !(eps > 0.0) || t->kind < 0 || t->kind > 2) return 1;
    const int64_t D = t->D;
    double* x = (double*)malloc(sizeof(double) * D);
    double* y = (double*)malloc(sizeof(double) * D);
    double* gx = (double*)malloc(sizeof(double) * D);
    double* gy = (double*)malloc(sizeof(double) * D);
    double* work = (double*)malloc(sizeof(double) * (t->N > 0 ? t->N : 1));
    if (!x || !y || !gx || !gy || !work) {
        free(x), free(y), free(gx), free(gy), free(work);
        return 1;
    }
    memcpy(x, x0, sizeof(double) * D);
    const double sq = sqrt(2.0 * eps);
    for (int64_t s = 0; s < T; ++s) {
        const double* xi = Xi + s * D;
        const double lpx = logp_grad(t, x, gx, work);
        for (int64_t d = 0; d < D; ++d) y[d] = x[d] + eps * gx[d] + sq * xi[d];
        const double lpy = logp_grad(t, y, gy, work);
        const double lq_yx = logq(y, x, gx, eps, D);
        const double lq_xy = logq(x, y, gy, eps, D);
        const double a = (lpy + lq_xy) - (lpx + lq_yx);
        const int accept = U[s] > 0.0 ? log(U[s]) < a : 1;
        if (accept) memcpy(x, y, sizeof(double) * D);
        memcpy(S + s * D, x, sizeof(double) * D);
        acc[s] = (uint8_t)accept;
        la[s] = a;
    }
    free(x), free(y), free(gx), free(gy), free(work);
    return 0;
}
