/*
 * mds_oracle.c -- CPU restatement of the MassiveMDS serial engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * liboracle.so.  The HIP engine in massivemds_amd/ never links or calls it.
 *
 * It restates, loop for loop, the reference's serial no-SIMD engine
 * NewMultiDimensionalScaling<DoubleNoSimdTypeInfo, CpuAccumulate> (and the
 * FloatNoSimdTypeInfo variant), with libm for the transcendentals:
 *
 *   likelihood  src/NewMultiDimensionalScaling.hpp:193-203  getSumOfIncrements
 *               src/NewMultiDimensionalScaling.hpp:524-561  computeSumOfIncrementsGeneric
 *               src/NewMultiDimensionalScaling.hpp:494-522  innerLikelihoodLoop
 *   gradient    src/NewMultiDimensionalScaling.hpp:281-302  getLogLikelihoodGradient
 *               src/NewMultiDimensionalScaling.hpp:304-334  computeLogLikelihoodGradientGeneric
 *               src/NewMultiDimensionalScaling.hpp:421-461  innerGradientLoop
 *   masking     src/NewMultiDimensionalScaling.hpp:402-410  getMissing / mask (scalar)
 *   distance    src/Distance.hpp:77-88 (generic P), :130-140 (P == 2)
 *   log Phi     src/CDF.h:583-592  phi_new(x) = log(0.5 * erfc(-x / sqrt 2))
 *   phi         src/CDF.h:608-616  pdf_new(x) = exp(-x^2 / 2) / sqrt(2 pi)
 *   float copy  src/MemoryManagement.hpp:45-56 (double -> float keeps NaN)
 *
 * Pinning (see tests/test_oracle.py): the BEAST known answer held in the
 * reference's inst/extdata (non-truncated log-likelihood of the h3 data),
 * mpmath closed forms at 50 digits (R/Untitled.R:16-52 restated),
 * high-precision central finite differences, translation and rotation
 * invariants, tau scaling, and hand-computed N = 2 cases.
 *
 * Build: oracle/Makefile (gcc, -ffp-contract=off, no -ffast-math).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_1_SQRT_2PI 0.398942280401432677939946059934
#define ORACLE_SQRT1_2 0.707106781186547524400844362105

/* src/CDF.h:583-592 (double instantiation). */
static double phi_new_d(double x) { return log(0.5 * erfc(-x * ORACLE_SQRT1_2)); }
/* src/CDF.h:608-616 (double instantiation). */
static double pdf_new_d(double x) { return ORACLE_1_SQRT_2PI * exp(-0.5 * x * x); }

/* src/Distance.hpp:77-88 and :130-140: sum of squared differences, then sqrt. */
static double distance_d(const double* xi, const double* xj, int p) {
    double sum = 0.0;
    for (int k = 0; k < p; ++k) {
        const double diff = xi[k] - xj[k];
        sum += diff * diff;
    }
    return sqrt(sum);
}

static float distance_f(const float* xi, const float* xj, int p) {
    float sum = 0.0f;
    for (int k = 0; k < p; ++k) {
        const float diff = xi[k] - xj[k];
        sum += diff * diff;
    }
    return sqrtf(sum);
}

double mds_oracle_rows_sum_of_increments_f64(const double* obs, const double* x, int64_t n, int p,
                                             double precision, int truncated, int64_t rows);
void mds_oracle_rows_gradient_f64(const double* obs, const double* x, int64_t n, int p,
                                  double precision, int truncated, int64_t rows, double* gradient);

/*
 * FP64 getSumOfIncrements() on a freshly computed state.
 * src/NewMultiDimensionalScaling.hpp:524-561: row sums over j (left to
 * right), rows summed in order, the total halved; :198-202: truncated
 * returns the halved total, non-truncated returns 0.5 * precision * it.
 */
double mds_oracle_sum_of_increments_f64(const double* obs, const double* x,
                                        int64_t n, int p, double precision,
                                        int truncated) {
    return mds_oracle_rows_sum_of_increments_f64(obs, x, n, p, precision, truncated, n);
}

/* The same over the first `rows` rows only: a bounded sample of the serial
 * engine's work, for timing.  As in mds_oracle_simd_sum_of_increments, a
 * sample (rows < n) returns the partial, unhalved and unscaled, sum. */
double mds_oracle_rows_sum_of_increments_f64(const double* obs, const double* x,
                                             int64_t n, int p, double precision,
                                             int truncated, int64_t rows) {
    const double scale = 0.5 * precision;
    const double oneOverSd = sqrt(precision);
    const int64_t m = (rows > 0 && rows < n) ? rows : n;
    double total = 0.0;
    for (int64_t i = 0; i < m; ++i) {
        double rowSum = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            const double observation = obs[i * n + j];
            if (i == j || isnan(observation)) continue; /* :402-405 */
            const double distance = distance_d(x + i * p, x + j * p, p);
            const double residual = observation - distance;
            double squared = residual * residual;
            if (truncated) {
                squared *= scale;
                squared += phi_new_d(distance * oneOverSd);
            }
            rowSum += squared;
        }
        total += rowSum;
    }
    if (m < n) return total;
    total /= 2.0;
    return truncated ? total : 0.5 * precision * total;
}

/* src/NewMultiDimensionalScaling.hpp:304-334 with :421-461 (double). */
void mds_oracle_gradient_f64(const double* obs, const double* x, int64_t n, int p,
                             double precision, int truncated, double* gradient) {
    mds_oracle_rows_gradient_f64(obs, x, n, p, precision, truncated, n, gradient);
}

/* Gradient rows [0, rows) only (rows <= 0: all), for timing. */
void mds_oracle_rows_gradient_f64(const double* obs, const double* x, int64_t n, int p,
                                  double precision, int truncated, int64_t rows,
                                  double* gradient) {
    const double scale = precision;
    const double sqrtScale = sqrt(scale);
    const int64_t m = (rows > 0 && rows < n) ? rows : n;
    memset(gradient, 0, sizeof(double) * (size_t)(n * p));
    for (int64_t i = 0; i < m; ++i) {
        for (int64_t j = 0; j < n; ++j) {
            const double observation = obs[i * n + j];
            if (i == j || isnan(observation)) continue;
            const double distance = distance_d(x + i * p, x + j * p, p);
            double residual = observation - distance;
            if (truncated) {
                residual -= pdf_new_d(distance * sqrtScale) /
                            (exp(phi_new_d(distance * sqrtScale)) * sqrtScale);
            }
            const double contribution = residual * scale / distance;
            for (int d = 0; d < p; ++d) {
                gradient[i * p + d] += contribution * (x[i * p + d] - x[j * p + d]);
            }
        }
    }
}

/*
 * FP32 engine (FloatNoSimdTypeInfo).  Observations and locations are stored
 * as float (src/MemoryManagement.hpp:45-56); the likelihood's scale is a
 * float, oneOverSd stays double (src/NewMultiDimensionalScaling.hpp:750),
 * phi_new is evaluated in double and added to a float increment, row and
 * grand sums are float, the halving is double (:554-557).
 */
double mds_oracle_sum_of_increments_f32(const double* obs, const double* x,
                                        int64_t n, int p, double precision,
                                        int truncated) {
    const float scale = (float)(0.5 * precision);
    const double oneOverSd = sqrt(precision);
    float* xf = (float*)malloc(sizeof(float) * (size_t)(n * p));
    for (int64_t k = 0; k < n * p; ++k) xf[k] = (float)x[k];
    float total = 0.0f;
    for (int64_t i = 0; i < n; ++i) {
        float rowSum = 0.0f;
        for (int64_t j = 0; j < n; ++j) {
            const double o = obs[i * n + j];
            const float observation = isnan(o) ? NAN : (float)o;
            if (i == j || isnan(observation)) continue;
            const float distance = distance_f(xf + i * p, xf + j * p, p);
            const float residual = observation - distance;
            float squared = residual * residual;
            if (truncated) {
                squared *= scale;
                squared = (float)((double)squared + phi_new_d((double)distance * oneOverSd));
            }
            rowSum += squared;
        }
        total += rowSum;
    }
    free(xf);
    double halved = (double)total;
    halved /= 2.0;
    return truncated ? halved : 0.5 * precision * halved;
}

/* FP32 gradient: float scale and sqrt, phi/pdf in double rounded to float. */
void mds_oracle_gradient_f32(const double* obs, const double* x, int64_t n, int p,
                             double precision, int truncated, double* gradient) {
    const float scale = (float)precision;
    const float sqrtScale = sqrtf(scale);
    float* xf = (float*)malloc(sizeof(float) * (size_t)(n * p));
    float* g = (float*)calloc((size_t)(n * p), sizeof(float));
    for (int64_t k = 0; k < n * p; ++k) xf[k] = (float)x[k];
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t j = 0; j < n; ++j) {
            const double o = obs[i * n + j];
            const float observation = isnan(o) ? NAN : (float)o;
            if (i == j || isnan(observation)) continue;
            const float distance = distance_f(xf + i * p, xf + j * p, p);
            float residual = observation - distance;
            if (truncated) {
                const float z = distance * sqrtScale;
                const float pdf = (float)pdf_new_d((double)z);
                const float logPhi = (float)phi_new_d((double)z);
                residual -= pdf / (expf(logPhi) * sqrtScale);
            }
            const float contribution = residual * scale / distance;
            for (int d = 0; d < p; ++d) {
                g[i * p + d] += contribution * (xf[i * p + d] - xf[j * p + d]);
            }
        }
    }
    for (int64_t k = 0; k < n * p; ++k) gradient[k] = (double)g[k];
    free(g);
    free(xf);
}

/*
 * R-level log-likelihood, R/mds.R:24-28:
 *   0.5 * (log(tau) - log(2 pi)) * N (N - 1) / 2 - sumOfIncrements
 * (the count includes missing pairs, as in the reference).
 */
double mds_oracle_log_likelihood(double sumOfIncrements, int64_t n, double precision) {
    const double observationCount = (double)(n * (n - 1)) / 2.0;
    return 0.5 * (log(precision) - log(2.0 * 3.14159265358979323846264338328)) * observationCount - sumOfIncrements;
}
