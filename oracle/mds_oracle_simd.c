/*
 * mds_oracle_simd.c -- vectorised, multi-core CPU restatement used as the
 * CPU BASELINE that bench.py times beside the GPU engine.
 *
 * TEST INFRASTRUCTURE ONLY (see mds_oracle.c): never linked by the product.
 *
 * It stands in for the reference's "AVX + TBB" engines
 * (NewMultiDimensionalScaling<DoubleAvxTypeInfo|DoubleAvx512TypeInfo,
 * TbbAccumulate>, src/NewMultiDimensionalScaling.hpp:840-869), which cannot
 * be compiled here (xsimd 7.1.3 and the TBB headers are not in the image).
 * The structure follows the reference SIMD path:
 *   - rows i are distributed over threads (TBB parallel_reduce/parallel_for,
 *     src/NewMultiDimensionalScaling.hpp:708-744; here OpenMP, and the row
 *     sums are added in row order so the result is deterministic);
 *   - j runs in SIMD batches of W = 8 (AVX-512) or 4 (AVX2) lanes, with the
 *     distances computed per lane and packed (src/Distance.hpp:158-359) and
 *     the transcendentals vectorised (glibc libmvec erfc/log/exp standing in
 *     for xsimd's), then a scalar remainder loop (:321-332, :532-548);
 *   - likelihood and gradient are separate passes over the matrix, as in the
 *     reference (two API calls per evaluation).
 */
#include <immintrin.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ORACLE_1_SQRT_2PI 0.398942280401432677939946059934
#define ORACLE_SQRT1_2 0.707106781186547524400844362105

/* glibc libmvec vector variants (exported names, GLIBC_2.22 / 2.35). */
__m256d _ZGVdN4v_erfc(__m256d);
__m256d _ZGVdN4v_log(__m256d);
__m256d _ZGVdN4v_exp(__m256d);
__m512d _ZGVeN8v_erfc(__m512d);
__m512d _ZGVeN8v_log(__m512d);
__m512d _ZGVeN8v_exp(__m512d);

static double phi_new_d(double x) { return log(0.5 * erfc(-x * ORACLE_SQRT1_2)); }
static double pdf_new_d(double x) { return ORACLE_1_SQRT_2PI * exp(-0.5 * x * x); }

static inline double dist_scalar(const double* xi, const double* xj, int p) {
    double s = 0.0;
    for (int k = 0; k < p; ++k) {
        const double d = xi[k] - xj[k];
        s += d * d;
    }
    return sqrt(s);
}

/* Scalar remainder (SimdSize = 1) of the likelihood row loop. */
static double lik_tail(const double* obs, const double* x, int64_t n, int p, int64_t i,
                       int64_t j0, double scale, double oneOverSd, int truncated) {
    double sum = 0.0;
    for (int64_t j = j0; j < n; ++j) {
        const double o = obs[i * n + j];
        if (i == j || isnan(o)) continue;
        const double d = dist_scalar(x + i * p, x + j * p, p);
        double sq = (o - d) * (o - d);
        if (truncated) sq = sq * scale + phi_new_d(d * oneOverSd);
        sum += sq;
    }
    return sum;
}

static void grad_tail(const double* obs, const double* x, int64_t n, int p, int64_t i,
                      int64_t j0, double scale, double sqrtScale, int truncated, double* g) {
    for (int64_t j = j0; j < n; ++j) {
        const double o = obs[i * n + j];
        if (i == j || isnan(o)) continue;
        const double d = dist_scalar(x + i * p, x + j * p, p);
        double r = o - d;
        if (truncated)
            r -= pdf_new_d(d * sqrtScale) / (exp(phi_new_d(d * sqrtScale)) * sqrtScale);
        const double c = r * scale / d;
        for (int k = 0; k < p; ++k) g[k] += c * (x[i * p + k] - x[j * p + k]);
    }
}

/* ---------------- AVX-512 (W = 8) ---------------- */

__attribute__((target("avx512f,avx512dq")))
static double lik_row_avx512(const double* obs, const double* x, int64_t n, int p, int64_t i,
                             double scale, double oneOverSd, int truncated) {
    const int64_t vec = n - n % 8;
    __m512d acc = _mm512_setzero_pd();
    const __m512d vscale = _mm512_set1_pd(scale);
    const __m512d vsd = _mm512_set1_pd(oneOverSd);
    const __m512d vhalf = _mm512_set1_pd(0.5);
    const __m512d vmsqrt = _mm512_set1_pd(-ORACLE_SQRT1_2);
    double dbuf[8] __attribute__((aligned(64)));
    for (int64_t j = 0; j < vec; j += 8) {
        for (int l = 0; l < 8; ++l) dbuf[l] = dist_scalar(x + i * p, x + (j + l) * p, p);
        const __m512d d = _mm512_load_pd(dbuf);
        const __m512d o = _mm512_loadu_pd(obs + i * n + j);
        __mmask8 valid = _mm512_cmp_pd_mask(o, o, _CMP_ORD_Q);
        if (i >= j && i < j + 8) valid &= (__mmask8)~(1u << (i - j));
        if (!valid) continue;
        const __m512d r = _mm512_sub_pd(o, d);
        __m512d sq = _mm512_mul_pd(r, r);
        if (truncated) {
            const __m512d z = _mm512_mul_pd(d, vsd);
            const __m512d lphi = _ZGVeN8v_log(_mm512_mul_pd(vhalf, _ZGVeN8v_erfc(_mm512_mul_pd(z, vmsqrt))));
            sq = _mm512_add_pd(_mm512_mul_pd(sq, vscale), lphi);
        }
        acc = _mm512_add_pd(acc, _mm512_maskz_mov_pd(valid, sq));
    }
    double sum = _mm512_reduce_add_pd(acc);
    if (vec < n) sum += lik_tail(obs, x, n, p, i, vec, scale, oneOverSd, truncated);
    return sum;
}

__attribute__((target("avx512f,avx512dq")))
static void grad_row_avx512(const double* obs, const double* x, int64_t n, int p, int64_t i,
                            double scale, double sqrtScale, int truncated, double* g) {
    const int64_t vec = n - n % 8;
    const __m512d vscale = _mm512_set1_pd(scale);
    const __m512d vss = _mm512_set1_pd(sqrtScale);
    const __m512d vhalf = _mm512_set1_pd(0.5);
    const __m512d vmhalf = _mm512_set1_pd(-0.5);
    const __m512d vmsqrt = _mm512_set1_pd(-ORACLE_SQRT1_2);
    const __m512d vpdf = _mm512_set1_pd(ORACLE_1_SQRT_2PI);
    double dbuf[8] __attribute__((aligned(64)));
    double cbuf[8] __attribute__((aligned(64)));
    for (int k = 0; k < p; ++k) g[k] = 0.0;
    for (int64_t j = 0; j < vec; j += 8) {
        for (int l = 0; l < 8; ++l) dbuf[l] = dist_scalar(x + i * p, x + (j + l) * p, p);
        const __m512d d = _mm512_load_pd(dbuf);
        const __m512d o = _mm512_loadu_pd(obs + i * n + j);
        __mmask8 valid = _mm512_cmp_pd_mask(o, o, _CMP_ORD_Q);
        if (i >= j && i < j + 8) valid &= (__mmask8)~(1u << (i - j));
        if (!valid) continue;
        __m512d r = _mm512_sub_pd(o, d);
        if (truncated) {
            const __m512d z = _mm512_mul_pd(d, vss);
            const __m512d pdf = _mm512_mul_pd(vpdf, _ZGVeN8v_exp(_mm512_mul_pd(_mm512_mul_pd(vmhalf, z), z)));
            const __m512d lphi = _ZGVeN8v_log(_mm512_mul_pd(vhalf, _ZGVeN8v_erfc(_mm512_mul_pd(z, vmsqrt))));
            const __m512d den = _mm512_mul_pd(_ZGVeN8v_exp(lphi), vss);
            r = _mm512_sub_pd(r, _mm512_div_pd(pdf, den));
        }
        const __m512d c = _mm512_maskz_mov_pd(valid, _mm512_div_pd(_mm512_mul_pd(r, vscale), d));
        _mm512_store_pd(cbuf, c);
        for (int l = 0; l < 8; ++l)
            for (int k = 0; k < p; ++k) g[k] += cbuf[l] * (x[i * p + k] - x[(j + l) * p + k]);
    }
    if (vec < n) grad_tail(obs, x, n, p, i, vec, scale, sqrtScale, truncated, g);
}

/* ---------------- AVX2 (W = 4) ---------------- */

__attribute__((target("avx2,fma")))
static double lik_row_avx2(const double* obs, const double* x, int64_t n, int p, int64_t i,
                           double scale, double oneOverSd, int truncated) {
    const int64_t vec = n - n % 4;
    __m256d acc = _mm256_setzero_pd();
    const __m256d vscale = _mm256_set1_pd(scale);
    const __m256d vsd = _mm256_set1_pd(oneOverSd);
    const __m256d vhalf = _mm256_set1_pd(0.5);
    const __m256d vmsqrt = _mm256_set1_pd(-ORACLE_SQRT1_2);
    double dbuf[4] __attribute__((aligned(32)));
    for (int64_t j = 0; j < vec; j += 4) {
        for (int l = 0; l < 4; ++l) dbuf[l] = dist_scalar(x + i * p, x + (j + l) * p, p);
        const __m256d d = _mm256_load_pd(dbuf);
        const __m256d o = _mm256_loadu_pd(obs + i * n + j);
        __m256d valid = _mm256_cmp_pd(o, o, _CMP_ORD_Q);
        if (i >= j && i < j + 4) {
            double m[4] __attribute__((aligned(32)));
            _mm256_store_pd(m, valid);
            m[i - j] = 0.0;
            valid = _mm256_load_pd(m);
        }
        if (_mm256_movemask_pd(valid) == 0) continue;
        const __m256d r = _mm256_sub_pd(o, d);
        __m256d sq = _mm256_mul_pd(r, r);
        if (truncated) {
            const __m256d z = _mm256_mul_pd(d, vsd);
            const __m256d lphi = _ZGVdN4v_log(_mm256_mul_pd(vhalf, _ZGVdN4v_erfc(_mm256_mul_pd(z, vmsqrt))));
            sq = _mm256_add_pd(_mm256_mul_pd(sq, vscale), lphi);
        }
        acc = _mm256_add_pd(acc, _mm256_and_pd(valid, sq));
    }
    double t[4] __attribute__((aligned(32)));
    _mm256_store_pd(t, acc);
    double sum = (t[0] + t[1]) + (t[2] + t[3]);
    if (vec < n) sum += lik_tail(obs, x, n, p, i, vec, scale, oneOverSd, truncated);
    return sum;
}

__attribute__((target("avx2,fma")))
static void grad_row_avx2(const double* obs, const double* x, int64_t n, int p, int64_t i,
                          double scale, double sqrtScale, int truncated, double* g) {
    const int64_t vec = n - n % 4;
    const __m256d vscale = _mm256_set1_pd(scale);
    const __m256d vss = _mm256_set1_pd(sqrtScale);
    const __m256d vhalf = _mm256_set1_pd(0.5);
    const __m256d vmhalf = _mm256_set1_pd(-0.5);
    const __m256d vmsqrt = _mm256_set1_pd(-ORACLE_SQRT1_2);
    const __m256d vpdf = _mm256_set1_pd(ORACLE_1_SQRT_2PI);
    double dbuf[4] __attribute__((aligned(32)));
    double cbuf[4] __attribute__((aligned(32)));
    for (int k = 0; k < p; ++k) g[k] = 0.0;
    for (int64_t j = 0; j < vec; j += 4) {
        for (int l = 0; l < 4; ++l) dbuf[l] = dist_scalar(x + i * p, x + (j + l) * p, p);
        const __m256d d = _mm256_load_pd(dbuf);
        const __m256d o = _mm256_loadu_pd(obs + i * n + j);
        __m256d valid = _mm256_cmp_pd(o, o, _CMP_ORD_Q);
        if (i >= j && i < j + 4) {
            double m[4] __attribute__((aligned(32)));
            _mm256_store_pd(m, valid);
            m[i - j] = 0.0;
            valid = _mm256_load_pd(m);
        }
        if (_mm256_movemask_pd(valid) == 0) continue;
        __m256d r = _mm256_sub_pd(o, d);
        if (truncated) {
            const __m256d z = _mm256_mul_pd(d, vss);
            const __m256d pdf = _mm256_mul_pd(vpdf, _ZGVdN4v_exp(_mm256_mul_pd(_mm256_mul_pd(vmhalf, z), z)));
            const __m256d lphi = _ZGVdN4v_log(_mm256_mul_pd(vhalf, _ZGVdN4v_erfc(_mm256_mul_pd(z, vmsqrt))));
            const __m256d den = _mm256_mul_pd(_ZGVdN4v_exp(lphi), vss);
            r = _mm256_sub_pd(r, _mm256_div_pd(pdf, den));
        }
        const __m256d c = _mm256_and_pd(valid, _mm256_div_pd(_mm256_mul_pd(r, vscale), d));
        _mm256_store_pd(cbuf, c);
        for (int l = 0; l < 4; ++l)
            for (int k = 0; k < p; ++k) g[k] += cbuf[l] * (x[i * p + k] - x[(j + l) * p + k]);
    }
    if (vec < n) grad_tail(obs, x, n, p, i, vec, scale, sqrtScale, truncated, g);
}

/* Returns 512, 256 or 0: the SIMD width this host will use. */
int mds_oracle_simd_width(void) {
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq")) return 512;
    if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) return 256;
    return 0;
}

/*
 * getSumOfIncrements() of the SIMD + thread-parallel engine.  `threads` <= 0
 * means all cores (omp_get_max_threads()).  `rows` limits the evaluation to
 * the first `rows` rows (a bounded sample for timing; rows <= 0 means all),
 * in which case the return value is the partial (unhalved, unscaled) sum.
 */
double mds_oracle_simd_sum_of_increments(const double* obs, const double* x, int64_t n, int p,
                                         double precision, int truncated, int threads,
                                         int64_t rows) {
    const int width = mds_oracle_simd_width();
    const double scale = 0.5 * precision;
    const double oneOverSd = sqrt(precision);
    const int64_t m = (rows > 0 && rows < n) ? rows : n;
    double* rowSums = (double*)malloc(sizeof(double) * (size_t)m);
    const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 8) num_threads(nt)
    for (int64_t i = 0; i < m; ++i) {
        if (width == 512)
            rowSums[i] = lik_row_avx512(obs, x, n, p, i, scale, oneOverSd, truncated);
        else if (width == 256)
            rowSums[i] = lik_row_avx2(obs, x, n, p, i, scale, oneOverSd, truncated);
        else
            rowSums[i] = lik_tail(obs, x, n, p, i, 0, scale, oneOverSd, truncated);
    }
    double total = 0.0;
    for (int64_t i = 0; i < m; ++i) total += rowSums[i];
    free(rowSums);
    if (m < n) return total;
    total /= 2.0;
    return truncated ? total : 0.5 * precision * total;
}

/* getLogLikelihoodGradient() of the SIMD + thread-parallel engine (rows as above). */
void mds_oracle_simd_gradient(const double* obs, const double* x, int64_t n, int p,
                              double precision, int truncated, int threads, int64_t rows,
                              double* gradient) {
    const int width = mds_oracle_simd_width();
    const double scale = precision;
    const double sqrtScale = sqrt(precision);
    const int64_t m = (rows > 0 && rows < n) ? rows : n;
    const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 8) num_threads(nt)
    for (int64_t i = 0; i < m; ++i) {
        double* g = gradient + i * p;
        if (width == 512) {
            grad_row_avx512(obs, x, n, p, i, scale, sqrtScale, truncated, g);
        } else if (width == 256) {
            grad_row_avx2(obs, x, n, p, i, scale, sqrtScale, truncated, g);
        } else {
            for (int k = 0; k < p; ++k) g[k] = 0.0;
            grad_tail(obs, x, n, p, i, 0, scale, sqrtScale, truncated, g);
        }
    }
}

int mds_oracle_max_threads(void) { return omp_get_max_threads(); }
