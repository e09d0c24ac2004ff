/*
 * The oracle (serial FP64/FP32, row samples, SIMD + OpenMP) run under
 * AddressSanitizer and UndefinedBehaviorSanitizer on small cases with the
 * conditions its callers can produce: N = 1 and 2, ragged N, P from 1 to 16,
 * missing observations, an all-missing row, coincident points, truncation on
 * and off.  Besides memory and UB checks it asserts the invariants that hold
 * for any input: the serial and SIMD forms agree, the gradient rows of an
 * isolated location are zero, and the gradient sums to zero over locations
 * for symmetric data (translation invariance).  Exit status 0 = clean.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

double mds_oracle_sum_of_increments_f64(const double*, const double*, int64_t, int, double, int);
void mds_oracle_gradient_f64(const double*, const double*, int64_t, int, double, int, double*);
double mds_oracle_rows_sum_of_increments_f64(const double*, const double*, int64_t, int, double, int, int64_t);
void mds_oracle_rows_gradient_f64(const double*, const double*, int64_t, int, double, int, int64_t, double*);
double mds_oracle_sum_of_increments_f32(const double*, const double*, int64_t, int, double, int);
void mds_oracle_gradient_f32(const double*, const double*, int64_t, int, double, int, double*);
double mds_oracle_simd_sum_of_increments(const double*, const double*, int64_t, int, double, int, int, int64_t);
void mds_oracle_simd_gradient(const double*, const double*, int64_t, int, double, int, int, int64_t, double*);
double mds_oracle_log_likelihood(double, int64_t, double);

static uint64_t state = 88172645463325252ULL;
static double uniform01(void) {
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return (double)(state >> 11) * 0x1.0p-53;
}

static int failures = 0;
#define EXPECT(cond, ...)                      \
    do {                                       \
        if (!(cond)) {                         \
            fprintf(stderr, __VA_ARGS__);      \
            fputc('\n', stderr);               \
            ++failures;                        \
        }                                      \
    } while (0)

static void run_case(int64_t n, int p, int truncated, int missingRow, int coincident) {
    double* obs = malloc(sizeof(double) * (size_t)(n * n));
    double* x = malloc(sizeof(double) * (size_t)(n * p));
    double* g = malloc(sizeof(double) * (size_t)(n * p));
    double* g2 = malloc(sizeof(double) * (size_t)(n * p));
    for (int64_t k = 0; k < n * p; ++k) x[k] = 2.0 * uniform01() - 1.0;
    if (coincident && n > 2)
        for (int d = 0; d < p; ++d) x[2 * p + d] = x[1 * p + d];
    for (int64_t i = 0; i < n; ++i) {
        obs[i * n + i] = 0.0;
        for (int64_t j = i + 1; j < n; ++j) {
            const double v = uniform01() < 0.07 ? NAN : 3.0 * uniform01();
            obs[i * n + j] = obs[j * n + i] = v;
        }
    }
    if (missingRow >= 0 && missingRow < n)
        for (int64_t j = 0; j < n; ++j) obs[missingRow * n + j] = obs[j * n + missingRow] = NAN;
    if (coincident && n > 2) obs[1 * n + 2] = obs[2 * n + 1] = NAN;  /* keep the zero distance masked */
    const double tau = 1.7;
    const double s = mds_oracle_sum_of_increments_f64(obs, x, n, p, tau, truncated);
    const double sr = mds_oracle_rows_sum_of_increments_f64(obs, x, n, p, tau, truncated, n);
    const double ss = mds_oracle_simd_sum_of_increments(obs, x, n, p, tau, truncated, 2, 0);
    const double sf = mds_oracle_sum_of_increments_f32(obs, x, n, p, tau, truncated);
    const double ll = mds_oracle_log_likelihood(s, n, tau);
    EXPECT(isfinite(s) && isfinite(ll), "n=%lld p=%d: sum %g", (long long)n, p, s);
    EXPECT(s == sr, "n=%lld p=%d: row form %g != %g", (long long)n, p, sr, s);
    EXPECT(fabs(ss - s) <= 1e-12 * fabs(s) + 1e-300, "n=%lld p=%d: simd %g != %g", (long long)n, p, ss, s);
    EXPECT(fabs(sf - s) <= 1e-5 * fabs(s) + 1e-30, "n=%lld p=%d: fp32 %g != %g", (long long)n, p, sf, s);
    mds_oracle_gradient_f64(obs, x, n, p, tau, truncated, g);
    mds_oracle_simd_gradient(obs, x, n, p, tau, truncated, 2, 0, g2);
    double gmax = 0.0, diff = 0.0;
    for (int64_t k = 0; k < n * p; ++k) {
        EXPECT(isfinite(g[k]), "n=%lld p=%d: gradient entry %lld not finite", (long long)n, p, (long long)k);
        gmax = fmax(gmax, fabs(g[k]));
        diff = fmax(diff, fabs(g[k] - g2[k]));
    }
    EXPECT(diff <= 1e-12 * gmax + 1e-300, "n=%lld p=%d: simd gradient differs by %g", (long long)n, p, diff);
    for (int d = 0; d < p; ++d) {
        double col = 0.0;
        for (int64_t i = 0; i < n; ++i) col += g[i * p + d];
        EXPECT(fabs(col) <= 1e-12 * (gmax + 1.0) * (double)n, "n=%lld p=%d: sum of gradient %g", (long long)n, p,
               col);
    }
    if (missingRow >= 0 && missingRow < n)
        for (int d = 0; d < p; ++d)
            EXPECT(g[missingRow * p + d] == 0.0, "n=%lld p=%d: isolated row has gradient", (long long)n, p);
    mds_oracle_gradient_f32(obs, x, n, p, tau, truncated, g2);
    mds_oracle_rows_gradient_f64(obs, x, n, p, tau, truncated, n > 3 ? 3 : n, g2);
    free(obs);
    free(x);
    free(g);
    free(g2);
}

int main(void) {
    const int64_t sizes[] = {1, 2, 3, 17, 64, 65, 130};
    const int dims[] = {1, 2, 3, 5, 8, 16};
    for (size_t a = 0; a < sizeof sizes / sizeof sizes[0]; ++a)
        for (size_t b = 0; b < sizeof dims / sizeof dims[0]; ++b)
            for (int t = 0; t < 2; ++t) run_case(sizes[a], dims[b], t, (int)(sizes[a] / 2), sizes[a] > 3);
    if (failures) fprintf(stderr, "%d failures\n", failures);
    return failures ? 1 : 0;
}
