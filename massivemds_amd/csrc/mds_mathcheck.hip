// Host build of the kernels' fused truncation arithmetic (mds_trunc_math.hpp)
// so that its accuracy can be checked on the CPU against high-precision
// values.  The hardware reciprocal-square-root seeds are replaced by seeds of
// the same accuracy (FP64: the correctly rounded value cut to 26 significant
// bits, ~2^-26 like v_rsq_f64; FP32: float operations); the Newton step and
// the log Phi tables are the ones the device runs.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "mds_trunc_math.hpp"

namespace {
double cut26(double v) {
    uint64_t u;
    std::memcpy(&u, &v, sizeof u);
    u &= ~((uint64_t(1) << 26) - 1);  // keep 26 of the 52 stored fraction bits
    std::memcpy(&v, &u, sizeof v);
    return v;
}

struct SoftSeeds {
    static double rsq(double x) { return cut26(1.0 / std::sqrt(x)); }
    static float rsq(float x) { return 1.0f / std::sqrt(x); }
    static double fract(double x) { return x - std::floor(x); }
    static float fract(float x) { return x - std::floor(x); }
    static int mul24(int a, int b) { return a * b; }
};
}  // namespace

extern "C" {

// For each squared distance d2[k]: d, 1/d, log Phi(d sqrt tau) and
// phi / (Phi sqrt tau), as the FP64 kernels compute them.
void mds_mathcheck_f64(const double* d2, long long n, double tau, double* d, double* invd, double* logPhi,
                       double* gradTerm) {
    const double sqrtTau = std::sqrt(tau);
    const double sT = mdship::coeffs::kLnPhiTabStep * sqrtTau;
    const double gScale = mdship::coeffs::kLnPhiTabStep / sqrtTau;
    for (long long k = 0; k < n; ++k) {
        mdship::distance64<SoftSeeds, 1>(d2[k], d[k], invd[k]);
        const mdship::Trunc<double> t = mdship::trunc64_table<SoftSeeds>(d[k], sT, mdship::coeffs::kLnPhiTab);
        logPhi[k] = t.logPhi;
        gradTerm[k] = t.gradTerm * gScale;
    }
}

// FP32 forms, results widened to double.
void mds_mathcheck_f32(const double* d2, long long n, double tau, double* d, double* invd, double* logPhi,
                       double* gradTerm) {
    const float sT = (float)(mdship::coeffs::kLnPhiTab32Step * std::sqrt(tau));
    const double gScale = mdship::coeffs::kLnPhiTab32Step / std::sqrt(tau);
    for (long long k = 0; k < n; ++k) {
        float df, idf;
        const float d2f = (float)d2[k];
        mdship::distance32<SoftSeeds>(d2f, df, idf);
        const mdship::Trunc<float> t = mdship::trunc32_table<SoftSeeds>(df, sT, mdship::coeffs::kLnPhiTab32);
        d[k] = df;
        invd[k] = idf;
        logPhi[k] = t.logPhi;
        gradTerm[k] = t.gradTerm * gScale;
    }
}
}
