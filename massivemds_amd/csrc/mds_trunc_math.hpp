// Fused per-pair arithmetic of the truncated-normal MDS terms.
//
// FP64 (trunc64_table): log Phi(z) from a table of degree-5 polynomials on
// intervals of z of width 1/64 (held in LDS by the kernels), and
// phi(z) / Phi(z) as the same polynomial's derivative, since
// d/dz log Phi = phi / Phi; the fit is joint on both, to ~1e-16 (log Phi)
// and ~7e-15 (phi / Phi) absolute.
//
// The rational form below (trunc64, trunc32) needs no table; the FP32
// kernels use it, and the host check keeps the FP64 one as a second,
// independent route to the same values.
// For a pair at squared distance d2 and precision tau, with z = d sqrt(tau):
//   d, 1/d                 from one reciprocal square root (+ Newton step)
//   e = exp(-z^2 / 2)      own range-reduced polynomial (2^k scaling)
//   Q(z) = phi(z) m(z)     upper tail through the Mills ratio m = P / A
//   log Phi(z) = log(1 - Q) = 2 atanh(s),  s = -Q / (2 - Q)
//   phi(z) / (Phi(z) sqrt(tau))
// All divisions share one reciprocal: with E = phi * P, B = A - E = Phi A and
// C = A + B = (2 - Q) A, R = 1 / (B C) gives 1/B = C R and 1/C = B R.
// Coefficients come from gen_trunc_coeffs.py.  The reference computes the
// same quantities as log(erfc(-z / sqrt 2) / 2) and
// pdf(z) / (exp(log Phi) sqrt(tau)) (src/CDF.h:583-616,
// src/NewMultiDimensionalScaling.hpp:435-444).
//
// The `Seeds` policy supplies the initial reciprocal / reciprocal square
// root: the gfx950 instructions on the device, a float-precision stand-in
// for the host build that checks the algorithm on the CPU.
#pragma once

#include <hip/hip_runtime.h>

#include "mds_trunc_coeffs.h"

namespace mdship {

constexpr double kInvSqrt2Pi = 0.398942280401432677939946059934;  // 1 / sqrt(2 pi)
constexpr double kSqrtHalf = 0.707106781186547524400844362105;    // 1 / sqrt(2)

struct HwSeeds {
    __device__ static double rcp(double x) { return __builtin_amdgcn_rcp(x); }
    __device__ static double rsq(double x) { return __builtin_amdgcn_rsq(x); }
    __device__ static float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    __device__ static float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
    __device__ static float exp2(float x) { return __builtin_amdgcn_exp2f(x); }
};

template <int N, typename T>
__host__ __device__ inline T horner(const T (&c)[N], T x) {
    T v = c[N - 1];
#pragma unroll
    for (int k = N - 2; k >= 0; --k) v = fma(v, x, c[k]);
    return v;
}

// d = sqrt(d2) and 1/d.  d2 = 0 gives d = 0 and 1/d = NaN (there the
// reference's gradient term (r / d) (x_i - x_j) is inf * 0 = NaN as well, and
// masked pairs are zeroed after); a NaN d2 gives NaN for both.  The seed's relative error e0 (v_rsq_f64: ~1.7e-8) becomes
// ~1.5 e0^2 after NEWTON = 1 step (~5e-16) and rounding level after 2.
template <class S, int NEWTON = 1>
__host__ __device__ inline void distance64(double d2, double& d, double& invd) {
    double y = S::rsq(d2);
    const double h = 0.5 * d2;
#pragma unroll
    for (int it = 0; it < NEWTON; ++it) {
        const double hy = h * y;
        y = fma(y, fma(-hy, y, 0.5), y);
    }
    invd = y;
    d = d2 == 0.0 ? 0.0 : d2 * y;
}

template <class S>
__host__ __device__ inline void distance32(float d2, float& d, float& invd) {
    const float y = S::rsq(d2);
    invd = d2 == 0.0f ? __builtin_inff() : y;
    d = d2 == 0.0f ? 0.0f : d2 * y;
}

// exp(-w) for w >= 0.
__host__ __device__ inline double exp_neg64(double w) {
    w = fmin(w, 1000.0);
    const double k = rint(w * coeffs::kLog2e);
    double r = fma(-k, coeffs::kLn2Hi, w);
    r = fma(-k, coeffs::kLn2Lo, r);
    return ldexp(horner(coeffs::kExpNeg64, r), -(int)k);
}

template <typename Real>
struct Trunc {
    Real logPhi;    // log Phi(z)
    Real gradTerm;  // phi(z) / (Phi(z) sqrt(tau))
};

// halfTau = tau / 2, cg = 1 / (sqrt(2 pi) sqrt(tau)).
template <class S>
__host__ __device__ inline Trunc<double> trunc64(double d, double d2, double halfTau, double sqrtTau,
                                                 double cg) {
    const double e = exp_neg64(halfTau * d2);
    const double u = fmin(d * sqrtTau, coeffs::kMillsZMax) * coeffs::kMillsScale;
    const double pm = horner(coeffs::kMillsP64, u);
    const double A = horner(coeffs::kMillsQ64, u);
    const double E = e * kInvSqrt2Pi * pm;  // Q A
    const double B = A - E;                 // Phi A
    const double C = A + B;                 // (2 - Q) A
    const double bc = B * C;
    double R = S::rcp(bc);
#pragma unroll
    for (int it = 0; it < 2; ++it) R = fma(R, fma(-bc, R, 1.0), R);
    const double invB = C * R;
    const double invC = B * R;
    const double s = -E * invC;             // -Q / (2 - Q)
    Trunc<double> t;
    t.logPhi = 2.0 * s * horner(coeffs::kAtanh64, s * s);
    t.gradTerm = e * cg * A * invB;
    return t;
}

// Table form.  sT = kLnPhiTabStep sqrt(tau) (z in table units, zT = sT d);
// tab: kLnPhiTabRows rows of kLnPhiTabStride values, 16-byte aligned (LDS on
// the device), each row read as 16-byte pairs.  Returns log Phi(z) and, as
// gradTerm, its derivative in table units, d log Phi / d zT =
// phi(z) / (kLnPhiTabStep Phi(z)).  zT is clamped to the zero row, which also
// catches inf and NaN distances.
__host__ __device__ inline Trunc<double> trunc64_table(double d, double sT, const double* __restrict__ tab) {
    constexpr int kPairs = coeffs::kLnPhiTabStride / 2;
    static_assert(coeffs::kLnPhiTabStride % 2 == 0 && coeffs::kLnPhiTabDeg < coeffs::kLnPhiTabStride,
                  "rows are read as 16-byte pairs");
    const double zT = fmin(d * sT, coeffs::kLnPhiTabClamp);
    const double fl = floor(zT);
    const double t = zT - fl;
    const double2* row = reinterpret_cast<const double2*>(tab) + (int)(fl * kPairs);  // exact: fl < 2^20
    double c[coeffs::kLnPhiTabStride];
#pragma unroll
    for (int h = 0; h < kPairs; ++h) {
        const double2 v = row[h];
        c[2 * h] = v.x;
        c[2 * h + 1] = v.y;
    }
    double p = c[coeffs::kLnPhiTabDeg];
    double q = 0.0;
#pragma unroll
    for (int j = coeffs::kLnPhiTabDeg - 1; j >= 0; --j) {
        q = fma(q, t, p);
        p = fma(p, t, c[j]);
    }
    Trunc<double> r;
    r.logPhi = p;
    r.gradTerm = q;
    return r;
}

template <class S>
__host__ __device__ inline Trunc<float> trunc32(float d, float d2, float halfTau, float sqrtTau, float cg) {
    const float e = S::exp2(-halfTau * d2 * (float)coeffs::kLog2e);
    const float u = fminf(d * sqrtTau, (float)coeffs::kMillsZMax) * (float)coeffs::kMillsScale;
    const float pm = horner(coeffs::kMillsP32, u);
    const float A = horner(coeffs::kMillsQ32, u);
    const float E = e * (float)kInvSqrt2Pi * pm;
    const float B = A - E;
    const float C = A + B;
    const float R = S::rcp(B * C);
    const float invB = C * R;
    const float invC = B * R;
    const float s = -E * invC;
    Trunc<float> t;
    t.logPhi = 2.0f * s * horner(coeffs::kAtanh32, s * s);
    t.gradTerm = e * cg * A * invB;
    return t;
}

}  // namespace mdship
