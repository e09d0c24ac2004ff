// Fused per-pair arithmetic of the truncated-normal MDS terms.
//
// For a pair at squared distance d2 and precision tau, with z = d sqrt(tau),
// the likelihood needs log Phi(z) and the gradient phi(z) / Phi(z) (the
// reference computes log(erfc(-z / sqrt 2) / 2) and
// pdf(z) / (exp(log Phi) sqrt(tau)): src/CDF.h:583-616,
// src/NewMultiDimensionalScaling.hpp:435-444).  Both come from one table
// of log Phi: on each interval of z of width 1/S a polynomial in the
// position t in [0, 1) within the interval, fitted jointly to log Phi and to
// its derivative, so that phi / Phi = d log Phi / dz is the same
// polynomial's derivative (one row read gives both, with no exp, division
// or logarithm per pair).  A final zero row covers z >= 10, where
// |log Phi| < 8e-24 and phi / Phi < 8e-23.
//
//   FP64: S = 64, degree 5, six doubles per row (three 16-byte loads);
//         max |error| ~1e-16 (log Phi), ~7e-15 (phi / Phi).
//   FP32: S = 32, degree 3, four floats per row (one 16-byte load);
//         max |error| ~6e-8 (log Phi), ~1e-7 (phi / Phi).
//
// The kernels hold the tables in LDS; gen_trunc_coeffs.py fits them and
// reports the errors.  d and 1/d come from the hardware reciprocal square
// root (FP64: one Newton step).  The `Seeds` policy supplies that seed and the
// fractional part: the gfx950 instructions on the device, stand-ins (of the
// same accuracy for the seed, exact for the fractional part) for the host
// build that checks the arithmetic on the CPU.
#pragma once

#include <hip/hip_runtime.h>

#include "mds_trunc_coeffs.h"

namespace mdship {

struct HwSeeds {
    __device__ static double rsq(double x) { return __builtin_amdgcn_rsq(x); }
    __device__ static float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
    // x - floor(x), exact for finite x >= 0 (one instruction)
    __device__ static double fract(double x) { return __builtin_amdgcn_fract(x); }
    __device__ static float fract(float x) { return __builtin_amdgcn_fractf(x); }
    // a * b for 0 <= a, b < 2^23 (full-rate 24-bit multiply)
    __device__ static int mul24(int a, int b) { return __mul24(a, b); }
};

// 2^500: above 1/sqrt(d2) for every d2 > 2^-1000.
constexpr double kDistBig = 0x1p500;

// d = sqrt(d2) and 1/d.  d2 = 0 gives d = 0 and 1/d = NaN (there the
// reference's gradient term (r / d) (x_i - x_j) is inf * 0 = NaN as well, and
// masked pairs are zeroed after); a NaN d2 gives NaN for both, d2 = inf gives
// d = inf.  The seed's relative error e0 (v_rsq_f64: ~1.7e-8) becomes
// ~1.5 e0^2 after NEWTON = 1 step (~5e-16) and rounding level after 2.
// d multiplies d2 by min(1/d, 2^500) (IEEE minNum: a NaN operand yields the
// other): the NaN that the Newton step gives at d2 = 0 (0 * inf) and at
// d2 = inf becomes 2^500, so d = 0 and d = inf there, and every
// d2 > 2^-1000 keeps d = d2 / sqrt(d2) bit for bit.
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
    d = d2 * fmin(y, kDistBig);
}

// 2^64: above 1/sqrt(d2) for every normal FP32 d2.
constexpr float kDistBig32 = 0x1p64f;

// FP32: the seed alone (~1 ulp).  d2 = 0: 1/d = inf, d = 0 * 2^64 = 0;
// d2 = inf: 1/d = 0 and d = inf * 0 = NaN, as the plain product gives.
template <class S>
__host__ __device__ inline void distance32(float d2, float& d, float& invd) {
    const float y = S::rsq(d2);
    invd = y;
    d = d2 * fminf(y, kDistBig32);
}

template <typename Real>
struct Trunc {
    Real logPhi;    // log Phi(z)
    Real gradTerm;  // d log Phi / d(S z) = phi(z) / (S Phi(z))
};

// FP64.  sT = S sqrt(tau) (z in table units, zT = sT d); tab: kLnPhiTabRows
// rows of kLnPhiTabStride values, 16-byte aligned.  zT is clamped to the zero
// row, which also catches inf and NaN distances.  zT >= 0, so the row is the
// truncation of zT and t its fractional part.
template <class S = HwSeeds>
__host__ __device__ inline Trunc<double> trunc64_table(double d, double sT, const double* __restrict__ tab) {
    constexpr int kPairs = coeffs::kLnPhiTabStride / 2;
    static_assert(coeffs::kLnPhiTabStride % 2 == 0 && coeffs::kLnPhiTabDeg < coeffs::kLnPhiTabStride,
                  "rows are read as 16-byte pairs");
    const double zT = fmin(d * sT, coeffs::kLnPhiTabClamp);
    const double t = S::fract(zT);
    const double2* row = reinterpret_cast<const double2*>(tab) + S::mul24((int)zT, kPairs);
    double c[coeffs::kLnPhiTabStride];
#pragma unroll
    for (int h = 0; h < kPairs; ++h) {
        const double2 v = row[h];
        c[2 * h] = v.x;
        c[2 * h + 1] = v.y;
    }
    // value and derivative in one Horner pass
    double q = c[coeffs::kLnPhiTabDeg];
    double p = fma(q, t, c[coeffs::kLnPhiTabDeg - 1]);
#pragma unroll
    for (int j = coeffs::kLnPhiTabDeg - 2; j >= 0; --j) {
        q = fma(q, t, p);
        p = fma(p, t, c[j]);
    }
    Trunc<double> r;
    r.logPhi = p;
    r.gradTerm = q;
    return r;
}

// FP32: one 16-byte row; sT = S sqrt(tau) with S = kLnPhiTab32Step.
template <class S = HwSeeds>
__host__ __device__ inline Trunc<float> trunc32_table(float d, float sT, const float* __restrict__ tab) {
    static_assert(coeffs::kLnPhiTab32Stride == 4 && coeffs::kLnPhiTab32Deg == 3, "one 16-byte row");
    const float zT = fminf(d * sT, coeffs::kLnPhiTab32Clamp);
    const float t = S::fract(zT);
    const float4 c = reinterpret_cast<const float4*>(tab)[(int)zT];
    float p = fmaf(c.w, t, c.z);
    float q = fmaf(c.w, t, p);
    p = fmaf(p, t, c.y);
    q = fmaf(q, t, p);
    p = fmaf(p, t, c.x);
    Trunc<float> r;
    r.logPhi = p;
    r.gradTerm = q;
    return r;
}

}  // namespace mdship
