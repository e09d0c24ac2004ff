// Device-side building blocks for the gfx950 MDS kernels: wavefront
// reductions on DPP (no LDS, no atomics), compensated accumulators and the
// per-pair truncated-normal terms.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mds_internal.hpp"

namespace mdship {

constexpr double kInvSqrt2Pi = 0.398942280401432677939946059934;  // 1 / sqrt(2 pi)
constexpr double kSqrtHalf = 0.707106781186547524400844362105;    // 1 / sqrt(2)

// DPP control words (GFX9 encoding, valid on gfx950).
constexpr int kDppQuadXor1 = 0xB1;       // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;       // quad_perm [2,3,0,1]
constexpr int kDppRowMirror = 0x140;     // lane i <- lane 15-i within a row of 16
constexpr int kDppRowHalfMirror = 0x141; // lane i <- lane 7-i within 8 lanes
constexpr int kDppRowBcast15 = 0x142;    // lane 15 of row r -> row r+1
constexpr int kDppRowBcast31 = 0x143;    // lane 31 -> rows 2 and 3

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_mov(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a fully active wavefront.  Lane 63 receives the
// total (other lanes hold partial sums).  Fixed order, so bitwise
// reproducible.  Lanes masked off by the row masks receive 0 and keep their
// value.
template <typename T>
__device__ __forceinline__ T wave_sum_to_last(T v) {
    v += dpp_mov<kDppQuadXor1, 0xf>(v);
    v += dpp_mov<kDppQuadXor2, 0xf>(v);
    v += dpp_mov<kDppRowHalfMirror, 0xf>(v);
    v += dpp_mov<kDppRowMirror, 0xf>(v);
    v += dpp_mov<kDppRowBcast15, 0xa>(v);
    v += dpp_mov<kDppRowBcast31, 0xc>(v);
    return v;
}

__device__ __forceinline__ float read_last_lane(float v) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ double read_last_lane(double v) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// Per-lane accumulators: plain FP64, or Kahan-compensated FP32 whose value is
// promoted to FP64 when it leaves the lane.
template <typename Real>
struct LaneSum;

template <>
struct LaneSum<double> {
    double s = 0.0;
    __device__ __forceinline__ void add(double v) { s += v; }
    __device__ __forceinline__ double value() const { return s; }
};

template <>
struct LaneSum<float> {
    float s = 0.0f;
    float c = 0.0f;
    __device__ __forceinline__ void add(float v) {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

// Math wrappers resolving to the ocml routine of the matching precision.
__device__ __forceinline__ double m_sqrt(double v) { return sqrt(v); }
__device__ __forceinline__ float m_sqrt(float v) { return sqrtf(v); }
__device__ __forceinline__ double m_erfc(double v) { return erfc(v); }
__device__ __forceinline__ float m_erfc(float v) { return erfcf(v); }
__device__ __forceinline__ double m_log1p(double v) { return log1p(v); }
__device__ __forceinline__ float m_log1p(float v) { return log1pf(v); }
__device__ __forceinline__ double m_exp(double v) { return exp(v); }
__device__ __forceinline__ float m_exp(float v) { return expf(v); }
__device__ __forceinline__ bool m_isnan(double v) { return v != v; }
__device__ __forceinline__ bool m_isnan(float v) { return v != v; }

// Truncation terms at z = d * sqrt(tau) >= 0, with Q(z) = erfc(z / sqrt 2) / 2:
//   log Phi(z) = log1p(-Q)                              (likelihood)
//   phi(z) / (Phi(z) sqrt(tau))                         (gradient)
// The reference evaluates log(erfc(-z / sqrt 2) / 2) and recovers Phi with
// exp() of it (src/CDF.h:583-592, src/NewMultiDimensionalScaling.hpp:439-441);
// the forms here are algebraically identical, computed once per unordered
// pair and accurate to a few ulp as Phi -> 1.
template <typename Real>
struct TruncTerms {
    Real logPhi;
    Real gradTerm;
};

template <typename Real>
__device__ __forceinline__ TruncTerms<Real> trunc_terms(Real z, Real sqrtTau) {
    const Real q = Real(0.5) * m_erfc(z * Real(kSqrtHalf));
    const Real phi = Real(kInvSqrt2Pi) * m_exp(Real(-0.5) * z * z);
    TruncTerms<Real> t;
    t.logPhi = m_log1p(-q);
    t.gradTerm = phi / ((Real(1) - q) * sqrtTau);
    return t;
}

}  // namespace mdship
