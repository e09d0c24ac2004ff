// Device-side building blocks for the gfx950 MDS kernels: wavefront
// reductions and rotations on DPP (no LDS, no atomics) and blocked
// accumulators.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mds_internal.hpp"
#include "mds_trunc_math.hpp"

// Floating-point contraction only within one expression: whether a product
// is fused into a later sum must not depend on how a function was inlined
// (the kernels' results are bitwise independent of the launch shape, and
// different instantiations share this code).  Fused products the kernels
// rely on are written as fma().
#pragma clang fp contract(on)

namespace mdship {

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

// Lane permutations that write every lane (full-wave rotations): no lane
// keeps its old value, so none needs initialising.
template <int CTRL>
__device__ __forceinline__ float dpp_rotate(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}

template <int CTRL>
__device__ __forceinline__ double dpp_rotate(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, false);
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

// Per-lane accumulators: plain FP64, or FP32 group sums flushed into an FP64
// sum whose value leaves the lane in FP64.
template <typename Real>
struct LaneSum;

// add(v) adds a value; add_product(a, b) adds a * b with one rounding (an
// explicit fma, so that the result never depends on whether the compiler
// fuses a product into the sum).  flush() closes a group of steps; value()
// is read after a flush.
//
// FP64: a plain sum.  FP32: blocked summation: within a group the terms go
// into a plain FP32 group sum g, and each flush adds g, widened exactly, into
// an FP64 sum s (two instructions per group; s takes the two registers a
// compensated FP32 pair would, and its rounding error is far below the group
// sums').
template <>
struct LaneSum<double> {
    double s = 0.0;
    __device__ __forceinline__ void add(double v) { s += v; }
    __device__ __forceinline__ void add_product(double a, double b) { s = fma(a, b, s); }
    __device__ __forceinline__ void flush() {}
    __device__ __forceinline__ double value() const { return s; }
};

template <>
struct LaneSum<float> {
    double s = 0.0;
    float g = 0.0f;
    __device__ __forceinline__ void add(float v) { g += v; }
    __device__ __forceinline__ void add_product(float a, float b) { g = fmaf(a, b, g); }
    __device__ __forceinline__ void flush() {
        s += (double)g;
        g = 0.0f;
    }
    __device__ __forceinline__ double value() const { return s; }
};

__device__ __forceinline__ bool m_isnan(double v) { return v != v; }
__device__ __forceinline__ bool m_isnan(float v) { return v != v; }

// Column accumulators of the skewed tile schedule.  At step k lane l holds
// the partial sum of tile column (l + k) mod 64; rotate() moves every
// accumulator one lane down (DPP wave_rol:1, lane l <- lane l + 1) so that
// at step k + 1 lane l holds column (l + k + 1) mod 64.  After 64 steps lane
// l holds column l.  FP32 rotates only the group sum g each step; the FP64
// sum stays put during a group of STEPS steps and moves STEPS lanes at the
// flush (one ds_bpermute per 32-bit half), landing on g's columns.
constexpr int kDppWaveRol1 = 0x134;

template <typename Real>
struct ColSum;

template <>
struct ColSum<double> {
    double s = 0.0;
    __device__ __forceinline__ void add(double v) { s += v; }
    __device__ __forceinline__ void add_product(double a, double b) { s = fma(a, b, s); }
    __device__ __forceinline__ void rotate() { s = dpp_rotate<kDppWaveRol1>(s); }
    template <int STEPS>
    __device__ __forceinline__ void flush() {}
    __device__ __forceinline__ double value() const { return s; }
};

// lane l receives v from lane (l + SHIFT) mod 64 (LDS crossbar, no memory)
template <int SHIFT>
__device__ __forceinline__ float lane_rotate(float v) {
    const int src = ((int)(threadIdx.x % kWave) + SHIFT) & (kWave - 1);
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}

template <int SHIFT>
__device__ __forceinline__ double lane_rotate(double v) {
    const int src = (((int)(threadIdx.x % kWave) + SHIFT) & (kWave - 1)) << 2;
    const int lo = __builtin_amdgcn_ds_bpermute(src, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(src, __double2hiint(v));
    return __hiloint2double(hi, lo);
}

template <>
struct ColSum<float> {
    double s = 0.0;
    float g = 0.0f;
    __device__ __forceinline__ void add(float v) { g += v; }
    __device__ __forceinline__ void add_product(float a, float b) { g = fmaf(a, b, g); }
    __device__ __forceinline__ void rotate() { g = dpp_rotate<kDppWaveRol1>(g); }
    template <int STEPS>
    __device__ __forceinline__ void flush() {
        s = lane_rotate<STEPS>(s) + (double)g;
        g = 0.0f;
    }
    __device__ __forceinline__ double value() const { return s; }
};

}  // namespace mdship
