// Device functions shared by the single-instance and batched kernels: the
// per-pair arithmetic and the wavefront's walk over one 64 x 64 tile.
#pragma once

#include "mds_device.hpp"

namespace mdship {

// Steps per prefetch group of the tile walk.
constexpr int kGroup = 4;

// Per-evaluation constants of the pair arithmetic: tau, tau / 2, and
// sT = S sqrt(tau), z in the units of the log Phi table (S = 64 for FP64,
// 32 for FP32), which `tab` points to in LDS.  The truncation's gradient
// term g = d log Phi / d(S z) enters the pair coefficient as
// tau (phi / Phi) / sqrt(tau) = sT g.
template <typename Real>
struct PairConst {
    Real tau, halfTau, sT;
    const Real* tab;
};

template <typename Real>
struct TableOf;
template <>
struct TableOf<double> {
    static constexpr int kStep = coeffs::kLnPhiTabStep;
    static constexpr int kCount = coeffs::kLnPhiTabRows * coeffs::kLnPhiTabStride;
    __device__ static double value(int i) { return coeffs::kLnPhiTab[i]; }
};
template <>
struct TableOf<float> {
    static constexpr int kStep = coeffs::kLnPhiTab32Step;
    static constexpr int kCount = coeffs::kLnPhiTab32Rows * coeffs::kLnPhiTab32Stride;
    __device__ static float value(int i) { return coeffs::kLnPhiTab32[i]; }
};

// Per-evaluation constants from the precision (as the host holds it, or
// read from a sampler's device parameter block), computed in FP64.
template <typename Real>
__device__ __forceinline__ PairConst<Real> make_pair_const(double tau, const Real* tab) {
    return {(Real)tau, (Real)(0.5 * tau), (Real)(TableOf<Real>::kStep * sqrt(tau)), tab};
}

// Copies the log Phi table of this precision into LDS with every thread of
// the block and waits for it; called before any wavefront of the block
// leaves.
template <typename Real>
__device__ __forceinline__ void load_lnphi_table(Real* lds) {
    for (int i = threadIdx.x; i < TableOf<Real>::kCount; i += blockDim.x) lds[i] = TableOf<Real>::value(i);
    __syncthreads();
}

__device__ __forceinline__ Trunc<double> trunc_terms(double d, const PairConst<double>& c) {
    return trunc64_table(d, c.sT, c.tab);
}

__device__ __forceinline__ Trunc<float> trunc_terms(float d, const PairConst<float>& c) {
    return trunc32_table(d, c.sT, c.tab);
}

__device__ __forceinline__ void distance(double d2, double& d, double& invd) {
    distance64<HwSeeds, 1>(d2, d, invd);
}
__device__ __forceinline__ void distance(float d2, float& d, float& invd) { distance32<HwSeeds>(d2, d, invd); }

// Per-wavefront accumulators of one tile.
template <typename Real, int P>
struct TileAcc {
    LaneSum<Real> s;
    LaneSum<Real> row[P];
    ColSum<Real> col[P];
};

// One step of the skewed schedule for one lane: the pair (row i, column c)
// with observations dij = delta_ij and dji = delta_ji.
template <typename Real, int P, bool TRUNC, bool SYM, bool GRAD>
__device__ __forceinline__ void pair_step(TileAcc<Real, P>& acc, const Real (&xi)[P], const Real (&xj)[P],
                                          Real dij, Real dji, bool inTri, const PairConst<Real>& pc) {
    const Real tau = pc.tau, halfTau = pc.halfTau;
    const bool vij = inTri && !m_isnan(dij);
    const bool vji = inTri && !m_isnan(dji);
    Real diff[P];
    Real d2 = Real(0);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        diff[p] = xi[p] - xj[p];
        d2 = fma(diff[p], diff[p], d2);
    }
    Real d, invd;
    distance(d2, d, invd);
    const Real rij = dij - d;
    const Real rji = dji - d;

    Real lij, lji;
    Real gt = Real(0);
    if (TRUNC) {
        const Trunc<Real> tt = trunc_terms(d, pc);
        lij = halfTau * rij * rij + tt.logPhi;
        lji = halfTau * rji * rji + tt.logPhi;
        gt = tt.gradTerm;
    } else {
        lij = rij * rij;
        lji = rji * rji;
    }
    // symmetric data: each unordered pair counts twice; the factor 2 is
    // applied to the tile's sum (exact, so the result is the same)
    if (SYM)
        acc.s.add(vij ? lij : Real(0));
    else
        acc.s.add((vij ? lij : Real(0)) + (vji ? lji : Real(0)));

    if (GRAD) {
        // c = (r - phi / (Phi sqrt tau)) * tau / d for each ordered pair
        // (src/NewMultiDimensionalScaling.hpp:435-444); masked pairs give 0.
        const Real cij = vij ? (TRUNC ? fma(-gt, pc.sT, tau * rij) : tau * rij) * invd : Real(0);
        const Real cji =
            SYM ? cij : (vji ? (TRUNC ? fma(-gt, pc.sT, tau * rji) : tau * rji) * invd : Real(0));
#pragma unroll
        for (int p = 0; p < P; ++p) {
            acc.row[p].add_product(cij, diff[p]);
            acc.col[p].add_product(SYM ? -cij : -cji, diff[p]);
            acc.col[p].rotate();
        }
    }
}

// One wavefront per tile.  Skewed schedule: at step k lane l evaluates the
// pair (row 64 I + l, column 64 J + (l + k) mod 64); the engine stores each
// tile in that order, so every step is one coalesced 64-value load.  Row
// partials stay in the lane; column partials and the column locations
// rotate one lane per step (column partials end in the lane of their
// column).  Observations are prefetched one group of G steps ahead.
//
// tileBase: this tile's plane-0 values; x: [nb * 64][P] locations; writes
// *tileSum (lane 63) and the two gradient slots of part ([nb][nb][P][64]).
template <typename Real, int P, bool TRUNC, bool SYM, bool GRAD, int G>
__device__ __forceinline__ void tile_body(const Real* __restrict__ tileBase, int64_t planeStride,
                                          const Real* __restrict__ x, int I, int J, int nb,
                                          const PairConst<Real>& pc, double* __restrict__ tileSum,
                                          double* __restrict__ part) {
    static_assert(kTile % (2 * G) == 0, "group size must divide half a tile");
    const int lane = threadIdx.x % kWave;
    const bool diag = (I == J);
    const Real* tileStep = tileBase + lane;
    // Two register buffers of G steps each (ping-pong): while one group is
    // evaluated the next one is in flight.  The loads of a group are pinned
    // ahead of the arithmetic with a scheduling barrier so that the compiler
    // cannot sink them next to their use.
    Real bufA[G], bufAT[G], bufB[G], bufBT[G];
    auto load_group = [&](Real (&b)[G], Real (&bT)[G], int k0) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int kn = min(k0 + g, kTile - 1);  // past the end: re-read step 63
            b[g] = tileStep[kn * kTile];
            bT[g] = SYM ? b[g] : tileStep[planeStride + kn * kTile];
        }
    };

    Real xi[P], xj[P];
    const Real* xrow = x + ((int64_t)I * kTile + lane) * P;
    const Real* xcol = x + ((int64_t)J * kTile + lane) * P;  // column (lane + 0) mod 64
#pragma unroll
    for (int p = 0; p < P; ++p) {
        xi[p] = xrow[p];
        xj[p] = xcol[p];
    }

    TileAcc<Real, P> acc;
    auto run_group = [&](const Real (&b)[G], const Real (&bT)[G], int k0) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int k = k0 + g;
            // diagonal tiles keep i < j only: lane < (lane + k) mod 64
            const bool inTri = !diag || (k > 0 && lane + k < kTile);
            pair_step<Real, P, TRUNC, SYM, GRAD>(acc, xi, xj, b[g], bT[g], inTri, pc);
#pragma unroll
            for (int p = 0; p < P; ++p) xj[p] = dpp_rotate<kDppWaveRol1>(xj[p]);
        }
    };

    load_group(bufA, bufAT, 0);
    for (int k0 = 0; k0 < kTile; k0 += 2 * G) {
        load_group(bufB, bufBT, k0 + G);
        __builtin_amdgcn_sched_barrier(0);
        run_group(bufA, bufAT, k0);
        load_group(bufA, bufAT, k0 + 2 * G);
        __builtin_amdgcn_sched_barrier(0);
        run_group(bufB, bufBT, k0 + G);
        // close the group of 2 G steps (FP32: one compensated update per sum)
        acc.s.flush();
#pragma unroll
        for (int p = 0; p < P; ++p) {
            acc.row[p].flush();
            acc.col[p].template flush<2 * G>();
        }
    }

    const double s = wave_sum_to_last(acc.s.value());
    if (lane == kWave - 1) *tileSum = SYM ? 2.0 * s : s;

    if (GRAD) {
        double* rowSlot = part + ((int64_t)I * nb + J) * (P * kTile) + lane;
        if (diag) {
#pragma unroll
            for (int p = 0; p < P; ++p) rowSlot[p * kTile] = acc.row[p].value() + acc.col[p].value();
        } else {
            double* colSlot = part + ((int64_t)J * nb + I) * (P * kTile) + lane;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                rowSlot[p * kTile] = acc.row[p].value();
                colSlot[p * kTile] = acc.col[p].value();
            }
        }
    }
}

}  // namespace mdship
