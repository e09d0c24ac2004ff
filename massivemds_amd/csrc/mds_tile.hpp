// Device functions shared by the single-instance and batched kernels: the
// per-pair arithmetic and the wavefront's walk over one 64 x 64 tile.
#pragma once

#include "mds_device.hpp"

namespace mdship {

// Steps per prefetch group of the tile walk.
constexpr int kGroup = 4;

// Per-evaluation constants of the pair arithmetic: tau / 2, sT = S sqrt(tau)
// (z in the units of the log Phi table, S = 64 for FP64 and 32 for FP32,
// which `tab` points to in LDS) and sT / 2.  The truncation's gradient term
// g = d log Phi / d(S z) enters the pair coefficient as
// tau (phi / Phi) / sqrt(tau) = sT g.
template <typename Real>
struct PairConst {
    Real halfTau, sT, halfST;
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

// The value of the wavefront's first active lane, in scalar registers: for
// values that are the same in every lane, so that the walk's arithmetic
// takes them as scalar operands instead of reading vector registers.  The
// empty asm statement hides where the value came from: knowing it uniform,
// the compiler would otherwise drop the read and redo the FP arithmetic that
// produced it, which has no scalar form, in vector registers.  The read
// itself is the compiler's, so it keeps the wait states the hardware needs
// between the FP64 instruction that writes the value and the read.
__device__ __forceinline__ unsigned first_lane(unsigned v) {
    asm volatile("" : "+v"(v));
    return (unsigned)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ float wave_uniform(float v) {
    return __builtin_bit_cast(float, first_lane(__builtin_bit_cast(unsigned, v)));
}
__device__ __forceinline__ double wave_uniform(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = first_lane((unsigned)b);
    const unsigned hi = first_lane((unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Per-evaluation constants from the precision (as the host holds it, or
// read from a sampler's device parameter block), computed in FP64; the same
// in every lane, so held in scalar registers.
template <typename Real>
__device__ __forceinline__ PairConst<Real> make_pair_const(double tau, const Real* tab) {
    const double sT = TableOf<Real>::kStep * sqrt(tau);
    return {wave_uniform((Real)(0.5 * tau)), wave_uniform((Real)sT), wave_uniform((Real)(0.5 * sT)), tab};
}

// Copies the log Phi table of this precision into LDS with every thread of
// the block; the caller waits for it (__syncthreads) before any read.
template <typename Real>
__device__ __forceinline__ void copy_lnphi_table(Real* lds) {
    for (int i = threadIdx.x; i < TableOf<Real>::kCount; i += blockDim.x) lds[i] = TableOf<Real>::value(i);
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

// Column locations of a tile.  When a column's P coordinates fit in
// kStageBytes (FP32: P <= 10, FP64: P <= 2) each wavefront stages its tile's
// 64 columns in LDS twice over, so that step k of lane l reads column l + k
// with no lane index wrapping and no VALU instruction moving a column.  For
// P <= 2 the stage is coordinate-major ([P][128]: coordinate p of column c at
// p * 128 + c, both coordinates of a step within the 8-bit offsets of one
// ds_read2); for larger P it is column-major ([128][P]: at c * P + p), so
// that every read of a group of steps is an immediate offset from one address
// per group (coordinate-major rows 512 bytes apart would need an address add
// per coordinate and step).  In column-major reads the lanes are P words
// apart: conflict-free for P = 5 (an odd number of words) and P = 10 (of
// doublewords); even P may see two-way bank conflicts.
// Larger P keep the columns in registers and rotate them one lane per step
// (DPP, one VALU instruction per 32-bit word), which costs no LDS.  The limits
// are measured: staging P = 5 and 10 in FP32 took 9 % off the tile kernel
// (N = 10,000: 75.6 -> 68.7 us at P = 5, 119.2 -> 108.7 us at P = 10), while
// FP64 P = 5, whose table already fills a quarter of the LDS, lost occupancy
// and ran 4 % slower.
template <typename Real, int P>
struct ColumnStage {
    static constexpr int kStageBytes = sizeof(Real) == 4 ? 40 : 16;
    static constexpr bool kInLds = sizeof(Real) * P <= kStageBytes;
    static constexpr bool kColumnMajor = kInLds && P > 2;
    static constexpr int kElems = kInLds ? P * 2 * kTile : 1;  // per wavefront
    static constexpr int kMaxElems = kStageBytes / sizeof(Real) * 2 * kTile;  // any P that stages
    // offset of coordinate p of stage column c
    __device__ static constexpr int at(int c, int p) { return kColumnMajor ? c * P + p : p * 2 * kTile + c; }
};

// Lane l writes column 64 J + l of x ([nb * 64][P]) into both halves.
template <typename Real, int P>
__device__ __forceinline__ void stage_columns(Real* __restrict__ stage, const Real* __restrict__ x, int J) {
    using CS = ColumnStage<Real, P>;
    if constexpr (CS::kInLds) {
        const int lane = threadIdx.x % kWave;
        const Real* src = x + ((int64_t)J * kTile + lane) * P;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const Real v = src[p];
            stage[CS::at(lane, p)] = v;
            stage[CS::at(lane + kTile, p)] = v;
        }
    }
}

// Orders one wavefront's LDS accesses across a point where lanes start to
// read what other lanes wrote (or overwrite what they read): LDS executes a
// wavefront's instructions in order, so only the compiler must not move
// accesses across it.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-wavefront accumulators of one tile.
template <typename Real, int P>
struct TileAcc {
    LaneSum<Real> s;
    LaneSum<Real> row[P];
    ColSum<Real> col[P];
};

// One step of the skewed schedule for one lane: the pair (row i, column c)
// with observations dij = delta_ij and dji = delta_ji.  The gradient
// accumulators take half the pair coefficients (tau / 2 and sT / 2 in place
// of tau and sT: exact halvings), and the tile doubles its partials at the
// end, so that tau / 2 serves the likelihood and the gradient alike.
// pair_terms computes the step's terms for every lane; pair_accumulate adds
// them for the observed pairs only.
template <typename Real, int P>
struct PairTerms {
    Real diff[P];
    Real lij, lji, hij, hji;
    bool vij, vji;
};

template <typename Real, int P, bool TRUNC, bool SYM, bool GRAD>
__device__ __forceinline__ PairTerms<Real, P> pair_terms(const Real (&xi)[P], const Real (&xj)[P], Real dij,
                                                         Real dji, bool inTri, const PairConst<Real>& pc) {
    PairTerms<Real, P> t;
    const Real halfTau = pc.halfTau;
    t.vij = inTri && !m_isnan(dij);
    t.vji = inTri && !m_isnan(dji);
    Real d2 = Real(0);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        t.diff[p] = xi[p] - xj[p];
        d2 = fma(t.diff[p], t.diff[p], d2);
    }
    Real d, invd;
    distance(d2, d, invd);
    const Real rij = dij - d;
    const Real rji = dji - d;
    const Real aij = halfTau * rij;
    const Real aji = halfTau * rji;
    Real gt = Real(0);
    if (TRUNC) {
        const Trunc<Real> tt = trunc_terms(d, pc);
        t.lij = fma(aij, rij, tt.logPhi);
        t.lji = fma(aji, rji, tt.logPhi);
        gt = tt.gradTerm;
    } else {
        t.lij = rij * rij;
        t.lji = rji * rji;
    }
    // c / 2 = (r - phi / (Phi sqrt tau)) * (tau / 2) / d for each ordered
    // pair (src/NewMultiDimensionalScaling.hpp:435-444)
    t.hij = (TRUNC ? fma(-gt, pc.halfST, aij) : aij) * invd;
    t.hji = SYM ? t.hij : (TRUNC ? fma(-gt, pc.halfST, aji) : aji) * invd;
    return t;
}

// Unobserved pairs add nothing: the additions run as branches on the
// observation mask, so that the lanes of unobserved pairs are switched off
// for them (no select per term), and the column accumulators rotate after.
// Symmetric data: each unordered pair counts twice; the factor 2 is applied
// to the tile's sum (exact, so the result is the same).
template <typename Real, int P, bool SYM, bool GRAD>
__device__ __forceinline__ void pair_accumulate(TileAcc<Real, P>& acc, const PairTerms<Real, P>& t) {
    if (__builtin_expect(t.vij, 1)) {
        acc.s.add(t.lij);
        if (GRAD) {
#pragma unroll
            for (int p = 0; p < P; ++p) {
                acc.row[p].add_product(t.hij, t.diff[p]);
                if (SYM) acc.col[p].add_product(-t.hij, t.diff[p]);
            }
        }
    }
    if (!SYM) {
        if (__builtin_expect(t.vji, 1)) {
            acc.s.add(t.lji);
            if (GRAD) {
#pragma unroll
                for (int p = 0; p < P; ++p) acc.col[p].add_product(-t.hji, t.diff[p]);
            }
        }
    }
    if (GRAD) {
#pragma unroll
        for (int p = 0; p < P; ++p) acc.col[p].rotate();
    }
}

// One wavefront per tile.  Skewed schedule: at step k lane l evaluates the
// pair (row 64 I + l, column 64 J + (l + k) mod 64); the engine stores each
// tile in that order, so every step is one coalesced 64-value load.  Row
// partials stay in the lane; column partials rotate one lane per step (DPP)
// and end in the lane of their column.  Observations are prefetched one
// group of G steps ahead.
//
// Off-diagonal tiles (DIAG = false) take steps 0..63 with no triangle mask.
// A diagonal tile holds each unordered pair {a, b} of its block twice; steps
// 1..31 with every lane and step 32 with lanes 0..31 visit each exactly
// once (lane a at step b - a when b - a <= 32, else lane b at step
// 64 - (b - a), as the row of the pair with the observations read in that
// orientation), half the steps of the full walk.
//
// tileBase: this tile's plane-0 values; x: [nb * 64][P] locations; stage:
// this wavefront's staged columns (see ColumnStage; unused for larger P);
// writes *tileSum (lane 63) and the two gradient slots of part
// ([nb][nb][P][64] in the engine's real type: each slot is the sum of one
// 64 x 64 tile's terms for 64 rows, which the reduce adds in FP64; stored in
// FP32 by the FP32 engine, whose per-term arithmetic already carries FP32
// rounding, so that its partial traffic halves).
template <typename Real, int P, bool TRUNC, bool SYM, bool GRAD, bool DIAG, int G, bool TERMS_FIRST>
__device__ __forceinline__ void tile_body(const Real* __restrict__ tileBase, int64_t planeStride,
                                          const Real* __restrict__ x, const Real* __restrict__ stage, int I,
                                          int J, int nb, const PairConst<Real>& pc, double* __restrict__ tileSum,
                                          Real* __restrict__ part) {
    static_assert(kTile % (2 * G) == 0 && (kTile / 2) % (2 * G) == 0, "group size must divide a quarter tile");
    constexpr bool kStaged = ColumnStage<Real, P>::kInLds;
    constexpr int kFirst = DIAG ? 1 : 0;                 // first step
    constexpr int kSteps = DIAG ? kTile / 2 : kTile;     // steps walked
    constexpr int kLast = kFirst + kSteps - 1;
    constexpr int kFlushSteps = 8 * G;                   // steps per FP32 group sum
    constexpr int kFlushPasses = kFlushSteps / (2 * G);  // loop passes per FP32 group sum
    static_assert(kSteps % kFlushSteps == 0, "the walk must end on a flush");
    const int lane = threadIdx.x % kWave;
    const Real* tileStep = tileBase + lane;
    // Two register buffers of G steps each (ping-pong): while one group is
    // evaluated the next one is in flight.  The loads of a group are pinned
    // ahead of the arithmetic with a scheduling barrier so that the compiler
    // cannot sink them next to their use.
    Real bufA[G], bufAT[G], bufB[G], bufBT[G];
    auto load_group = [&](Real (&b)[G], Real (&bT)[G], int k0) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            b[g] = tileStep[(k0 + g) * kTile];
            bT[g] = SYM ? b[g] : tileStep[planeStride + (k0 + g) * kTile];
        }
    };

    Real xi[P], xj[P];
    const Real* xrow = x + ((int64_t)I * kTile + lane) * P;
    const Real* xcol = x + ((int64_t)J * kTile + ((lane + kFirst) & (kTile - 1))) * P;  // column of step kFirst
#pragma unroll
    for (int p = 0; p < P; ++p) {
        xi[p] = xrow[p];
        if constexpr (!kStaged) xj[p] = xcol[p];
    }
    using CS = ColumnStage<Real, P>;
    const Real* stageLane = stage + CS::at(lane, 0);

    TileAcc<Real, P> acc;
    auto run_group = [&](const Real (&b)[G], const Real (&bT)[G], int k0) {
        Real xs[G][P];
        if constexpr (kStaged) {
            const Real* stageGroup = stageLane + CS::at(k0, 0);
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int p = 0; p < P; ++p) xs[g][p] = stageGroup[CS::at(g, p)];
        }
        // TERMS_FIRST: the group's terms first (one block, in which the
        // steps' work interleaves), then their additions step by step;
        // otherwise each step's additions right after its terms (the
        // compiler then moves the whole step under the observation mask).
        // The additions keep their order either way, so do the results.
        // Measured on the tile kernel at N = 10,000 against selects of
        // zeros: FP32 -4 % (P = 2) to -7.5 % (P = 10) with the terms first,
        // +5 % the other way; FP64 -2 % step by step, +5 % the other way (its
        // terms of four steps at once contend for the LDS table), but -3 %
        // with the terms first where one wavefront per SIMD walks (small
        // instances, whose walks are latency-bound).
        constexpr bool kGroupTermsFirst = TERMS_FIRST;
        PairTerms<Real, P> terms[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int k = k0 + g;
            const bool inTri = !DIAG || k < kTile / 2 || lane < kTile / 2;
            if constexpr (kStaged) {
                terms[g] = pair_terms<Real, P, TRUNC, SYM, GRAD>(xi, xs[g], b[g], bT[g], inTri, pc);
            } else {
                terms[g] = pair_terms<Real, P, TRUNC, SYM, GRAD>(xi, xj, b[g], bT[g], inTri, pc);
#pragma unroll
                for (int p = 0; p < P; ++p) xj[p] = dpp_rotate<kDppWaveRol1>(xj[p]);
            }
            if constexpr (!kGroupTermsFirst) pair_accumulate<Real, P, SYM, GRAD>(acc, terms[g]);
        }
        if constexpr (kGroupTermsFirst) {
#pragma unroll
            for (int g = 0; g < G; ++g) pair_accumulate<Real, P, SYM, GRAD>(acc, terms[g]);
        }
    };

    load_group(bufA, bufAT, kFirst);
#pragma nounroll
    for (int k0 = kFirst; k0 <= kLast; k0 += 2 * G) {
        load_group(bufB, bufBT, k0 + G);
        __builtin_amdgcn_sched_barrier(0);
        run_group(bufA, bufAT, k0);
        // the walk's last pass re-reads its last group instead of reading
        // past it (one clamp for the group: its loads share one address)
        load_group(bufA, bufAT, min(k0 + 2 * G, kLast + 1 - G));
        __builtin_amdgcn_sched_barrier(0);
        run_group(bufB, bufBT, k0 + G);
        // every fourth pass (8 G steps) the FP32 group sums enter their FP64
        // sums (FP64: nothing to do); the walk's step count is a multiple of
        // 8 G, so its last pass flushes
        if (((k0 - kFirst) / (2 * G)) % kFlushPasses == kFlushPasses - 1) {
            acc.s.flush();
#pragma unroll
            for (int p = 0; p < P; ++p) {
                acc.row[p].flush();
                acc.col[p].template flush<kFlushSteps>();
            }
        }
    }

    const double s = wave_sum_to_last(acc.s.value());
    if (lane == kWave - 1) *tileSum = SYM ? 2.0 * s : s;

    if (GRAD) {
        // the accumulators hold half the gradient terms: doubled here (exact)
        Real* rowSlot = part + ((int64_t)I * nb + J) * (P * kTile) + lane;
        if constexpr (DIAG) {
            // after the walk lane l holds column (l + kFirst + kSteps) mod 64
            constexpr int kAlign = (2 * kTile - kFirst - kSteps) % kTile;
#pragma unroll
            for (int p = 0; p < P; ++p)
                rowSlot[p * kTile] = (Real)(2.0 * (acc.row[p].value() + lane_rotate<kAlign>(acc.col[p].value())));
        } else {
            Real* colSlot = part + ((int64_t)J * nb + I) * (P * kTile) + lane;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                rowSlot[p * kTile] = (Real)(2.0 * acc.row[p].value());
                colSlot[p * kTile] = (Real)(2.0 * acc.col[p].value());
            }
        }
    }
}

// Both walks of the tile (I, J): diagonal tiles take the half walk.
// TERMS_FIRST: the schedule of a group's steps (see tile_body); FP32 walks
// take the terms first, FP64 walks step by step except in the
// small-instance kernels.
template <typename Real, int P, bool TRUNC, bool SYM, bool GRAD, int G, bool TERMS_FIRST = (sizeof(Real) == 4)>
__device__ __forceinline__ void tile_walk(const Real* __restrict__ tileBase, int64_t planeStride,
                                          const Real* __restrict__ x, const Real* __restrict__ stage, int I,
                                          int J, int nb, const PairConst<Real>& pc, double* __restrict__ tileSum,
                                          Real* __restrict__ part) {
    if (I == J)
        tile_body<Real, P, TRUNC, SYM, GRAD, true, G, TERMS_FIRST>(tileBase, planeStride, x, stage, I, J, nb, pc,
                                                                   tileSum, part);
    else
        tile_body<Real, P, TRUNC, SYM, GRAD, false, G, TERMS_FIRST>(tileBase, planeStride, x, stage, I, J, nb, pc,
                                                                    tileSum, part);
}

}  // namespace mdship
