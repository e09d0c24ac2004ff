// gfx950 kernels for the pairwise truncated-normal MDS log-likelihood and its
// gradient.
//
// Work decomposition.  The N x N observation matrix is stored by the engine
// as packed 64 x 64 tiles of its upper block triangle (tile (I, J), I <= J),
// each tile in skewed order (step k, lane l) -> (row l, column (l + k) mod 64)
// so that one wavefront reads 64 consecutive values per step.  Symmetric data keep one plane; asymmetric
// data keep a second plane with the transposed tile (delta_ji), so the
// reference's row-wise semantics (row i uses delta_ij, row j uses delta_ji;
// src/NewMultiDimensionalScaling.hpp:421-461, 494-522) are reproduced while
// every transcendental is evaluated once per unordered pair.
//
// One wavefront owns one tile: lane l is row i = 64 I + l and visits the 64
// columns in 64 steps (diagonal tiles: 32 steps, each unordered pair once).
// Row partials of the gradient stay in registers; column partials rotate
// across the wavefront with DPP; column locations are read from LDS (small
// P) or rotate with the partials.
// Each tile writes its likelihood partial and its two 64-row gradient
// partials to fixed slots; a second kernel adds them in a fixed order in
// FP64.  No atomics anywhere, so results are bitwise reproducible.
#include "mds_hmc_device.hpp"
#include "mds_internal.hpp"
#include "mds_tile.hpp"

namespace mdship {

// Wavefronts per workgroup of the single-instance tile kernel.  FP64 with
// staged columns (P <= 2) runs eight, measured 2-4.5 % faster at
// N = 5,000-20,000: the 31 KB log Phi table is copied once per eight tiles
// instead of four.  The symmetric P = 2 walk's 76 VGPRs allow 6 wavefronts
// per SIMD, so three eight-wavefront workgroups per CU, as many as their
// 47 KB of LDS allow.  Everything else keeps four.
template <typename Real, int P>
constexpr int tile_block_waves() {
    return (sizeof(Real) == 8 && ColumnStage<Real, P>::kInLds) ? 2 * kWavesPerBlock : kWavesPerBlock;
}

// Small instances, at most kSmallTiles tiles (N <= 2,816: at most one
// four-wavefront workgroup per CU of the MI355X's 256), run the
// eight-wavefront kernels as four-wavefront workgroups: their few tiles then
// spread over twice as many CUs, one wavefront per SIMD, and a pass lasts
// about one tile walk (measured, FP64 P = 2: 19.1 -> 16.0 us per tile pass
// at N = 500-1,500, 4-10 % faster at N = 2,100-2,700, 8 % slower at
// N = 2,900, 5-8 % slower from N = 3,000 up).
// Their FP64 walks compute each group's terms ahead of its additions (3 %
// faster there, one wavefront per SIMD; slower in fuller launches).  Each
// tile's additions keep their order in either form, so do the results.
constexpr int kSmallTiles = 1024;

// Single instance: one wavefront per tile of the upper block triangle.  The
// truncated forms copy the log Phi table into LDS and every wavefront stages
// its tile's column locations (small P); all WAVES wavefronts of the block
// take part in the barrier before any leaves; TERMS_FIRST as in tile_walk.
template <typename Real, int P, bool TRUNC, bool SYM, bool GRAD, int WAVES, bool TERMS_FIRST>
__device__ __forceinline__ void tile_block(Real* __restrict__ lnPhiTab, Real (*colStage)[ColumnStage<Real, P>::kElems],
                                           const Real* __restrict__ delta, const Real* __restrict__ x,
                                           const uint32_t* __restrict__ tiles, int numTiles, int nb,
                                           int64_t planeStride, double tau, double* __restrict__ tileSums,
                                           Real* __restrict__ part, const double* __restrict__ devTau) {
    const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES + threadIdx.x / kWave);
    const uint32_t code = t < numTiles ? tiles[t] : 0u;
    Real* stage = colStage[threadIdx.x / kWave];
    if (t < numTiles) stage_columns<Real, P>(stage, x, (int)(code & 0xffffu));
    if constexpr (TRUNC) copy_lnphi_table(lnPhiTab);
    __syncthreads();
    if (t >= numTiles) return;
    const PairConst<Real> pc = make_pair_const<Real>(devTau ? *devTau : tau, lnPhiTab);
    tile_walk<Real, P, TRUNC, SYM, GRAD, kGroup, TERMS_FIRST>(delta + (int64_t)t * kTileElems, planeStride, x,
                                                              stage, (int)(code >> 16), (int)(code & 0xffffu), nb,
                                                              pc, tileSums + t, part);
}

template <typename Real, int P, bool TRUNC, bool SYM, bool GRAD>
__global__ __launch_bounds__((tile_block_waves<Real, P>() * kWave)) void tile_kernel(
    const Real* __restrict__ delta, const Real* __restrict__ x,
    const uint32_t* __restrict__ tiles, int numTiles, int nb, int64_t planeStride,
    double tau, double* __restrict__ tileSums, Real* __restrict__ part, const double* __restrict__ devTau) {
    constexpr int kWaves = tile_block_waves<Real, P>();
    __shared__ __align__(16) Real lnPhiTab[TRUNC ? TableOf<Real>::kCount : 1];
    __shared__ __align__(16) Real colStage[kWaves][ColumnStage<Real, P>::kElems];
    tile_block<Real, P, TRUNC, SYM, GRAD, kWaves, (sizeof(Real) == 4)>(lnPhiTab, colStage, delta, x, tiles,
                                                                       numTiles, nb, planeStride, tau, tileSums,
                                                                       part, devTau);
}

// The four-wavefront form of the eight-wavefront kernels, for small instances.
template <typename Real, int P, bool TRUNC, bool SYM, bool GRAD>
__global__ __launch_bounds__(kWavesPerBlock * kWave) void tile_kernel_small(
    const Real* __restrict__ delta, const Real* __restrict__ x,
    const uint32_t* __restrict__ tiles, int numTiles, int nb, int64_t planeStride,
    double tau, double* __restrict__ tileSums, Real* __restrict__ part, const double* __restrict__ devTau) {
    __shared__ __align__(16) Real lnPhiTab[TRUNC ? TableOf<Real>::kCount : 1];
    __shared__ __align__(16) Real colStage[kWavesPerBlock][ColumnStage<Real, P>::kElems];
    tile_block<Real, P, TRUNC, SYM, GRAD, kWavesPerBlock, true>(lnPhiTab, colStage, delta, x, tiles, numTiles,
                                                                nb, planeStride, tau, tileSums, part, devTau);
}

constexpr int kReduceWaves = 8;
constexpr int kReduceThreads = kReduceWaves * kWave;

// Block b < nb: gradient rows of row block b, 8 wavefronts splitting the
// column blocks J (J = w mod 8), combined through LDS in fixed order.
// Last block: the likelihood, a fixed-order FP64 sum of the tile partials,
// halved as in src/NewMultiDimensionalScaling.hpp:556, then scaled by
// sumScale (1, or 0.5 tau for the untruncated value of :198-202), into
// sumOut and, when sumHost is set, also there (pinned host memory the caller
// reads after a stream wait, with no copy command); likewise the gradient
// rows into grad and gradHost.
// KICK >= 0: each row's thread then applies the sampler's kick of that form
// to its row (kick_row, reading back the gradient it has just written), so
// that a leapfrog step needs no kick launch of its own; the reduce blocks
// read no locations, so moving them here races with nothing.
template <typename PartT, int P, int KICK>
__global__ __launch_bounds__(kReduceThreads) void reduce_kernel(
    const PartT* __restrict__ part, const double* __restrict__ tileSums, int numTiles,
    int nb, int n, int pOut, double* __restrict__ grad, double* __restrict__ gradHost, double* __restrict__ sumOut,
    double* __restrict__ sumHost, double sumScale,
    PartT* __restrict__ kickX, double* __restrict__ kickMom, const double* __restrict__ kickParams) {
    __shared__ double red[kReduceThreads * (P > 1 ? P : 1)];
    const int tid = threadIdx.x;
    // Loads are issued kAhead iterations at a time and added in the same
    // order as a plain loop would (latency hidden, results unchanged).
    constexpr int kAhead = 4;
    if (blockIdx.x == gridDim.x - 1) {
        double acc = 0.0;
        int k = tid;
        for (; k + (2 * kAhead - 1) * kReduceThreads < numTiles; k += 2 * kAhead * kReduceThreads) {
            double v[2 * kAhead];
#pragma unroll
            for (int u = 0; u < 2 * kAhead; ++u) v[u] = tileSums[k + u * kReduceThreads];
#pragma unroll
            for (int u = 0; u < 2 * kAhead; ++u) acc += v[u];
        }
        for (; k < numTiles; k += kReduceThreads) acc += tileSums[k];
        red[tid] = acc;
        __syncthreads();
        for (int s = kReduceThreads / 2; s > 0; s >>= 1) {
            if (tid < s) red[tid] += red[tid + s];
            __syncthreads();
        }
        if (tid == 0) {
            const double v = (0.5 * red[0]) * sumScale;  // x * 1.0 == x: the plain sum keeps its bits
            *sumOut = v;
            if (sumHost) *sumHost = v;
        }
        return;
    }
    const int K = blockIdx.x;
    const int wave = tid / kWave;
    const int lane = tid % kWave;
    double acc[P];
#pragma unroll
    for (int p = 0; p < P; ++p) acc[p] = 0.0;
    const PartT* base = part + (int64_t)K * nb * (P * kTile) + lane;
    int J = wave;
    for (; J + (kAhead - 1) * kReduceWaves < nb; J += kAhead * kReduceWaves) {
        double v[kAhead][P];
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            const PartT* slot = base + (int64_t)(J + u * kReduceWaves) * (P * kTile);
#pragma unroll
            for (int p = 0; p < P; ++p) v[u][p] = (double)slot[p * kTile];
        }
#pragma unroll
        for (int u = 0; u < kAhead; ++u)
#pragma unroll
            for (int p = 0; p < P; ++p) acc[p] += v[u][p];
    }
    for (; J < nb; J += kReduceWaves) {
        const PartT* slot = base + (int64_t)J * (P * kTile);
#pragma unroll
        for (int p = 0; p < P; ++p) acc[p] += (double)slot[p * kTile];
    }
#pragma unroll
    for (int p = 0; p < P; ++p) red[(p * kReduceWaves + wave) * kWave + lane] = acc[p];
    __syncthreads();
    if (wave == 0) {
        const int row = K * kTile + lane;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            double v = 0.0;
#pragma unroll
            for (int w = 0; w < kReduceWaves; ++w) v += red[(p * kReduceWaves + w) * kWave + lane];
            if (row < n && p < pOut) {
                grad[(int64_t)row * pOut + p] = v;
                if (gradHost) gradHost[(int64_t)row * pOut + p] = v;
            }
        }
        if constexpr (KICK >= 0) kick_row<PartT, KICK>(kickX, kickMom, grad, kickParams, n, pOut, P, row);
    }
}

template <typename Real, int P, bool TRUNC, bool SYM, bool GRAD>
static hipError_t go(const TileLaunch& a, hipStream_t s) {
    constexpr int kWaves = tile_block_waves<Real, P>();
    if constexpr (kWaves != kWavesPerBlock) {
        if (a.numTiles <= kSmallTiles) {
            const dim3 grid((a.numTiles + kWavesPerBlock - 1) / kWavesPerBlock);
            hipLaunchKernelGGL((tile_kernel_small<Real, P, TRUNC, SYM, GRAD>), grid, dim3(kWavesPerBlock * kWave), 0,
                               s, (const Real*)a.delta, (const Real*)a.x, a.tiles, a.numTiles, a.nb, a.planeStride,
                               a.tau, a.tileSums, (Real*)a.part, a.devTau);
            return hipGetLastError();
        }
    }
    const dim3 grid((a.numTiles + kWaves - 1) / kWaves);
    hipLaunchKernelGGL((tile_kernel<Real, P, TRUNC, SYM, GRAD>), grid, dim3(kWaves * kWave), 0, s,
                       (const Real*)a.delta, (const Real*)a.x, a.tiles, a.numTiles, a.nb, a.planeStride, a.tau,
                       a.tileSums, (Real*)a.part, a.devTau);
    return hipGetLastError();
}

template <typename Real, int P>
static hipError_t dispatch_flags(const TileLaunch& a, hipStream_t s) {
    if (a.truncated) {
        if (a.symmetric)
            return a.gradient ? go<Real, P, true, true, true>(a, s) : go<Real, P, true, true, false>(a, s);
        return a.gradient ? go<Real, P, true, false, true>(a, s) : go<Real, P, true, false, false>(a, s);
    }
    if (a.symmetric)
        return a.gradient ? go<Real, P, false, true, true>(a, s) : go<Real, P, false, true, false>(a, s);
    return a.gradient ? go<Real, P, false, false, true>(a, s) : go<Real, P, false, false, false>(a, s);
}

template <typename Real>
static hipError_t dispatch_dim(const TileLaunch& a, hipStream_t s) {
    switch (a.kdim) {
        case 1: return dispatch_flags<Real, 1>(a, s);
        case 2: return dispatch_flags<Real, 2>(a, s);
        case 3: return dispatch_flags<Real, 3>(a, s);
        case 4: return dispatch_flags<Real, 4>(a, s);
        case 5: return dispatch_flags<Real, 5>(a, s);
        case 6: return dispatch_flags<Real, 6>(a, s);
        case 8: return dispatch_flags<Real, 8>(a, s);
        case 10: return dispatch_flags<Real, 10>(a, s);
        case 12: return dispatch_flags<Real, 12>(a, s);
        case 16: return dispatch_flags<Real, 16>(a, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_tiles_f64(const TileLaunch& a, hipStream_t s) { return dispatch_dim<double>(a, s); }
hipError_t launch_tiles_f32(const TileLaunch& a, hipStream_t s) { return dispatch_dim<float>(a, s); }

template <typename PartT, int P, int KICK>
static void reduce_go(const ReduceLaunch& a, hipStream_t s) {
    const dim3 grid(a.gradient ? a.nb + 1 : 1);
    hipLaunchKernelGGL((reduce_kernel<PartT, P, KICK>), grid, dim3(kReduceThreads), 0, s, (const PartT*)a.part,
                       a.tileSums, a.numTiles, a.nb, a.n, a.p, a.grad, a.gradHost, a.sumOut, a.sumHost, a.sumScale,
                       (PartT*)a.kickX,
                       a.kickMom, a.kickParams);
}

template <typename PartT, int P>
static hipError_t reduce_kick(const ReduceLaunch& a, hipStream_t s) {
    switch (a.kickForm) {
        case -1: reduce_go<PartT, P, -1>(a, s); break;
        case kKick: reduce_go<PartT, P, kKick>(a, s); break;
        case kKick2Drift: reduce_go<PartT, P, kKick2Drift>(a, s); break;
        default: return hipErrorInvalidValue;  // the kick that starts a trajectory has no gradient before it
    }
    return hipGetLastError();
}

template <int P>
static hipError_t reduce_p(const ReduceLaunch& a, hipStream_t s) {
    if (a.kickForm >= 0 && (!a.gradient || !a.kickX || !a.kickMom || !a.kickParams)) return hipErrorInvalidValue;
    return a.fp32 ? reduce_kick<float, P>(a, s) : reduce_kick<double, P>(a, s);
}

hipError_t launch_reduce(const ReduceLaunch& a, hipStream_t s) {
    switch (a.kdim) {
        case 1: return reduce_p<1>(a, s);
        case 2: return reduce_p<2>(a, s);
        case 3: return reduce_p<3>(a, s);
        case 4: return reduce_p<4>(a, s);
        case 5: return reduce_p<5>(a, s);
        case 6: return reduce_p<6>(a, s);
        case 8: return reduce_p<8>(a, s);
        case 10: return reduce_p<10>(a, s);
        case 12: return reduce_p<12>(a, s);
        case 16: return reduce_p<16>(a, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace mdship
