// O(N) change of the sum of increments when one location row moves: the
// incremental path the reference sketches for single-location updates
// (src/NewMultiDimensionalScaling.hpp:146-157, 563-616, where it is disabled)
// built for the packed tile layout.
//
// Thread j (one per location) finds delta_ij and delta_ji in the packed
// upper-triangle tiles, computes the pair's distance with the row's old and
// new location, and returns the change of the pair's terms (the same terms as
// the tile kernel: 0.5 tau r^2 + log Phi(d sqrt tau) with truncation, r^2
// without), in FP64 for both engine precisions.  Each block sums its threads
// in a fixed order (DPP wavefront sums, then the four wavefronts in order).
#include "mds_internal.hpp"
#include "mds_tile.hpp"

namespace mdship {

// Index of tile (I, J), I <= J, in the order make_tiles builds: the
// off-diagonal tiles by I then J, then the diagonal tiles.
__device__ __forceinline__ int64_t tile_index(int I, int J, int nb) {
    if (I == J) return (int64_t)nb * (nb - 1) / 2 + I;
    return (int64_t)I * (nb - 1) - (int64_t)I * (I - 1) / 2 + (J - I - 1);
}

// delta(a, b) and delta(b, a) for a != b from the packed planes.
template <typename Real>
__device__ __forceinline__ void lookup_pair(const Real* __restrict__ delta, int64_t planeStride, bool symmetric,
                                            int a, int b, int nb, double& dab, double& dba) {
    const int ba = a / kTile, bb = b / kTile;
    // the tile's row block holds lane l, its column block column c
    const bool aIsRow = ba <= bb;
    const int I = aIsRow ? ba : bb, J = aIsRow ? bb : ba;
    const int l = (aIsRow ? a : b) % kTile, c = (aIsRow ? b : a) % kTile;
    const int k = (c - l) & (kTile - 1);
    const int64_t at = tile_index(I, J, nb) * kTileElems + (int64_t)k * kTile + l;
    const double v0 = (double)delta[at];                              // delta(row, col)
    const double v1 = symmetric ? v0 : (double)delta[planeStride + at];  // delta(col, row)
    dab = aIsRow ? v0 : v1;
    dba = aIsRow ? v1 : v0;
}

__device__ __forceinline__ double pair_term(double obs, double d, double halfTau, double sT, bool truncated) {
    const double r = obs - d;
    if (!truncated) return r * r;
    return fma(halfTau * r, r, trunc64_table(d, sT, coeffs::kLnPhiTab).logPhi);
}

template <typename Real, int P>
__global__ __launch_bounds__(kRowThreads) void row_delta_kernel(RowDeltaLaunch a) {
    const Real* delta = (const Real*)a.delta;
    const Real* x = (const Real*)a.x;
    const Real* xOld = (const Real*)a.xOld;
    const int j = blockIdx.x * kRowThreads + threadIdx.x;
    const double halfTau = 0.5 * a.tau;
    const double sT = coeffs::kLnPhiTabStep * sqrt(a.tau);
    double change = 0.0;
    if (j < a.n && j != a.row) {
        double dij, dji;
        lookup_pair(delta, a.planeStride, a.symmetric, a.row, j, a.nb, dij, dji);
        double sNew = 0.0, sOld = 0.0;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const double xj = (double)x[(int64_t)j * P + p];
            const double dn = (double)x[(int64_t)a.row * P + p] - xj;
            const double dold = (double)xOld[p] - xj;
            sNew = fma(dn, dn, sNew);
            sOld = fma(dold, dold, sOld);
        }
        const double dNew = sqrt(sNew), dOld = sqrt(sOld);
        if (dij == dij)
            change += pair_term(dij, dNew, halfTau, sT, a.truncated) - pair_term(dij, dOld, halfTau, sT, a.truncated);
        if (dji == dji)
            change += pair_term(dji, dNew, halfTau, sT, a.truncated) - pair_term(dji, dOld, halfTau, sT, a.truncated);
    }
    __shared__ double waveSums[kRowThreads / kWave];
    const double w = wave_sum_to_last(change);
    if (threadIdx.x % kWave == kWave - 1) waveSums[threadIdx.x / kWave] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < kRowThreads / kWave; ++k) s += waveSums[k];
        a.partial[blockIdx.x] = s;
    }
}

template <typename Real>
static hipError_t row_delta_dim(const RowDeltaLaunch& a, hipStream_t s) {
    const dim3 grid(row_delta_blocks(a.n)), block(kRowThreads);
#define MDS_ROW_CASE(D)                                                                     \
    case D:                                                                                 \
        hipLaunchKernelGGL((row_delta_kernel<Real, D>), grid, block, 0, s, a);              \
        return hipGetLastError();
    switch (a.kdim) {
        MDS_ROW_CASE(1)
        MDS_ROW_CASE(2)
        MDS_ROW_CASE(3)
        MDS_ROW_CASE(4)
        MDS_ROW_CASE(5)
        MDS_ROW_CASE(6)
        MDS_ROW_CASE(8)
        MDS_ROW_CASE(10)
        MDS_ROW_CASE(12)
        MDS_ROW_CASE(16)
        default:
            return hipErrorInvalidValue;
    }
#undef MDS_ROW_CASE
}

hipError_t launch_row_delta(const RowDeltaLaunch& a, hipStream_t s) {
    if (a.row < 0 || a.row >= a.n || a.nb * kTile < a.n) return hipErrorInvalidValue;
    return a.fp32 ? row_delta_dim<float>(a, s) : row_delta_dim<double>(a, s);
}

}  // namespace mdship
