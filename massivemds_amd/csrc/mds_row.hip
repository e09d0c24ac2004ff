// O(N) change of the sum of increments when one location row moves: the
// incremental path the reference sketches for single-location updates
// (src/NewMultiDimensionalScaling.hpp:146-157, 563-616, where it is disabled)
// built for the packed tile layout.
//
// Thread j (one per location) finds delta_ij and delta_ji in the packed
// upper-triangle tiles, computes the pair's distance with the row's old
// location (still in the location array) and its new one (a kernel
// argument), and returns the change of the pair's terms (the same terms as
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

// The row's previous location is read from x (every workgroup reads it
// before it signals completion); the last workgroup to finish -- after all
// the others' reads -- writes the new location into x and resets the
// counter for the next launch.  The counter orders the write only; the
// change itself is summed per workgroup in a fixed order, each workgroup
// stores its partial straight into coherent pinned host memory, and the
// host adds them in order after the stream synchronises.
template <typename Real, int P>
__global__ __launch_bounds__(kRowThreads) void row_delta_kernel(RowDeltaLaunch a) {
    const Real* delta = (const Real*)a.delta;
    Real* x = (Real*)a.x;
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
            const double dn = (p < a.p ? a.xNew[p] : 0.0) - xj;
            const double dold = (double)x[(int64_t)a.row * P + p] - xj;
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
    __syncthreads();  // also: every thread of the workgroup has read the old row
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < kRowThreads / kWave; ++k) s += waveSums[k];
        a.partial[blockIdx.x] = s;  // pinned host memory
        __threadfence_system();
        const unsigned int prev = atomicAdd(a.done, 1u);
        if (prev == gridDim.x - 1) {
            __threadfence();
            for (int p = 0; p < a.p; ++p) x[(int64_t)a.row * P + p] = (Real)a.xNew[p];
            atomicExch(a.done, 0u);
        }
    }
}

template <typename Real>
__global__ void row_write_kernel(RowWrite w) {
    const int p = threadIdx.x;
    if (p < w.p) ((Real*)w.x)[(int64_t)w.row * w.kdim + p] = (Real)w.values[p];
}

hipError_t launch_row_write(const RowWrite& w, hipStream_t s) {
    if (w.p < 1 || w.p > kMaxDim || w.p > w.kdim || w.row < 0) return hipErrorInvalidValue;
    if (w.fp32)
        hipLaunchKernelGGL(row_write_kernel<float>, dim3(1), dim3(kMaxDim), 0, s, w);
    else
        hipLaunchKernelGGL(row_write_kernel<double>, dim3(1), dim3(kMaxDim), 0, s, w);
    return hipGetLastError();
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
    if (a.row < 0 || a.row >= a.n || a.nb * kTile < a.n || a.p < 1 || a.p > a.kdim || !a.done)
        return hipErrorInvalidValue;
    return a.fp32 ? row_delta_dim<float>(a, s) : row_delta_dim<double>(a, s);
}

}  // namespace mdship
