// Host-side declarations shared by the kernel launchers and the engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdship {

constexpr int kWave = 64;                 // CDNA wavefront width
constexpr int kTile = 64;                 // tile edge: 64 rows (one per lane) x 64 columns
constexpr int kTileElems = kTile * kTile;
constexpr int kWavesPerBlock = 4;
constexpr int kBlockThreads = kWave * kWavesPerBlock;

// Kernel embedding dimensions that are compiled; a location dimension P runs
// on the smallest of these >= P with zero-padded coordinates (padding adds 0
// to every distance and gradient).
constexpr int kKernelDims[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 16};
constexpr int kMaxDim = 16;

inline int kernel_dim(int p) {
    for (int d : kKernelDims)
        if (d >= p) return d;
    return -1;
}

// One fused evaluation over the packed upper-triangle tiles.
struct TileLaunch {
    const void* delta;       // packed tiles, Real, [planes][numTiles][64 cols][64 rows]
    const void* x;           // Real [nb * 64][kdim], zero padded
    const uint32_t* tiles;   // (I << 16) | J for I <= J, one per tile
    int numTiles;
    int nb;                  // 64-row blocks
    int kdim;                // kernel dimension
    int64_t planeStride;     // elements between the two planes (asymmetric data)
    double tau;
    bool truncated;
    bool symmetric;
    bool gradient;
    double* tileSums;        // [numTiles]
    void* part;              // [nb][nb][kdim][64] gradient partials, Real
    const double* devTau = nullptr;  // if set, the precision is read from here
};

struct ReduceLaunch {
    const void* part;        // gradient partials in the engine's real type (fp32: float)
    bool fp32;
    const double* tileSums;
    int numTiles;
    int nb;
    int kdim;
    int n;
    int p;                   // output dimension (<= kdim)
    bool gradient;
    double* grad;            // [n][p]
    double* gradHost = nullptr;  // optional second copy of it (pinned host memory)
    double* sumOut;          // 1 double: (0.5 * sum of tile sums) * sumScale
    double* sumHost = nullptr;  // optional second copy of it (pinned host memory)
    double sumScale = 1.0;
    // Optional HMC kick of each gradient row right after it is written
    // (kickForm: a KickForm of the sampler, or -1 for none): the locations
    // (the engine's real type, [n][kdim]), the momenta [n][p] and the
    // sampler's parameter block.
    int kickForm = -1;
    void* kickX = nullptr;
    double* kickMom = nullptr;
    const double* kickParams = nullptr;
};

// Change of the raw sum of increments when one location row moves (O(N)):
// per block of kRowThreads locations j, the partial
// sum_j [term(delta_ij, d_new) - term(delta_ij, d_old) + the same for delta_ji]
// over observed entries, written to partial[block] (device-accessible; the
// engine passes coherent pinned host memory) -- the caller adds the
// partials in order and halves.
constexpr int kRowThreads = 256;

struct RowDeltaLaunch {
    const void* delta;       // packed tiles as in TileLaunch
    void* x;                 // Real [nb * 64][kdim]: row `row` still holds the previous location;
                             // the last workgroup to finish writes xNew there
    double xNew[kMaxDim];    // the row's new location (the engine's real type, widened), P values
    int row;
    int n;
    int nb;
    int kdim;
    int p;
    int64_t planeStride;
    double tau;
    bool truncated;
    bool symmetric;
    bool fp32;               // Real = float (else double)
    double* partial;         // [blocks(n)]
    unsigned int* done;      // workgroups finished; 0 before the launch, reset by the last one
};

// Writes one location row (the engine's real type; padding dimensions stay 0).
struct RowWrite {
    void* x;
    double values[kMaxDim];
    int row, kdim, p;
    bool fp32;
};
hipError_t launch_row_write(const RowWrite& w, hipStream_t s);

inline int row_delta_blocks(int n) { return (n + kRowThreads - 1) / kRowThreads; }

hipError_t launch_row_delta(const RowDeltaLaunch& a, hipStream_t s);
hipError_t launch_tiles_f64(const TileLaunch& a, hipStream_t s);
hipError_t launch_tiles_f32(const TileLaunch& a, hipStream_t s);
hipError_t launch_reduce(const ReduceLaunch& a, hipStream_t s);

}  // namespace mdship
