// Host helpers shared by the single-instance engine and the batch: errors,
// a small thread pool loop, the tile list and the packing of observations
// into the kernels' tile layout.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "mds_hip.h"
#include "mds_internal.hpp"

namespace mdship {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        throw Error(MDS_HIP_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    }
}

inline int host_threads() {
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(hc == 0 ? 1u : hc, 16u));
}

template <typename F>
inline void parallel_for(size_t n, F&& f) {
    const int nt = (int)std::min<size_t>((size_t)host_threads(), std::max<size_t>(n, 1));
    if (nt <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> pool;
    const size_t chunk = (n + nt - 1) / nt;
    for (int k = 0; k < nt; ++k) {
        const size_t a = std::min(n, k * chunk), b = std::min(n, a + chunk);
        pool.emplace_back([&f, a, b] { f(a, b); });
    }
    for (auto& th : pool) th.join();
}

// Tile order: off-diagonal tiles (I < J) row-major, then the diagonal tiles
// (half the pairs) last, so the heavier work is dispatched first.
inline std::vector<uint32_t> make_tiles(int nb) {
    std::vector<uint32_t> t;
    t.reserve((size_t)nb * (nb + 1) / 2);
    for (int I = 0; I < nb; ++I)
        for (int J = I + 1; J < nb; ++J) t.push_back(((uint32_t)I << 16) | (uint32_t)J);
    for (int I = 0; I < nb; ++I) t.push_back(((uint32_t)I << 16) | (uint32_t)I);
    return t;
}

template <typename Real>
inline bool same_value(Real a, Real b) {
    return (a != a && b != b) || std::memcmp(&a, &b, sizeof(Real)) == 0;
}

// Packs the row-major N x N matrix into 64 x 64 tiles of the upper block
// triangle in the kernels' skewed order: element k * 64 + l of tile (I, J)
// holds delta(64 I + l, 64 J + (l + k) mod 64) in plane 0 and the transposed
// entry delta(64 J + (l + k) mod 64, 64 I + l) in plane 1.  Entries outside
// the matrix are NaN (missing).  Returns whether the stored values are
// symmetric, in which case plane 1 is not needed.
template <typename Real>
inline bool pack_tiles(const double* src, int n, const std::vector<uint32_t>& tiles, Real* plane0,
                       Real* plane1) {
    std::atomic<bool> asym(false);
    const Real nan = std::numeric_limits<Real>::quiet_NaN();
    parallel_for(tiles.size(), [&](size_t t0, size_t t1) {
        bool local = false;
        for (size_t t = t0; t < t1; ++t) {
            const int I = (int)(tiles[t] >> 16), J = (int)(tiles[t] & 0xffffu);
            Real* a = plane0 + t * kTileElems;
            Real* b = plane1 + t * kTileElems;
            // plane 0: row i fixed, columns visited in order (contiguous reads)
            for (int l = 0; l < kTile; ++l) {
                const int64_t i = (int64_t)I * kTile + l;
                for (int k = 0; k < kTile; ++k) {
                    const int64_t j = (int64_t)J * kTile + ((l + k) & (kTile - 1));
                    a[k * kTile + l] = (i < n && j < n) ? (Real)src[i * n + j] : nan;
                }
            }
            // plane 1: source row j fixed, l varies (contiguous reads)
            for (int c = 0; c < kTile; ++c) {
                const int64_t j = (int64_t)J * kTile + c;
                for (int l = 0; l < kTile; ++l) {
                    const int64_t i = (int64_t)I * kTile + l;
                    const int k = (c - l) & (kTile - 1);
                    const Real v = (i < n && j < n) ? (Real)src[j * n + i] : nan;
                    b[k * kTile + l] = v;
                    if (i < j && !same_value(v, a[k * kTile + l])) local = true;
                }
            }
        }
        if (local) asym.store(true);
    });
    return !asym.load();
}

}  // namespace mdship
