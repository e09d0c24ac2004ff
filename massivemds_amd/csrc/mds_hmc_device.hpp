// Device arithmetic of the HMC sampler, shared by the single-chain sampler
// (mds_hmc.hip) and the batched one (mds_batch.hip): the same operations in
// the same order, so a chain gives the same bits in either.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mds_device.hpp"  // floating-point contraction within single expressions, before any code here
#include "mds_hmc.hpp"

namespace mdship {

// splitmix64 finaliser, as mdsgen::mix on the host.
__device__ __forceinline__ uint64_t dmix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Standard normal momenta from a counter-based stream keyed by (seed,
// iteration): the same splitmix64 / Box-Muller construction as the host
// generator; `pair` writes mom[2 pair] and mom[2 pair + 1].
__device__ __forceinline__ void momentum_pair(double* __restrict__ mom, int count, uint64_t seed, uint64_t iteration,
                                              int pair) {
    const int k = 2 * pair;
    if (k >= count) return;
    const uint64_t state = dmix(seed * 0x100000001B3ULL + (iteration + 1) * 0xC2B2AE3D27D4EB4FULL);
    const uint64_t z1 = dmix(state + (uint64_t)(k + 1) * 0x9E3779B97F4A7C15ULL);
    const uint64_t z2 = dmix(state + (uint64_t)(k + 2) * 0x9E3779B97F4A7C15ULL);
    const double u1 = ((double)(z1 >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = ((double)(z2 >> 11) + 0.5) * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    mom[k] = r * c;
    if (k + 1 < count) mom[k + 1] = r * s;
}

// Row i: g = grad log-likelihood + grad log prior = grad_ll - x_i T.
// The reference's leapfrog (R/Untitled.R:611-620) takes two half kicks per
// step, mom += eps / 2 * g before and after the drift x += eps * mom; the
// forms: KICK mom += halfEps * g; KICK_DRIFT the kick, then the drift;
// KICK2_DRIFT the half kick ending one step and the one starting the next
// (same locations, so the same g; two separate additions, as two launches
// would make them), then the drift.
enum KickForm { kKick = 0, kKickDrift = 1, kKick2Drift = 2 };

template <typename Real, int FORM>
__device__ __forceinline__ void kick_row(Real* __restrict__ x, double* __restrict__ mom,
                                         const double* __restrict__ grad, const double* __restrict__ params, int n,
                                         int p, int kdim, int i) {
    if (i >= n) return;
    const double eps = params[kHmcEps];
    const double halfEps = 0.5 * eps;
    const double* T = params + kHmcTrait;
    Real* xi = x + (int64_t)i * kdim;
    for (int d = 0; d < p; ++d) {
        double prior = 0.0;
        for (int e = 0; e < p; ++e) prior += (double)xi[e] * T[e * p + d];
        const double g = grad[(int64_t)i * p + d] - prior;
        double m = mom[(int64_t)i * p + d];
        m = fma(halfEps, g, m);
        if (FORM == kKick2Drift) m = fma(halfEps, g, m);
        mom[(int64_t)i * p + d] = m;
    }
    if (FORM != kKick) {
        for (int d = 0; d < p; ++d) xi[d] = (Real)((double)xi[d] + eps * mom[(int64_t)i * p + d]);
    }
}

// out[0] = sum mom^2 / 2, out[1] = sum_i x_i^T T x_i (fixed-order
// reductions over one block of kEnergyThreads threads: thread t adds rows t,
// t + kEnergyThreads, ... in turn, then a fixed tree).
constexpr int kEnergyThreads = 512;

// Thread t's rows for P coordinates: the loads of kAhead rows are issued
// before their additions, which keep the row order and each row's
// expressions (the same bits as adding row after row; the block is
// otherwise bound by one load's latency per row).
template <typename Real, int P>
__device__ __forceinline__ void energy_rows(const Real* __restrict__ x, const double* __restrict__ mom,
                                            const double* __restrict__ T, int n, int kdim, double& k, double& q) {
    constexpr int kAhead = P <= 2 ? 8 : (P <= 4 ? 4 : (P <= 8 ? 2 : 1));
    for (int i0 = threadIdx.x; i0 < n; i0 += kAhead * kEnergyThreads) {
        double m[kAhead][P], xv[kAhead][P];
#pragma unroll
        for (int a = 0; a < kAhead; ++a) {
            const int i = i0 + a * kEnergyThreads;
#pragma unroll
            for (int d = 0; d < P; ++d) {
                m[a][d] = i < n ? mom[(int64_t)i * P + d] : 0.0;
                xv[a][d] = i < n ? (double)x[(int64_t)i * kdim + d] : 0.0;
            }
        }
#pragma unroll
        for (int a = 0; a < kAhead; ++a) {
            if (i0 + a * kEnergyThreads >= n) break;
#pragma unroll
            for (int d = 0; d < P; ++d) {
                const double md = m[a][d];
                k += 0.5 * md * md;
                double row = 0.0;
#pragma unroll
                for (int e = 0; e < P; ++e) row += xv[a][e] * T[e * P + d];
                q += row * xv[a][d];
            }
        }
    }
}

template <typename Real>
__device__ __forceinline__ void energy_block(const Real* __restrict__ x, const double* __restrict__ mom,
                                             const double* __restrict__ params, int n, int p, int kdim,
                                             double* __restrict__ out, double (*red)[kEnergyThreads]) {
    const double* T = params + kHmcTrait;
    double k = 0.0, q = 0.0;
    switch (p) {
        case 1: energy_rows<Real, 1>(x, mom, T, n, kdim, k, q); break;
        case 2: energy_rows<Real, 2>(x, mom, T, n, kdim, k, q); break;
        case 3: energy_rows<Real, 3>(x, mom, T, n, kdim, k, q); break;
        case 4: energy_rows<Real, 4>(x, mom, T, n, kdim, k, q); break;
        case 5: energy_rows<Real, 5>(x, mom, T, n, kdim, k, q); break;
        case 6: energy_rows<Real, 6>(x, mom, T, n, kdim, k, q); break;
        case 7: energy_rows<Real, 7>(x, mom, T, n, kdim, k, q); break;
        case 8: energy_rows<Real, 8>(x, mom, T, n, kdim, k, q); break;
        case 9: energy_rows<Real, 9>(x, mom, T, n, kdim, k, q); break;
        case 10: energy_rows<Real, 10>(x, mom, T, n, kdim, k, q); break;
        case 11: energy_rows<Real, 11>(x, mom, T, n, kdim, k, q); break;
        case 12: energy_rows<Real, 12>(x, mom, T, n, kdim, k, q); break;
        case 13: energy_rows<Real, 13>(x, mom, T, n, kdim, k, q); break;
        case 14: energy_rows<Real, 14>(x, mom, T, n, kdim, k, q); break;
        case 15: energy_rows<Real, 15>(x, mom, T, n, kdim, k, q); break;
        case 16: energy_rows<Real, 16>(x, mom, T, n, kdim, k, q); break;
        default: break;  // the samplers admit 1 <= P <= kMaxDim = 16
    }
    red[0][threadIdx.x] = k;
    red[1][threadIdx.x] = q;
    __syncthreads();
    for (int s = kEnergyThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0][0];
        out[1] = red[1][0];
    }
}

// The trace's rows (kTraceRows of them, indices in rows), widened to double:
// out[k * p + d], one thread per value.
template <typename Real>
__device__ __forceinline__ void gather_rows(const Real* __restrict__ x, int kdim, int p, const TraceRows& rows,
                                            double* __restrict__ out) {
    const int t = threadIdx.x;
    if (t < kTraceRows * p) out[t] = (double)x[(int64_t)rows.r[t / p] * kdim + t % p];
}

// Entry (a, b) of X^T X: the block sums x_ia x_ib over i in a fixed order
// (strided thread sums, then a fixed tree), so entries (a, b) and (b, a)
// agree bitwise.
constexpr int kGramThreads = 256;

template <typename Real>
__device__ __forceinline__ void gram_entry(const Real* __restrict__ x, int n, int kdim, int p, int a, int b,
                                           double* __restrict__ out, double* red) {
    // the loads of kAhead rows ahead of their additions, which keep the
    // thread's row order (the same bits as one row after another)
    constexpr int kAhead = 8;
    double acc = 0.0;
    for (int i0 = threadIdx.x; i0 < n; i0 += kAhead * kGramThreads) {
        double va[kAhead], vb[kAhead];
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            const int i = i0 + u * kGramThreads;
            va[u] = i < n ? (double)x[(int64_t)i * kdim + a] : 0.0;
            vb[u] = i < n ? (double)x[(int64_t)i * kdim + b] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            if (i0 + u * kGramThreads >= n) break;
            acc += va[u] * vb[u];
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kGramThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[a * p + b] = red[0];
}

}  // namespace mdship
