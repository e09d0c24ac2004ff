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
// reductions over one block of kEnergyThreads threads).
constexpr int kEnergyThreads = 512;

template <typename Real>
__device__ __forceinline__ void energy_block(const Real* __restrict__ x, const double* __restrict__ mom,
                                             const double* __restrict__ params, int n, int p, int kdim,
                                             double* __restrict__ out, double (*red)[kEnergyThreads]) {
    const double* T = params + kHmcTrait;
    double k = 0.0, q = 0.0;
    for (int i = threadIdx.x; i < n; i += kEnergyThreads) {
        const Real* xi = x + (int64_t)i * kdim;
        for (int d = 0; d < p; ++d) {
            const double m = mom[(int64_t)i * p + d];
            k += 0.5 * m * m;
            double row = 0.0;
            for (int e = 0; e < p; ++e) row += (double)xi[e] * T[e * p + d];
            q += row * (double)xi[d];
        }
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
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kGramThreads)
        acc += (double)x[(int64_t)i * kdim + a] * (double)x[(int64_t)i * kdim + b];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kGramThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[a * p + b] = red[0];
}

}  // namespace mdship
