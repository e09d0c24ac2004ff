// Device-resident Hamiltonian Monte Carlo for Bayesian MDS, the GPU
// counterpart of the reference's R sampler hmcsampler (R/Untitled.R:477-729)
// with Potential (:427-445) and the matrix-normal prior dmatrixnorm (:68-95)
// for the case without a tree (row covariance I_N):
//
//   potential U(X) = -(log prior(X | T) + log-likelihood(X, tau)),
//   log prior = -tr(T X^T X) / 2 - (N P / 2) log 2 pi + (N / 2) log det T,
//   log-likelihood = (log tau - log 2 pi) N (N - 1) / 4 - sumOfIncrements
//   (R/mds.R:24-28), momentum ~ N(0, I), leapfrog with step trajectory / L.
//
// Locations, momenta and gradients stay on the device; the trajectory of L
// leapfrog steps (kicks and drifts, the fused likelihood/gradient kernels)
// is captured once as a hipGraph and replayed once per iteration.  The precision and the
// trait precision live in a device parameter block, so the graph stays valid
// when they are updated.  The Hamiltonian and the Metropolis log-ratio are
// formed on the host in FP64.
//
// Between HMC moves:
//   * tau: with truncation, random-walk Metropolis with the reference's
//     proposal |U(tau - w, tau + w)| (R/Untitled.R:668-691), evaluated at the
//     current (accepted) locations; without truncation the flat-prior
//     posterior of tau is Gamma(1 + N(N-1)/4, C / 2) with C the unscaled
//     sum of squared residuals, and tau is drawn from it (Gibbs);
//   * T: Gibbs draw from Wishart(d0 + N, (T0 + X^T X)^-1) (R/Untitled.R:694-700),
//     d0 = P, T0 = I_P.
// Unlike the reference the current potential is recomputed after every tau
// or T update, and the tau step uses the current locations after a
// rejection.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "mds_datagen.hpp"
#include "mds_device.hpp"
#include "mds_engine.hpp"
#include "mds_hmc.hpp"
#include "mds_hmc_device.hpp"
#include "mds_mlp.hpp"

namespace mdship {

// ------------------------------------------------------------------ kernels

// The start of an iteration in one launch: standard normal momenta from a
// counter-based stream keyed by (seed, iteration), one pair per thread
// (momentum_pair), and copies of the locations and of the gradient the
// trajectory starts from (put back after a rejected move).
template <typename Real>
__global__ void start_kernel(double* __restrict__ mom, int count, uint64_t seed, uint64_t iteration,
                             const Real* __restrict__ x, Real* __restrict__ xSave, int64_t xCount,
                             const double* __restrict__ grad, double* __restrict__ gradSave, int64_t gCount) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    momentum_pair(mom, count, seed, iteration, (int)t);
    if (t < xCount) xSave[t] = x[t];
    if (t < gCount) gradSave[t] = grad[t];
}

// The copies back after a rejected move.
template <typename Real>
__global__ void restore_kernel(const Real* __restrict__ xSave, Real* __restrict__ x, int64_t xCount,
                               const double* __restrict__ gradSave, double* __restrict__ grad, int64_t gCount) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < xCount) x[t] = xSave[t];
    if (t < gCount) grad[t] = gradSave[t];
}

template <typename Real, int FORM>
__global__ void kick_kernel(Real* __restrict__ x, double* __restrict__ mom, const double* __restrict__ grad,
                            const double* __restrict__ params, int n, int p, int kdim) {
    kick_row<Real, FORM>(x, mom, grad, params, n, p, kdim, blockIdx.x * blockDim.x + threadIdx.x);
}

// out[0..1] as energy_block; with sumIn, also *sumOut = *sumIn (the sum of
// the pass before, carried to the host with the energies).
template <typename Real>
__global__ __launch_bounds__(kEnergyThreads) void energy_kernel(const Real* __restrict__ x,
                                                                const double* __restrict__ mom,
                                                                const double* __restrict__ params, int n, int p,
                                                                int kdim, double* __restrict__ out,
                                                                const double* __restrict__ sumIn,
                                                                double* __restrict__ sumOut) {
    __shared__ double red[2][kEnergyThreads];
    energy_block<Real>(x, mom, params, n, p, kdim, out, red);
    if (sumIn && threadIdx.x == 0) *sumOut = *sumIn;
}

template <typename Real>
__global__ void gather_rows_kernel(const Real* __restrict__ x, int kdim, int p, TraceRows rows,
                                   double* __restrict__ out) {
    gather_rows(x, kdim, p, rows, out);
}

template <typename Real>
__global__ __launch_bounds__(kGramThreads) void gram_kernel(const Real* __restrict__ x, int n, int kdim, int p,
                                                            double* __restrict__ out) {
    __shared__ double red[kGramThreads];
    gram_entry<Real>(x, n, kdim, p, blockIdx.x / p, blockIdx.x % p, out, red);
}

// ------------------------------------------------------------------ host

// Marsaglia & Tsang (2000), "A simple method for generating gamma variables".
double gamma_draw(mdsgen::Stream& s, double shape) {
    if (shape < 1.0) {
        const double u = s.uniform();
        return gamma_draw(s, shape + 1.0) * std::pow(u, 1.0 / shape);
    }
    const double d = shape - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
    for (;;) {
        double z, v;
        do {
            const double u1 = s.uniform(), u2 = s.uniform();
            z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
            v = 1.0 + c * z;
        } while (v <= 0.0);
        v = v * v * v;
        const double u = s.uniform();
        if (std::log(u) < 0.5 * z * z + d - d * v + d * std::log(v)) return d * v;
    }
}

// Cholesky factor (lower) of a symmetric positive-definite p x p matrix.
bool cholesky(const std::vector<double>& a, int p, std::vector<double>& l) {
    l.assign((size_t)p * p, 0.0);
    for (int i = 0; i < p; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = a[i * p + j];
            for (int k = 0; k < j; ++k) s -= l[i * p + k] * l[j * p + k];
            if (i == j) {
                if (s <= 0.0) return false;
                l[i * p + i] = std::sqrt(s);
            } else {
                l[i * p + j] = s / l[j * p + j];
            }
        }
    return true;
}

std::vector<double> spd_inverse(const std::vector<double>& a, int p) {
    std::vector<double> l;
    if (!cholesky(a, p, l)) throw Error(MDS_HIP_ERR_STATE, "matrix not positive definite");
    // inverse of L, then (L^-1)^T L^-1
    std::vector<double> li((size_t)p * p, 0.0);
    for (int j = 0; j < p; ++j) {
        li[j * p + j] = 1.0 / l[j * p + j];
        for (int i = j + 1; i < p; ++i) {
            double s = 0.0;
            for (int k = j; k < i; ++k) s -= l[i * p + k] * li[k * p + j];
            li[i * p + j] = s / l[i * p + i];
        }
    }
    std::vector<double> inv((size_t)p * p, 0.0);
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j) {
            double s = 0.0;
            for (int k = std::max(i, j); k < p; ++k) s += li[k * p + i] * li[k * p + j];
            inv[i * p + j] = s;
        }
    return inv;
}

double log_det_spd(const std::vector<double>& a, int p) {
    std::vector<double> l;
    if (!cholesky(a, p, l)) throw Error(MDS_HIP_ERR_STATE, "matrix not positive definite");
    double s = 0.0;
    for (int i = 0; i < p; ++i) s += 2.0 * std::log(l[i * p + i]);
    return s;
}

// Wishart(df, scale) by the Bartlett decomposition: with scale = L L^T and
// A lower triangular, A_ii^2 ~ chi^2(df - i), A_ij ~ N(0, 1) below the
// diagonal, W = L A A^T L^T.
std::vector<double> wishart_draw(mdsgen::Stream& s, double df, const std::vector<double>& scale, int p) {
    std::vector<double> l;
    if (!cholesky(scale, p, l)) throw Error(MDS_HIP_ERR_STATE, "Wishart scale not positive definite");
    std::vector<double> A((size_t)p * p, 0.0);
    for (int i = 0; i < p; ++i) {
        A[i * p + i] = std::sqrt(2.0 * gamma_draw(s, 0.5 * (df - i)));
        for (int j = 0; j < i; ++j) {
            const double u1 = s.uniform(), u2 = s.uniform();
            A[i * p + j] = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
        }
    }
    std::vector<double> la((size_t)p * p, 0.0);  // L A
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j) {
            double v = 0.0;
            for (int k = 0; k < p; ++k) v += l[i * p + k] * A[k * p + j];
            la[i * p + j] = v;
        }
    std::vector<double> w((size_t)p * p, 0.0);
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j) {
            double v = 0.0;
            for (int k = 0; k < p; ++k) v += la[i * p + k] * la[j * p + k];
            w[i * p + j] = v;
        }
    return w;
}

Hmc::Hmc(Engine& engine, const mds_hip_hmc_options& options) : e(engine), opt(options) {
    if (!e.dataSet) throw Error(MDS_HIP_ERR_STATE, "pairwise data not set");
    if (!e.precisionSet) throw Error(MDS_HIP_ERR_STATE, "precision not set");
    if (opt.iterations < 1 || opt.leapfrogSteps < 1 || opt.burnIn < 0 || opt.burnIn >= opt.iterations)
        throw Error(MDS_HIP_ERR_ARGUMENT, "invalid iteration counts");
    if (!(opt.trajectory > 0.0)) throw Error(MDS_HIP_ERR_ARGUMENT, "trajectory must be > 0");
    if (opt.traitInvWeight < 1) throw Error(MDS_HIP_ERR_ARGUMENT, "traitInvWeight must be >= 1");
    n = e.N;
    p = e.P;
    if (opt.useSurrogate || opt.collectSnapshots) {
        if (!opt.surrogate || !opt.surrogate->s)
            throw Error(MDS_HIP_ERR_ARGUMENT, "surrogate handle required for useSurrogate / collectSnapshots");
        sur = opt.surrogate->s;
        if (sur->mlp.n != n || sur->mlp.p != p)
            throw Error(MDS_HIP_ERR_ARGUMENT, "surrogate built for another N or P");
        if (sur->mlp.device != e.device) throw Error(MDS_HIP_ERR_ARGUMENT, "surrogate on another device");
    }
    const size_t np = (size_t)n * p;
    const size_t xBytes = (size_t)e.npad * e.kdim * e.realSize;
    try {
        check(hipMalloc((void**)&dMom, np * sizeof(double)), "hipMalloc(momentum)");
        check(hipMalloc((void**)&dGradSave, np * sizeof(double)), "hipMalloc(gradient copy)");
        check(hipMalloc(&dXSave, xBytes), "hipMalloc(location copy)");
        check(hipMalloc((void**)&dParams, kHmcParamCount * sizeof(double)), "hipMalloc(parameters)");
        check(hipHostMalloc((void**)&hParams, kHmcParamCount * sizeof(double), hipHostMallocDefault),
              "hipHostMalloc");
        check(hipHostMalloc(&hX, xBytes, hipHostMallocDefault), "hipHostMalloc");
        // written by kernels (coherent pinned memory, read after a stream wait)
        check(hipHostMalloc((void**)&hScalars, kScalarSlots * sizeof(double), hipHostMallocCoherent),
              "hipHostMalloc");
        check(hipHostMalloc((void**)&hRows, kTraceRows * kMaxDim * sizeof(double), hipHostMallocCoherent),
              "hipHostMalloc");
        check(hipHostMalloc((void**)&hGram, kMaxDim * kMaxDim * sizeof(double), hipHostMallocCoherent),
              "hipHostMalloc");
        check(hipHostGetDevicePointer((void**)&kScalars, hScalars, 0), "hipHostGetDevicePointer");
        check(hipHostGetDevicePointer((void**)&kRows, hRows, 0), "hipHostGetDevicePointer");
        check(hipHostGetDevicePointer((void**)&kGram, hGram, 0), "hipHostGetDevicePointer");
        check(hipEventCreateWithFlags(&uploadDone, hipEventDisableTiming), "hipEventCreate");
    } catch (...) {
        release();
        throw;
    }
    traceRows = trace_rows(n);
    std::memset(hParams, 0, kHmcParamCount * sizeof(double));
    tau = e.precision;
    trait.assign((size_t)p * p, 0.0);
    for (int d = 0; d < p; ++d) trait[d * p + d] = 1.0;  // solve(traitVcv) with traitVcv = I_P
    eps = opt.stepSize > 0.0 ? opt.stepSize : opt.trajectory / opt.leapfrogSteps;
}

Hmc::~Hmc() { release(); }

// Frees what the constructor allocated (also after a failure part way).
void Hmc::release() {
    (void)hipStreamSynchronize(e.stream);
    for (int m = 0; m < 3; ++m) {
        if (graphExec[m]) (void)hipGraphExecDestroy(graphExec[m]);
        if (graph[m]) (void)hipGraphDestroy(graph[m]);
        graphExec[m] = nullptr;
        graph[m] = nullptr;
    }
    (void)hipFree(dMom);
    (void)hipFree(dGradSave);
    (void)hipFree(dXSave);
    (void)hipFree(dParams);
    (void)hipHostFree(hParams);
    (void)hipHostFree(hScalars);
    (void)hipHostFree(hX);
    (void)hipHostFree(hRows);
    (void)hipHostFree(hGram);
    if (uploadDone) (void)hipEventDestroy(uploadDone);
    dMom = dGradSave = dParams = nullptr;
    dXSave = hX = nullptr;
    hParams = hScalars = hRows = hGram = kScalars = kRows = kGram = nullptr;
    uploadDone = nullptr;
}

// The pinned block is rewritten only after the previous upload from it has
// run (the copy reads host memory when it executes, not when it is queued).
void Hmc::uploadParams() {
    if (uploadPending) check(hipEventSynchronize(uploadDone), "hipEventSynchronize(parameters)");
    paramsStale = false;
    hParams[kHmcTau] = tau;
    hParams[kHmcEps] = eps;
    for (int k = 0; k < p * p; ++k) hParams[kHmcTrait + k] = trait[k];
    check(hipMemcpyAsync(dParams, hParams, kHmcParamCount * sizeof(double), hipMemcpyHostToDevice, e.stream),
          "hipMemcpy(parameters)");
    check(hipEventRecord(uploadDone, e.stream), "hipEventRecord(parameters)");
    uploadPending = true;
}

template <int FORM>
void Hmc::kick() {
    const int threads = 256, blocks = (n + threads - 1) / threads;
    if (e.fp32)
        hipLaunchKernelGGL((kick_kernel<float, FORM>), dim3(blocks), dim3(threads), 0, e.stream,
                           (float*)e.dX[e.cur], dMom, e.dGrad, dParams, n, p, e.kdim);
    else
        hipLaunchKernelGGL((kick_kernel<double, FORM>), dim3(blocks), dim3(threads), 0, e.stream,
                           (double*)e.dX[e.cur], dMom, e.dGrad, dParams, n, p, e.kdim);
    check(hipGetLastError(), "kick kernel");
}

// One trajectory of opt.leapfrogSteps steps from the gradient now in
// e.dGrad: kick + drift, then per step the gradient and the kicks.  With
// exact gradients the reduce applies each step's kicks to the rows it has
// just written (the same kick_row arithmetic as the kick kernel, so the same
// bits), one launch fewer per step; collecting snapshots needs the gradient
// at the locations before the drift, and the surrogate has no reduce.
void Hmc::trajectory(int mode) {
    kick<kKickDrift>();
    for (int s = 0; s < opt.leapfrogSteps; ++s) {
        const int form = s + 1 < opt.leapfrogSteps ? kKick2Drift : kKick;
        if (mode == kExact) {
            ReduceKick k;
            k.form = form;
            k.mom = dMom;
            k.params = dParams;
            e.evaluate(true, dParams + kHmcTau, k);
            continue;
        }
        gradient(mode);
        if (form == kKick2Drift)
            kick<kKick2Drift>();
        else
            kick<kKick>();
    }
}

void Hmc::gradientAtCurrent() { e.evaluate(true, dParams + kHmcTau); }

// Gradient of the log-likelihood at the current locations into e.dGrad:
// exact (fused kernels; with kExactCollect the (x, gradient) pair is also
// stored for the surrogate's training), or the surrogate's prediction.
void Hmc::gradient(int mode) {
    if (mode == kSurrogate) {
        if (e.fp32)
            sur->mlp.predictInto<float>((const float*)e.dX[e.cur], e.kdim, e.dGrad, e.stream);
        else
            sur->mlp.predictInto<double>((const double*)e.dX[e.cur], e.kdim, e.dGrad, e.stream);
        return;
    }
    gradientAtCurrent();
    if (mode == kExactCollect)
        launch_snapshot(e, e.dGrad, sur->store.sx, sur->store.sg, sur->store.dSlot, sur->store.capacity, e.stream);
}

// The whole trajectory as one graph: one launch per iteration.  Every exit
// ends the capture and restores the engine's timing switch, so a failure
// while capturing leaves the stream usable.
void Hmc::buildGraph(int mode) {
    if (graphExec[mode]) return;
    const bool timing = e.timing;
    e.timing = false;
    const hipError_t begun = hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal);
    if (begun != hipSuccess) {
        e.timing = timing;
        check(begun, "hipStreamBeginCapture");
    }
    hipGraph_t g = nullptr;
    try {
        trajectory(mode);
    } catch (...) {
        if (hipStreamEndCapture(e.stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        e.timing = timing;
        throw;
    }
    const hipError_t ended = hipStreamEndCapture(e.stream, &g);
    e.timing = timing;
    check(ended, "hipStreamEndCapture");
    graph[mode] = g;
    check(hipGraphInstantiate(&graphExec[mode], graph[mode], nullptr, nullptr, 0), "hipGraphInstantiate");
}

// energies at the current locations: hScalars[kSlotK] = K, [kSlotQuad] =
// x^T T x sum, [kSlotSum] = the last pass's sumOfIncrements (0.5 * sum of
// tile partials); enqueueEnergies leaves the wait to the caller
void Hmc::readEnergies() {
    enqueueEnergies();
    check(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
}

void Hmc::enqueueEnergies() { launchEnergy(kSlotK, true); }

// The energy kernel at the current locations and momenta into the slots
// from `slot` on; withSum: the last pass's sum into kSlotSum as well.
void Hmc::launchEnergy(int slot, bool withSum) {
    const double* sumIn = withSum ? e.dSum : nullptr;
    if (e.fp32)
        hipLaunchKernelGGL((energy_kernel<float>), dim3(1), dim3(kEnergyThreads), 0, e.stream,
                           (const float*)e.dX[e.cur], dMom, dParams, n, p, e.kdim, kScalars + slot, sumIn,
                           kScalars + kSlotSum);
    else
        hipLaunchKernelGGL((energy_kernel<double>), dim3(1), dim3(kEnergyThreads), 0, e.stream,
                           (const double*)e.dX[e.cur], dMom, dParams, n, p, e.kdim, kScalars + slot, sumIn,
                           kScalars + kSlotSum);
    check(hipGetLastError(), "energy kernel");
}

// The work the rest of an iteration does at the locations the HMC step
// keeps, enqueued ahead of the wait on the HMC step's energies: the trace
// rows (record), X^T X (trait step) and, for tauStar > 0, the likelihood at
// the precision proposal tauStar into hScalars[kSlotSumStar] (tauStar passed
// to the kernels by value: the device parameter block keeps tau).
void Hmc::enqueueAtKept(bool traitStep, bool recordIt, double tauStar) {
    if (traitStep) gatherGram();
    if (recordIt) gatherTraceRows();
    if (tauStar > 0.0) {
        PassOut out;
        out.tau = tauStar;
        out.sum = kScalars + kSlotSumStar;
        e.evaluate(false, nullptr, ReduceKick(), out);
    }
}

double mds_log_likelihood(double rawSum, double t, int n, bool truncated) {
    // rawSum = 0.5 * sum of tile partials; truncated: sumOfIncrements itself,
    // otherwise scaled by 0.5 tau (src/NewMultiDimensionalScaling.hpp:198-202)
    const double s = truncated ? rawSum : 0.5 * t * rawSum;
    const double pairs = (double)n * (n - 1) / 2.0;
    return 0.5 * (std::log(t) - std::log(2.0 * M_PI)) * pairs - s;
}

double matrix_normal_log_prior(double quad, int n, int p, const std::vector<double>& trait) {
    return -0.5 * quad - 0.5 * n * p * std::log(2.0 * M_PI) + 0.5 * n * log_det_spd(trait, p);
}

double Hmc::logLikelihood(double rawSum, double t) const { return mds_log_likelihood(rawSum, t, n, e.truncated); }

double Hmc::logPrior(double quad) const { return matrix_normal_log_prior(quad, n, p, trait); }

// Enqueues the trace rows (record) and X^T X (trait) of the current
// locations, written by the kernels into pinned host memory; the caller
// synchronises before reading.
void Hmc::gatherTraceRows() {
    if (e.fp32)
        hipLaunchKernelGGL(gather_rows_kernel<float>, dim3(1), dim3(kTraceRows * kMaxDim), 0, e.stream,
                           (const float*)e.dX[e.cur], e.kdim, p, traceRows, kRows);
    else
        hipLaunchKernelGGL(gather_rows_kernel<double>, dim3(1), dim3(kTraceRows * kMaxDim), 0, e.stream,
                           (const double*)e.dX[e.cur], e.kdim, p, traceRows, kRows);
    check(hipGetLastError(), "trace rows kernel");
}

void Hmc::startIteration(int it) {
    const int count = n * p;
    const int64_t xCount = (int64_t)e.npad * e.kdim, gCount = (int64_t)n * p;
    const int64_t work = std::max<int64_t>(std::max<int64_t>((count + 1) / 2, xCount), gCount);
    constexpr int kThreads = 256;
    const dim3 grid((unsigned)((work + kThreads - 1) / kThreads));
    if (e.fp32)
        hipLaunchKernelGGL(start_kernel<float>, grid, dim3(kThreads), 0, e.stream, dMom, count, (uint64_t)opt.seed,
                           (uint64_t)it, (const float*)e.dX[e.cur], (float*)dXSave, xCount, e.dGrad, dGradSave,
                           gCount);
    else
        hipLaunchKernelGGL(start_kernel<double>, grid, dim3(kThreads), 0, e.stream, dMom, count, (uint64_t)opt.seed,
                           (uint64_t)it, (const double*)e.dX[e.cur], (double*)dXSave, xCount, e.dGrad, dGradSave,
                           gCount);
    check(hipGetLastError(), "start kernel");
}

void Hmc::restoreStart() {
    const int64_t xCount = (int64_t)e.npad * e.kdim, gCount = (int64_t)n * p;
    constexpr int kThreads = 256;
    const dim3 grid((unsigned)((std::max(xCount, gCount) + kThreads - 1) / kThreads));
    if (e.fp32)
        hipLaunchKernelGGL(restore_kernel<float>, grid, dim3(kThreads), 0, e.stream, (const float*)dXSave,
                           (float*)e.dX[e.cur], xCount, dGradSave, e.dGrad, gCount);
    else
        hipLaunchKernelGGL(restore_kernel<double>, grid, dim3(kThreads), 0, e.stream, (const double*)dXSave,
                           (double*)e.dX[e.cur], xCount, dGradSave, e.dGrad, gCount);
    check(hipGetLastError(), "restore kernel");
}

void Hmc::gatherGram() {
    if (e.fp32)
        hipLaunchKernelGGL(gram_kernel<float>, dim3(p * p), dim3(kGramThreads), 0, e.stream, (const float*)e.dX[e.cur],
                           n, e.kdim, p, kGram);
    else
        hipLaunchKernelGGL(gram_kernel<double>, dim3(p * p), dim3(kGramThreads), 0, e.stream,
                           (const double*)e.dX[e.cur], n, e.kdim, p, kGram);
    check(hipGetLastError(), "gram kernel");
}

void Hmc::copyLocationsToHost(std::vector<double>& out) {
    const size_t bytes = (size_t)n * e.kdim * e.realSize;
    check(hipMemcpyAsync(hX, e.dX[e.cur], bytes, hipMemcpyDeviceToHost, e.stream), "hipMemcpy(locations)");
    check(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
    out.resize((size_t)n * p);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < p; ++d)
            out[(size_t)i * p + d] = e.fp32 ? (double)((const float*)hX)[(size_t)i * e.kdim + d]
                                            : ((const double*)hX)[(size_t)i * e.kdim + d];
}

// rows: the trace rows in the order of traceRows (0, 1, n/4, n/3, n/2,
// 2n/3, 3n/4, n - 1), p values each
void Hmc::record(double* row, const double* rows, double logPost) const {
    write_trace_row(row, rows, p, logPost, tau, trait);
}

TraceRows trace_rows(int n) {
    TraceRows t;
    const int r[kTraceRows] = {0, 1 % n, n / 4, n / 3, n / 2, 2 * n / 3, 3 * n / 4, n - 1};
    for (int k = 0; k < kTraceRows; ++k) t.r[k] = r[k];
    return t;
}

void write_trace_row(double* row, const double* rows, int p, double logPost, double tau,
                     const std::vector<double>& trait) {
    row[0] = logPost;
    row[1] = tau;
    double tr = 0.0;
    for (int d = 0; d < p; ++d) tr += trait[d * p + d];
    row[2] = tr;
    // distances 0-1, 0-n/2, n/3-2n/3, 1-(n-1), n/4-3n/4; norms of 0, n/2, n-1
    const int pairs[5][2] = {{0, 1}, {0, 4}, {3, 5}, {1, 7}, {2, 6}};
    for (int k = 0; k < 5; ++k) {
        double s = 0.0;
        for (int d = 0; d < p; ++d) {
            const double v = rows[pairs[k][0] * p + d] - rows[pairs[k][1] * p + d];
            s += v * v;
        }
        row[3 + k] = std::sqrt(s);
    }
    const int norms[3] = {0, 4, 7};
    for (int k = 0; k < 3; ++k) {
        double s = 0.0;
        for (int d = 0; d < p; ++d) s += rows[norms[k] * p + d] * rows[norms[k] * p + d];
        row[8 + k] = std::sqrt(s);
    }
}

void Hmc::run(const double* x0, double* trace, double* finalX, mds_hip_hmc_stats* stats) {
    using clock = std::chrono::steady_clock;
    const auto tStart = clock::now();
    mdsgen::Stream rng(opt.seed, 77);
    e.updateLocations(-1, x0, (size_t)n * p);
    uploadParams();
    const int startSnapshots = sur ? sur->store.count() : 0;
    const int exactMode = opt.collectSnapshots ? kExactCollect : kExact;
    int gradMode = exactMode;  // source of the gradient now in e.dGrad
    gradient(exactMode);
    check(hipMemsetAsync(dMom, 0, (size_t)n * p * sizeof(double), e.stream), "hipMemset(momentum)");
    readEnergies();
    double logLik = logLikelihood(hScalars[kSlotSum], tau);
    double quad = hScalars[kSlotQuad];
    double U = -(logLik + logPrior(quad));

    // dual averaging of the step size (Hoffman & Gelman 2014, sec. 3.2)
    const double daMu = std::log(10.0 * eps), daGamma = 0.05, daT0 = 10.0, daKappa = 0.75;
    double daH = 0.0, daLogEpsBar = 0.0;

    long long accepted = 0, proposedTau = 0, acceptedTau = 0, evaluations = 1;
    long long surrogateIts = 0, exactIts = 0, surrogateAccepted = 0;
    int windowIts = 0, windowAcc = 0, fallbackLeft = 0, failedWindows = 0;
    PrecisionWidth width;
    width.w = opt.precisionProposalWidth;
    width.adapt = opt.adaptPrecisionWidth != 0;
    auto tSampling = clock::now();
    for (int it = 0; it < opt.iterations; ++it) {
        if (it == opt.burnIn) {
            tSampling = clock::now();
            if (opt.adaptStepSize && opt.burnIn > 0) {
                eps = std::exp(daLogEpsBar);
                uploadParams();
            }
        }
        const int mode = (opt.useSurrogate && fallbackLeft == 0) ? kSurrogate : exactMode;
        if (mode != gradMode && !(mode != kSurrogate && gradMode != kSurrogate)) {
            gradient(mode);  // the trajectory starts from this source's gradient
            if (mode != kSurrogate) ++evaluations;
        }
        gradMode = mode;
        if (opt.useGraph) buildGraph(mode);
        // the momenta and the copies a rejection restores, then the starting
        // kinetic energy, read with the energies after the trajectory (no
        // wait here)
        startIteration(it);
        launchEnergy(kSlotK0, false);
        if (mode == kExactCollect)  // snapshots of a rejected trajectory are dropped
            check(hipMemcpyAsync(sur->store.dSlot + 1, sur->store.dSlot, sizeof(double), hipMemcpyDeviceToDevice,
                                 e.stream), "hipMemcpy(slot)");
        if (graphExec[mode])
            check(hipGraphLaunch(graphExec[mode], e.stream), "hipGraphLaunch");
        else
            trajectory(mode);
        if (mode == kSurrogate) {
            // the exact likelihood (FP64-reduced) judges the proposal
            e.evaluate(false, dParams + kHmcTau);
            ++evaluations;
            ++surrogateIts;
        } else {
            evaluations += opt.leapfrogSteps;
            ++exactIts;
        }
        enqueueEnergies();
        // the trace rows and X^T X of the kept locations (the precision step
        // does not move them)
        const bool traitStep = opt.learnTraitPrecision && it < opt.iterations - 1 &&
                               ((it + 1) % opt.traitInvWeight == 0);
        const bool recordIt = it >= opt.burnIn && trace;
        // The random draws up to the precision's Metropolis test do not depend
        // on the device's results, so they are made here, in the sampler's
        // order (the HMC test's uniform, the precision proposal, its test's
        // uniform), and with exact gradients and a truncated likelihood the
        // rest of the iteration's device work is enqueued at the proposed
        // locations before the one wait: kept in ~93 % of iterations, redone
        // at the restored locations otherwise.  Every value is the one the
        // steps would give in turn.
        const bool speculate = mode != kSurrogate && opt.learnPrecision && e.truncated;
        const double u = rng.uniform();
        double tauStar = 0.0, uTau = 0.0;
        if (speculate) {
            const double lo = tau - width.w, hi = tau + width.w;
            tauStar = std::fabs(lo + (hi - lo) * rng.uniform());
            if (tauStar > 0.0) uTau = rng.uniform();
            enqueueAtKept(traitStep, recordIt, tauStar);
        }
        check(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
        const double K0 = hScalars[kSlotK0];
        const double K1 = hScalars[kSlotK];
        const double logLik1 = logLikelihood(hScalars[kSlotSum], tau);
        const double quad1 = hScalars[kSlotQuad];
        const double U1 = -(logLik1 + logPrior(quad1));
        const double ratio = (U + K0) - (U1 + K1);  // R/Untitled.R:630-637
        const bool acc = std::isfinite(ratio) && ratio > std::min(0.0, std::log(u));
        if (acc && mode == kSurrogate) ++surrogateAccepted;
        if (acc) {
            U = U1;
            logLik = logLik1;
            quad = quad1;
            ++accepted;
        } else {
            restoreStart();
            if (mode == kExactCollect)
                check(hipMemcpyAsync(sur->store.dSlot, sur->store.dSlot + 1, sizeof(double), hipMemcpyDeviceToDevice,
                                     e.stream), "hipMemcpy(slot)");
            if (speculate) {
                enqueueAtKept(traitStep, recordIt, tauStar);
                check(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
            }
        }
        if (opt.adaptStepSize && it < opt.burnIn) {
            const double alpha = std::isfinite(ratio) ? std::min(1.0, std::exp(std::min(ratio, 0.0))) : 0.0;
            const double m = it + 1.0;
            daH = (1.0 - 1.0 / (m + daT0)) * daH + (opt.targetAccept - alpha) / (m + daT0);
            const double logEps = daMu - std::sqrt(m) / daGamma * daH;
            const double eta = std::pow(m, -daKappa);
            daLogEpsBar = eta * logEps + (1.0 - eta) * daLogEpsBar;
            eps = std::exp(logEps);
            uploadParams();
        }
        if (opt.useSurrogate && opt.fallbackWindow > 0) {
            if (fallbackLeft > 0) {
                --fallbackLeft;
            } else {
                ++windowIts;
                windowAcc += acc ? 1 : 0;
                if (windowIts == opt.fallbackWindow) {
                    // a failed window hands over to exact gradients for a
                    // window that doubles with each consecutive failure
                    if ((double)windowAcc / windowIts < opt.fallbackAcceptRate) {
                        fallbackLeft = opt.fallbackWindow << std::min(failedWindows, 10);
                        ++failedWindows;
                    } else {
                        failedWindows = 0;
                    }
                    windowIts = windowAcc = 0;
                }
            }
        }

        if (!speculate) {  // read after the next wait
            if (traitStep) gatherGram();
            if (recordIt) gatherTraceRows();
        }

        if (speculate) {
            // random-walk Metropolis at the kept locations, from the pass made
            // ahead (its likelihood in hScalars[kSlotSumStar])
            ++proposedTau;
            if (tauStar > 0.0) {
                ++evaluations;
                const double logLikStar = logLikelihood(hScalars[kSlotSumStar], tauStar);
                const double r = logLikStar - logLik;  // the prior does not involve tau
                const bool accTau = std::isfinite(r) && r > std::min(0.0, std::log(uTau));
                width.update(accTau, it < opt.burnIn);
                if (accTau) {
                    tau = tauStar;
                    logLik = logLikStar;
                    U = -(logLik + logPrior(quad));
                    ++acceptedTau;
                    // the gradient at the new tau, passed by value; the device
                    // parameter block follows at the end of the iteration
                    PassOut out;
                    out.tau = tau;
                    e.evaluate(true, nullptr, ReduceKick(), out);
                    paramsStale = true;
                    ++evaluations;
                }
            }
        } else if (opt.learnPrecision) {
            if (gradMode == kSurrogate) {
                gradient(exactMode);  // precision moves use the exact likelihood and gradient
                gradMode = exactMode;
                ++evaluations;
            }
            if (e.truncated) {
                // random-walk Metropolis at the current locations
                const double lo = tau - width.w, hi = tau + width.w;
                const double tauStar = std::fabs(lo + (hi - lo) * rng.uniform());
                ++proposedTau;
                if (tauStar > 0.0) {
                    const double saved = tau;
                    tau = tauStar;
                    uploadParams();
                    e.evaluate(false, dParams + kHmcTau);
                    ++evaluations;
                    check(hipMemcpyAsync(hScalars + kSlotSum, e.dSum, sizeof(double), hipMemcpyDeviceToHost,
                                         e.stream),
                          "hipMemcpy");
                    check(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
                    const double logLikStar = logLikelihood(hScalars[kSlotSum], tauStar);
                    const double r = logLikStar - logLik;  // the prior does not involve tau
                    const bool accTau = std::isfinite(r) && r > std::min(0.0, std::log(rng.uniform()));
                    width.update(accTau, it < opt.burnIn);
                    if (accTau) {
                        logLik = logLikStar;
                        U = -(logLik + logPrior(quad));
                        ++acceptedTau;
                        gradientAtCurrent();  // gradient at the new tau
                        ++evaluations;
                    } else {
                        // the likelihood-only pass left the gradient at the old
                        // tau in place
                        tau = saved;
                        uploadParams();
                    }
                }
            } else {
                // Gibbs: tau | X ~ Gamma(1 + N(N-1)/4, C / 2) with C the unscaled
                // sum of squared residuals (this engine's raw sum at the current X)
                gradientAtCurrent();
                ++evaluations;
                check(hipMemcpyAsync(hScalars + kSlotSum, e.dSum, sizeof(double), hipMemcpyDeviceToHost, e.stream),
                      "hipMemcpy");
                check(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
                const double C = hScalars[kSlotSum];
                const double pairs = (double)n * (n - 1) / 2.0;
                tau = gamma_draw(rng, 1.0 + 0.5 * pairs) / (0.5 * C);
                ++proposedTau;
                ++acceptedTau;
                uploadParams();
                gradientAtCurrent();
                ++evaluations;
                logLik = logLikelihood(C, tau);
                U = -(logLik + logPrior(quad));
            }
        }

        if (!speculate && (traitStep || recordIt)) check(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
        if (traitStep) {
            std::vector<double> m((size_t)p * p, 0.0);  // T0 + X^T X, T0 = I_P
            for (int d = 0; d < p; ++d) m[d * p + d] = 1.0;
            for (int k = 0; k < p * p; ++k) m[k] += hGram[k];
            const double df = (opt.priorDf > 0 ? opt.priorDf : p) + n;
            trait = wishart_draw(rng, df, spd_inverse(m, p), p);
            uploadParams();
            quad = 0.0;  // sum_i x_i^T T x_i = sum_ab T_ab (X^T X)_ab
            for (int k = 0; k < p * p; ++k) quad += trait[k] * hGram[k];
            U = -(logLik + logPrior(quad));
        }
        if (paramsStale) uploadParams();  // an accepted precision without a trait step
        if (recordIt) record(trace + (size_t)(it - opt.burnIn) * kHmcTraceColumns, hRows, -U);
    }
    check(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
    const auto tEnd = clock::now();
    if (finalX) {
        std::vector<double> xHost;
        copyLocationsToHost(xHost);
        std::memcpy(finalX, xHost.data(), xHost.size() * sizeof(double));
    }
    // leave the engine consistent: new precision, caches recomputed on demand
    e.precision = tau;
    e.makeDirty();
    if (stats) {
        stats->acceptRate = (double)accepted / opt.iterations;
        stats->precisionAcceptRate = proposedTau ? (double)acceptedTau / proposedTau : 0.0;
        stats->samplingSeconds = std::chrono::duration<double>(tEnd - tSampling).count();
        stats->totalSeconds = std::chrono::duration<double>(tEnd - tStart).count();
        stats->evaluations = (double)evaluations;
        stats->finalPrecision = tau;
        stats->stepSize = eps;
        stats->finalPotential = U;
        stats->surrogateIterations = (double)surrogateIts;
        stats->exactIterations = (double)exactIts;
        stats->snapshots = sur ? (double)(sur->store.count() - startSnapshots) : 0.0;
        stats->surrogateAcceptRate = surrogateIts ? (double)surrogateAccepted / surrogateIts : 0.0;
        stats->precisionProposalWidth = width.w;
    }
}

}  // namespace mdship

extern "C" {

int mds_hip_hmc_default_options(mds_hip_hmc_options* o) {
    if (!o) return MDS_HIP_ERR_ARGUMENT;
    o->iterations = 2000;
    o->burnIn = 500;
    o->leapfrogSteps = 20;
    o->trajectory = 0.1;
    o->learnPrecision = 1;
    o->learnTraitPrecision = 1;
    o->traitInvWeight = 1;
    o->precisionProposalWidth = 0.01;
    o->priorDf = 0.0;
    o->seed = 666;
    o->useGraph = 1;
    o->adaptStepSize = 0;
    o->targetAccept = 0.8;
    o->stepSize = 0.0;
    o->surrogate = nullptr;
    o->useSurrogate = 0;
    o->collectSnapshots = 0;
    o->fallbackAcceptRate = 0.2;
    o->fallbackWindow = 20;
    o->adaptPrecisionWidth = 0;
    return MDS_HIP_OK;
}

int mds_hip_hmc_run(mds_hip_handle h, const mds_hip_hmc_options* options, const double* initialLocations,
                    size_t length, double* trace, size_t traceLength, double* finalLocations,
                    mds_hip_hmc_stats* stats) {
    if (!h || !h->e) return MDS_HIP_ERR_ARGUMENT;
    mdship::Engine& e = *h->e;
    try {
        (void)hipSetDevice(e.device);
        if (!options || !initialLocations || length != (size_t)e.N * e.P)
            throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "hmc: options and N*P initial locations required");
        if (trace && traceLength != (size_t)(options->iterations - options->burnIn) * mdship::kHmcTraceColumns)
            throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "hmc: trace length must be (iterations - burnIn) * columns");
        mdship::Hmc hmc(e, *options);
        hmc.run(initialLocations, trace, finalLocations, stats);
        return MDS_HIP_OK;
    } catch (const mdship::Error& err) {
        e.lastError = err.what();
        return err.code;
    } catch (const std::exception& err) {
        e.lastError = err.what();
        return MDS_HIP_ERR_DEVICE;
    }
}

}  // extern "C"
