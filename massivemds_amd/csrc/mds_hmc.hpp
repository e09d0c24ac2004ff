// Device-resident HMC sampler over an Engine (see mds_hmc.hip).
#pragma once

#include <cmath>
#include <vector>

#include "mds_datagen.hpp"
#include "mds_engine.hpp"
#include "mds_hip.h"
#include "mds_mlp.hpp"

namespace mdship {

// Device parameter block: precision, step size, trait precision (P x P).
constexpr int kHmcTau = 0;
constexpr int kHmcEps = 1;
constexpr int kHmcTrait = 2;
constexpr int kHmcParamCount = kHmcTrait + kMaxDim * kMaxDim;
constexpr int kHmcTraceColumns = MDS_HIP_HMC_TRACE_COLUMNS;

// The locations the trace records (R/Untitled.R's invariant functionals):
// rows 0, 1, n/4, n/3, n/2, 2n/3, 3n/4, n - 1.
constexpr int kTraceRows = 8;
struct TraceRows {
    int r[kTraceRows];
};

// Width w of the precision's random-walk proposal |U(tau - w, tau + w)|
// (R/Untitled.R:668-691).  With adaptation, during burn-in only, log w moves
// by (accepted - 0.44) / sqrt(proposals so far) after each evaluated
// proposal (Robbins-Monro toward the 0.44 acceptance of a one-dimensional
// random walk); after burn-in w stays fixed, so the sampled part of the
// chain runs one fixed kernel.
struct PrecisionWidth {
    double w = 0.0;
    bool adapt = false;
    long long proposals = 0;
    void update(bool accepted, bool burnIn) {
        if (!adapt || !burnIn) return;
        ++proposals;
        w *= std::exp(((accepted ? 1.0 : 0.0) - 0.44) / std::sqrt((double)proposals));
    }
};

// Host helpers of the precision and trait-precision updates (mds_hmc.hip).
// Marsaglia & Tsang (2000) gamma variate of the given shape (scale 1).
double gamma_draw(mdsgen::Stream& s, double shape);
// Cholesky factor (lower) of a symmetric positive-definite p x p matrix;
// false when it is not positive definite.
bool cholesky(const std::vector<double>& a, int p, std::vector<double>& l);
std::vector<double> spd_inverse(const std::vector<double>& a, int p);
double log_det_spd(const std::vector<double>& a, int p);
// Wishart(df, scale) by the Bartlett decomposition.
std::vector<double> wishart_draw(mdsgen::Stream& s, double df, const std::vector<double>& scale, int p);
// The sampler's log-likelihood from the engine's raw sum (R/mds.R:24-28) and
// the matrix-normal log prior with row covariance I_N (R/Untitled.R:68-95),
// quad = sum_i x_i^T T x_i.
double mds_log_likelihood(double rawSum, double t, int n, bool truncated);
double matrix_normal_log_prior(double quad, int n, int p, const std::vector<double>& trait);
// The trace's location rows of an N-location run, and one trace row from
// them (MDS_HIP_HMC_TRACE_COLUMNS values).
TraceRows trace_rows(int n);
void write_trace_row(double* row, const double* rows, int p, double logPost, double tau,
                     const std::vector<double>& trait);

struct Hmc {
    Engine& e;
    mds_hip_hmc_options opt;
    int n = 0, p = 0;
    double tau = 0.0, eps = 0.0;
    std::vector<double> trait;
    double* dMom = nullptr;
    double* dGradSave = nullptr;
    void* dXSave = nullptr;
    double* dParams = nullptr;
    double* hParams = nullptr;
    hipEvent_t uploadDone = nullptr;  // after the last copy from hParams
    bool uploadPending = false;
    bool paramsStale = false;  // tau changed on the host since the last upload
    // Results the host reads, written by the kernels themselves into pinned
    // coherent memory (no copy per value): hScalars by slot (ScalarSlot),
    // hRows the trace rows [kTraceRows][p], hGram X^T X [p][p]; kScalars,
    // kRows and kGram are the same memory's device addresses.
    enum ScalarSlot { kSlotK = 0, kSlotQuad = 1, kSlotSum = 2, kSlotSumStar = 3, kSlotK0 = 4, kScalarSlots = 6 };
    double* hScalars = nullptr;
    double* kScalars = nullptr;
    double* hRows = nullptr;
    double* kRows = nullptr;
    double* hGram = nullptr;
    double* kGram = nullptr;
    void* hX = nullptr;
    TraceRows traceRows{};
    // gradient sources of a leapfrog step
    enum Mode { kExact = 0, kExactCollect = 1, kSurrogate = 2 };
    Surrogate* sur = nullptr;
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t graphExec[3] = {nullptr, nullptr, nullptr};

    Hmc(Engine& engine, const mds_hip_hmc_options& options);
    ~Hmc();
    void release();
    void uploadParams();
    template <int FORM>
    void kick();
    void trajectory(int mode);
    void gradientAtCurrent();
    void gradient(int mode);
    void buildGraph(int mode);
    void readEnergies();
    double logLikelihood(double rawSum, double t) const;
    double logPrior(double quad) const;
    void copyLocationsToHost(std::vector<double>& out);
    void gatherTraceRows();
    void enqueueEnergies();
    void launchEnergy(int slot, bool withSum);
    void startIteration(int it);
    void restoreStart();
    void enqueueAtKept(bool traitStep, bool recordIt, double tauStar);
    void gatherGram();
    void record(double* row, const double* rows, double logPost) const;
    void run(const double* x0, double* trace, double* finalX, mds_hip_hmc_stats* stats);
};

}  // namespace mdship
