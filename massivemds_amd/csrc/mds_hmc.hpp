// Device-resident HMC sampler over an Engine (see mds_hmc.hip).
#pragma once

#include <vector>

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
    double* dEnergy = nullptr;
    double* hParams = nullptr;
    hipEvent_t uploadDone = nullptr;  // after the last copy from hParams
    bool uploadPending = false;
    double* hScalars = nullptr;
    void* hX = nullptr;
    double* dRows = nullptr;  // trace rows [kTraceRows][p]
    double* dGram = nullptr;  // X^T X [p][p]
    double* hRows = nullptr;  // pinned
    double* hGram = nullptr;  // pinned
    TraceRows traceRows{};
    // gradient sources of a leapfrog step
    enum Mode { kExact = 0, kExactCollect = 1, kSurrogate = 2 };
    Surrogate* sur = nullptr;
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t graphExec[3] = {nullptr, nullptr, nullptr};

    Hmc(Engine& engine, const mds_hip_hmc_options& options);
    ~Hmc();
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
    void gatherGram();
    void record(double* row, const double* rows, double logPost) const;
    void run(const double* x0, double* trace, double* finalX, mds_hip_hmc_stats* stats);
};

}  // namespace mdship
