// Per-location MLP surrogate of the gradient rows (see mds_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "mds_engine.hpp"

namespace mdship {

constexpr int kMlpIn = 32;      // P + 16 embedding values, zero padded
constexpr int kMlpHidden = 256;
constexpr int kMlpOut = 16;     // P outputs, zero padded
constexpr int kMlpEmbed = 16;
constexpr int kMlpMaxP = 16;
constexpr int kSplits = 32;     // K slices of the weight-gradient products

struct Mlp {
    int n = 0, p = 0, device = 0, rows = 0;
    hipStream_t stream = nullptr;
    std::vector<int> sizes, offsets;
    int paramCount = 0;
    long long adamStep = 0;
    float* dParams = nullptr;  // W1 (32x256), b1, W2 (256x256), b2, W3 (256x16), b3, embeddings (n x 16)
    float* dGrads = nullptr;
    float* dM = nullptr;
    float* dV = nullptr;
    std::vector<float> norm;   // 1/sigma_x, 1/sigma_g, sigma_g (kMlpMaxP each)
    float* dNorm = nullptr;
    float* dCenter = nullptr;  // per-location centres: x (n x p), then g (n x p)
    __bf16 *w1t = nullptr, *w1rm = nullptr, *w2t = nullptr, *w2rm = nullptr, *w3t = nullptr, *w3rm = nullptr;
    __bf16 *a0 = nullptr, *a0t = nullptr, *h1 = nullptr, *h1t = nullptr, *h2 = nullptr, *h2t = nullptr;
    __bf16 *dOut = nullptr, *dOutT = nullptr, *dH2 = nullptr, *dH2t = nullptr, *dH1 = nullptr, *dH1t = nullptr;
    float* out = nullptr;
    float* dA0 = nullptr;
    float* lossPart = nullptr;
    float* splitBuf = nullptr;

    Mlp(int n, int p, int device, uint64_t seed);
    ~Mlp();
    void release();
    void convertWeights();
    void forward(bool keepTransposed, hipStream_t s);
    void backward();
    void weightGrad(const __bf16* aT, int k, const __bf16* yT, int nOut, float* dst);
    void setNormalisation(const float* xc, const float* gc, const float* sx, const float* sg, bool centresOnDevice);
    void lossAndGradient(const float* x, const float* g);
    void trainStep(const float* x, const float* g, float lr);
    double loss();
    template <typename XT>
    void predictInto(const XT* x, int xStride, double* grad, hipStream_t s);
};

// Device store of (x, gradient) snapshots in FP32, [capacity][n][p] each;
// the next slot index is kept on the device so that collection can run
// inside a captured graph.
struct SnapshotStore {
    int n = 0, p = 0, capacity = 0, device = 0;
    float* sx = nullptr;
    float* sg = nullptr;
    double* dSlot = nullptr;  // [0] next slot, [1] the value at the current trajectory's start
    SnapshotStore(int n, int p, int capacity, int device);
    ~SnapshotStore();
    int count();
};

struct Surrogate {
    Mlp mlp;
    SnapshotStore store;
    uint64_t seed;
    std::string lastError;
    Surrogate(int n, int p, int capacity, uint64_t seed, int device)
        : mlp(n, p, device, seed), store(n, p, capacity, device), seed(seed) {}
    double train(int epochs, double lr);
};

void launch_snapshot(const Engine& e, const double* grad, float* sx, float* sg, const double* slotPtr,
                     int capacity, hipStream_t s);

}  // namespace mdship

struct mds_hip_surrogate {
    mdship::Surrogate* s;
};
