// Per-location MLP surrogate of the log-likelihood gradient rows (no
// counterpart in the reference): input (x_i centred and scaled, e_i) of
// P + 16 values padded to 32, two hidden layers of 256 with tanh, output the
// P gradient components (centred and scaled, padded to 16).  The centres are
// per location (the mean of x_i and of g_i over the training snapshots), the
// scales per dimension: the posterior moves each x_i by far less than the
// spread of the configuration, so a global centre would leave the network's
// BF16 inputs unable to resolve those moves.  Weights are kept in FP32 (Adam
// masters) and in BF16 copies for the MFMA matrix cores; products accumulate
// in FP32.
//
// All matrix products run on one hand-written kernel, C = A * B with A
// (M x K) row-major and B supplied as B^T (N x K) row-major, so that every
// lane of a wavefront reads 8 consecutive BF16 values of each operand for
// v_mfma_f32_16x16x32_bf16 (lane l: A[m = l % 16][k = 8 (l / 16) + 0..7],
// B[k = 8 (l / 16) + 0..7][n = l % 16], C[row = 4 (l / 16) + j][col = l % 16]).
// Fused epilogues: bias + tanh (forward), tanh derivative (backward), FP32
// store, and split-K partials for the weight gradients (reduced in a fixed
// order: training is deterministic).
//
// Training step = one snapshot of all N locations, so each location's
// embedding row receives exactly one gradient row (no scatter, no atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mds_datagen.hpp"
#include "mds_engine.hpp"
#include "mds_mlp.hpp"

namespace mdship {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

enum GemmEpilogue : int {
    kEpiStoreF32 = 0,    // C (+ bias) -> FP32 [M x ldc]
    kEpiTanhBf16 = 1,    // tanh(C + bias) -> BF16 row-major and transposed
    kEpiDtanhBf16 = 2,   // C * (1 - h^2), h from a BF16 row-major activation -> BF16 both layouts
    kEpiSplitK = 3,      // FP32 partial of this K slice -> out + z * M * ldc
};

struct GemmArgs {
    const __bf16* A;   // M x K, row stride lda
    const __bf16* Bt;  // N x K, row stride ldb
    int M, N, K, lda, ldb;
    int kSlice;        // K per blockIdx.z (split-K), else K
    const float* bias; // N, may be null
    float* outF;       // FP32 output, row stride ldc
    int ldc;
    __bf16* outB;      // BF16 row-major output, row stride ldc
    __bf16* outT;      // BF16 transposed output, row stride ldt (may be null)
    int ldt;
    const __bf16* act; // activation h for the tanh derivative, row stride ldc
};

// Workgroup: 4 wavefronts, each 16 rows x TN columns (TN / 16 accumulators).
template <int TN, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int m0 = (blockIdx.x * 4 + wave) * 16;
    const int n0 = blockIdx.y * TN;
    if (m0 >= g.M) return;
    const int kBegin = blockIdx.z * g.kSlice;
    const int kEnd = min(g.K, kBegin + g.kSlice);
    f32x4 acc[TN / 16];
#pragma unroll
    for (int t = 0; t < TN / 16; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const __bf16* aRow = g.A + (int64_t)(m0 + lane % 16) * g.lda + (lane / 16) * 8;
    for (int k0 = kBegin; k0 < kEnd; k0 += 32) {
        const bf16x8 a = *(const bf16x8*)(aRow + k0);
#pragma unroll
        for (int t = 0; t < TN / 16; ++t) {
            const bf16x8 b =
                *(const bf16x8*)(g.Bt + (int64_t)(n0 + t * 16 + lane % 16) * g.ldb + (lane / 16) * 8 + k0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < TN / 16; ++t) {
        const int col = n0 + t * 16 + lane % 16;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = m0 + (lane / 16) * 4 + j;
            float v = acc[t][j];
            if (EPI == kEpiStoreF32) {
                if (g.bias) v += g.bias[col];
                g.outF[(int64_t)row * g.ldc + col] = v;
            } else if (EPI == kEpiTanhBf16) {
                v = tanhf(v + g.bias[col]);
                const __bf16 h = (__bf16)v;
                g.outB[(int64_t)row * g.ldc + col] = h;
                if (g.outT) g.outT[(int64_t)col * g.ldt + row] = h;
            } else if (EPI == kEpiDtanhBf16) {
                const float h = (float)g.act[(int64_t)row * g.ldc + col];
                const __bf16 d = (__bf16)(v * (1.0f - h * h));
                g.outB[(int64_t)row * g.ldc + col] = d;
                if (g.outT) g.outT[(int64_t)col * g.ldt + row] = d;
            } else {
                g.outF[(int64_t)blockIdx.z * g.M * g.ldc + (int64_t)row * g.ldc + col] = v;
            }
        }
    }
}

// out[i] = sum over z of parts[z][i] (fixed order).
__global__ void sum_slices_kernel(const float* __restrict__ parts, int slices, int count, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.f;
    for (int z = 0; z < slices; ++z) s += parts[(int64_t)z * count + i];
    out[i] = s;
}

// out[c] = sum over rows of the transposed BF16 matrix row c (length m).
__global__ __launch_bounds__(256) void row_sum_bf16_kernel(const __bf16* __restrict__ t, int ld, int m,
                                                           float* __restrict__ out) {
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int i = threadIdx.x; i < m; i += 256) s += (float)t[(int64_t)c * ld + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

// FP32 masters -> BF16 copies: W (rows x cols, row-major) to Wrm (rows x
// ldr, zero padded) and Wt (cols x ldt, zero padded).
__global__ void convert_weights_kernel(const float* __restrict__ w, int rows, int cols, __bf16* __restrict__ wrm,
                                       int ldr, __bf16* __restrict__ wt, int ldt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    const int r = i / cols, c = i % cols;
    const __bf16 v = (__bf16)w[i];
    wrm[(int64_t)r * ldr + c] = v;
    wt[(int64_t)c * ldt + r] = v;
}

// Network input rows: [ (x_i - xc_i) / sigma_x (P), e_i (16), 0 ... ] as
// BF16, row-major (rows x 32) and transposed (32 x rows).  x has row stride
// xStride (FP32 snapshot or engine locations in FP64/FP32); xc is n x p.
template <typename XT>
__global__ void assemble_input_kernel(const XT* __restrict__ x, int xStride, int n, int p,
                                      const float* __restrict__ norm, const float* __restrict__ xc,
                                      const float* __restrict__ emb, __bf16* __restrict__ a0,
                                      __bf16* __restrict__ a0t, int rows) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    for (int c = 0; c < kMlpIn; ++c) {
        float v = 0.f;
        if (i < n) {
            if (c < p)
                v = ((float)x[(int64_t)i * xStride + c] - xc[(int64_t)i * p + c]) * norm[c];
            else if (c < p + kMlpEmbed)
                v = emb[(int64_t)i * kMlpEmbed + (c - p)];
        }
        const __bf16 b = (__bf16)v;
        a0[(int64_t)i * kMlpIn + c] = b;
        if (a0t) a0t[(int64_t)c * rows + i] = b;
    }
}

// Loss gradient of the mean squared error over the p outputs:
// dOut = 2 (out - t) / (n p) with t = (g - gc) / sigma_g; BF16 row-major
// (rows x 32, zero padded) and transposed (16 x rows).  lossPart[i] = row i's
// squared error.
__global__ void loss_grad_kernel(const float* __restrict__ out, const float* __restrict__ g, int n, int p,
                                 const float* __restrict__ norm, const float* __restrict__ gc,
                                 __bf16* __restrict__ dOut,
                                 __bf16* __restrict__ dOutT, int rows, float* __restrict__ lossPart) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    float l = 0.f;
    const float scale = 2.0f / ((float)n * (float)p);
    for (int c = 0; c < 32; ++c) {
        float d = 0.f;
        if (i < n && c < p) {
            const float t = (g[(int64_t)i * p + c] - gc[(int64_t)i * p + c]) * norm[kMlpMaxP + c];
            const float e = out[(int64_t)i * kMlpOut + c] - t;
            l += e * e;
            d = scale * e;
        }
        dOut[(int64_t)i * 32 + c] = (__bf16)d;
        if (c < kMlpOut) dOutT[(int64_t)c * rows + i] = (__bf16)d;
    }
    lossPart[i] = l;
}

// Adam (Kingma & Ba 2015) on FP32 parameters.
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ grad, float* __restrict__ m,
                            float* __restrict__ v, int count, float lr, float b1, float b2, float c1, float c2,
                            float epsilon) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + epsilon);
}

// Embedding gradient rows: dA0[i][p .. p + 16) -> dEmb[i][0 .. 16).
__global__ void embed_grad_kernel(const float* __restrict__ dA0, int n, int p, float* __restrict__ dEmb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * kMlpEmbed) return;
    const int r = i / kMlpEmbed, c = i % kMlpEmbed;
    dEmb[i] = dA0[(int64_t)r * kMlpIn + p + c];
}

// Prediction: gradient rows g_i = gc_i + out_i * sigma_g (into the engine's
// FP64 gradient buffer).
__global__ void denormalize_kernel(const float* __restrict__ out, int n, int p, const float* __restrict__ norm,
                                   const float* __restrict__ gc, double* __restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * p) return;
    const int r = i / p, c = i % p;
    grad[i] = (double)gc[i] + (double)out[(int64_t)r * kMlpOut + c] * (double)norm[2 * kMlpMaxP + c];
}

// Per-element mean and sum of squared deviations over `count` snapshots of
// n p values each (two passes over the store, for the centres and scales).
__global__ void snapshot_moments_kernel(const float* __restrict__ store, int count, int64_t np,
                                        float* __restrict__ mean, float* __restrict__ ssd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double s = 0.0;
    for (int k = 0; k < count; ++k) s += store[(int64_t)k * np + i];
    const double m = s / count;
    double q = 0.0;
    for (int k = 0; k < count; ++k) {
        const double d = store[(int64_t)k * np + i] - m;
        q += d * d;
    }
    mean[i] = (float)m;
    ssd[i] = (float)q;
}

// Snapshot collection: x (engine layout) and the FP64 gradient into FP32 slot
// `slot` of the store; the slot index is read from the device (advanced by
// snapshot_advance_kernel), so the step can be replayed from a graph.
template <typename XT>
__global__ void snapshot_kernel(const XT* __restrict__ x, int xStride, const double* __restrict__ grad, int n,
                                int p, float* __restrict__ sx, float* __restrict__ sg,
                                const double* __restrict__ slotPtr, int capacity) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = (int)slotPtr[0];
    if (i >= n * p || slot < 0 || slot >= capacity) return;
    const int r = i / p, c = i % p;
    sx[(int64_t)slot * n * p + i] = (float)x[(int64_t)r * xStride + c];
    sg[(int64_t)slot * n * p + i] = (float)grad[i];
}

__global__ void snapshot_advance_kernel(double* __restrict__ slotPtr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) slotPtr[0] += 1.0;
}

// ------------------------------------------------------------------ host

static void launch_gemm(const GemmArgs& g, int tn, int epi, int splits, hipStream_t s) {
    if (g.M % 16 || g.N % tn || g.K % 32 || g.kSlice % 32)
        throw Error(MDS_HIP_ERR_ARGUMENT, "gemm shape not padded");
    const dim3 grid((g.M + 63) / 64, g.N / tn, splits), block(256);
#define MDS_GEMM_CASE(TN, E)                                                                   \
    if (tn == TN && epi == E) {                                                                \
        hipLaunchKernelGGL((gemm_bf16_kernel<TN, E>), grid, block, 0, s, g);                   \
        check(hipGetLastError(), "gemm kernel");                                               \
        return;                                                                                \
    }
    MDS_GEMM_CASE(64, kEpiTanhBf16)
    MDS_GEMM_CASE(64, kEpiDtanhBf16)
    MDS_GEMM_CASE(64, kEpiStoreF32)
    MDS_GEMM_CASE(64, kEpiSplitK)
    MDS_GEMM_CASE(16, kEpiStoreF32)
    MDS_GEMM_CASE(16, kEpiSplitK)
#undef MDS_GEMM_CASE
    throw Error(MDS_HIP_ERR_ARGUMENT, "unsupported gemm configuration");
}

static inline int blocks_for(int64_t count, int threads = 256) { return (int)((count + threads - 1) / threads); }

Mlp::Mlp(int n_, int p_, int device_, uint64_t seed) : n(n_), p(p_), device(device_) {
    if (p < 1 || p > kMlpMaxP) throw Error(MDS_HIP_ERR_ARGUMENT, "surrogate dimension must be in [1, 16]");
    if (n < 1) throw Error(MDS_HIP_ERR_ARGUMENT, "location count must be >= 1");
    rows = (n + 63) / 64 * 64;
    check(hipSetDevice(device), "hipSetDevice");
    check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    // parameters: W1 32x256, b1 256, W2 256x256, b2 256, W3 256x16, b3 16, emb n x 16
    sizes = {kMlpIn * kMlpHidden, kMlpHidden, kMlpHidden * kMlpHidden, kMlpHidden, kMlpHidden * kMlpOut, kMlpOut,
             n * kMlpEmbed};
    offsets.resize(sizes.size());
    int total = 0;
    for (size_t k = 0; k < sizes.size(); ++k) {
        offsets[k] = total;
        total += sizes[k];
    }
    paramCount = total;
    // Glorot-uniform weights on the real fan-in / fan-out, zero biases, small
    // random embeddings; padded rows / columns stay zero
    std::vector<float> host(total, 0.f);
    mdsgen::Stream rng(seed, 31);
    auto uni = [&](float a) { return (float)((2.0 * rng.uniform() - 1.0) * a); };
    const int fanIn1 = p + kMlpEmbed;
    const float a1 = std::sqrt(6.f / (fanIn1 + kMlpHidden));
    for (int r = 0; r < fanIn1; ++r)
        for (int c = 0; c < kMlpHidden; ++c) host[offsets[0] + r * kMlpHidden + c] = uni(a1);
    const float a2 = std::sqrt(6.f / (2 * kMlpHidden));
    for (int i = 0; i < kMlpHidden * kMlpHidden; ++i) host[offsets[2] + i] = uni(a2);
    const float a3 = std::sqrt(6.f / (kMlpHidden + p));
    for (int r = 0; r < kMlpHidden; ++r)
        for (int c = 0; c < p; ++c) host[offsets[4] + r * kMlpOut + c] = uni(a3);
    for (int i = 0; i < n * kMlpEmbed; ++i) host[offsets[6] + i] = uni(0.1f);
    norm.assign(3 * kMlpMaxP, 1.f);
    try {
        check(hipMalloc((void**)&dParams, (size_t)total * sizeof(float)), "hipMalloc(mlp)");
        check(hipMalloc((void**)&dGrads, (size_t)total * sizeof(float)), "hipMalloc(mlp grads)");
        check(hipMalloc((void**)&dM, (size_t)total * sizeof(float)), "hipMalloc(adam m)");
        check(hipMalloc((void**)&dV, (size_t)total * sizeof(float)), "hipMalloc(adam v)");
        check(hipMemcpy(dParams, host.data(), (size_t)total * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemset(dM, 0, (size_t)total * sizeof(float)), "hipMemset");
        check(hipMemset(dV, 0, (size_t)total * sizeof(float)), "hipMemset");
        check(hipMemset(dGrads, 0, (size_t)total * sizeof(float)), "hipMemset");
        check(hipMalloc((void**)&dNorm, norm.size() * sizeof(float)), "hipMalloc(norm)");
        check(hipMemcpy(dNorm, norm.data(), norm.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMalloc((void**)&dCenter, 2 * (size_t)n * p * sizeof(float)), "hipMalloc(centres)");
        check(hipMemset(dCenter, 0, 2 * (size_t)n * p * sizeof(float)), "hipMemset");
        auto bf = [&](__bf16** ptr, size_t count) {
            check(hipMalloc((void**)ptr, count * sizeof(__bf16)), "hipMalloc(bf16)");
            check(hipMemset(*ptr, 0, count * sizeof(__bf16)), "hipMemset");
        };
        bf(&w1t, (size_t)kMlpHidden * kMlpIn);
        bf(&w1rm, (size_t)kMlpIn * kMlpHidden);
        bf(&w2t, (size_t)kMlpHidden * kMlpHidden);
        bf(&w2rm, (size_t)kMlpHidden * kMlpHidden);
        bf(&w3t, (size_t)kMlpOut * kMlpHidden);
        bf(&w3rm, (size_t)kMlpHidden * 32);
        bf(&a0, (size_t)rows * kMlpIn);
        bf(&a0t, (size_t)kMlpIn * rows);
        bf(&h1, (size_t)rows * kMlpHidden);
        bf(&h1t, (size_t)kMlpHidden * rows);
        bf(&h2, (size_t)rows * kMlpHidden);
        bf(&h2t, (size_t)kMlpHidden * rows);
        bf(&dOut, (size_t)rows * 32);
        bf(&dOutT, (size_t)kMlpOut * rows);
        bf(&dH2, (size_t)rows * kMlpHidden);
        bf(&dH2t, (size_t)kMlpHidden * rows);
        bf(&dH1, (size_t)rows * kMlpHidden);
        bf(&dH1t, (size_t)kMlpHidden * rows);
        check(hipMalloc((void**)&out, (size_t)rows * kMlpOut * sizeof(float)), "hipMalloc(out)");
        check(hipMalloc((void**)&dA0, (size_t)rows * kMlpIn * sizeof(float)), "hipMalloc(dA0)");
        check(hipMalloc((void**)&lossPart, (size_t)rows * sizeof(float)), "hipMalloc(loss)");
        check(hipMalloc((void**)&splitBuf, (size_t)kSplits * kMlpHidden * kMlpHidden * sizeof(float)),
              "hipMalloc(split-K)");
        convertWeights();
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    } catch (...) {
        release();
        throw;
    }
}

Mlp::~Mlp() { release(); }

void Mlp::release() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* ptr : {(void*)dParams, (void*)dGrads, (void*)dM, (void*)dV, (void*)dNorm, (void*)dCenter,
                      (void*)w1t, (void*)w1rm,
                      (void*)w2t, (void*)w2rm, (void*)w3t, (void*)w3rm, (void*)a0, (void*)a0t, (void*)h1,
                      (void*)h1t, (void*)h2, (void*)h2t, (void*)dOut, (void*)dOutT, (void*)dH2, (void*)dH2t,
                      (void*)dH1, (void*)dH1t, (void*)out, (void*)dA0, (void*)lossPart, (void*)splitBuf})
        (void)hipFree(ptr);
    dParams = dGrads = dM = dV = dNorm = dCenter = out = dA0 = lossPart = splitBuf = nullptr;
    w1t = w1rm = w2t = w2rm = w3t = w3rm = a0 = a0t = h1 = h1t = h2 = h2t = nullptr;
    dOut = dOutT = dH2 = dH2t = dH1 = dH1t = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
}

void Mlp::convertWeights() {
    hipLaunchKernelGGL(convert_weights_kernel, dim3(blocks_for(kMlpIn * kMlpHidden)), dim3(256), 0, stream,
                       dParams + offsets[0], kMlpIn, kMlpHidden, w1rm, kMlpHidden, w1t, kMlpIn);
    hipLaunchKernelGGL(convert_weights_kernel, dim3(blocks_for(kMlpHidden * kMlpHidden)), dim3(256), 0, stream,
                       dParams + offsets[2], kMlpHidden, kMlpHidden, w2rm, kMlpHidden, w2t, kMlpHidden);
    hipLaunchKernelGGL(convert_weights_kernel, dim3(blocks_for(kMlpHidden * kMlpOut)), dim3(256), 0, stream,
                       dParams + offsets[4], kMlpHidden, kMlpOut, w3rm, 32, w3t, kMlpHidden);
    check(hipGetLastError(), "convert weights");
}

void Mlp::forward(bool keepTransposed, hipStream_t s) {
    GemmArgs g{};
    // layer 1: h1 = tanh(a0 W1 + b1)
    g.A = a0; g.lda = kMlpIn; g.Bt = w1t; g.ldb = kMlpIn;
    g.M = rows; g.N = kMlpHidden; g.K = kMlpIn; g.kSlice = kMlpIn;
    g.bias = dParams + offsets[1]; g.outB = h1; g.ldc = kMlpHidden;
    g.outT = keepTransposed ? h1t : nullptr; g.ldt = rows;
    launch_gemm(g, 64, kEpiTanhBf16, 1, s);
    // layer 2: h2 = tanh(h1 W2 + b2)
    g.A = h1; g.lda = kMlpHidden; g.Bt = w2t; g.ldb = kMlpHidden; g.K = kMlpHidden; g.kSlice = kMlpHidden;
    g.bias = dParams + offsets[3]; g.outB = h2; g.outT = keepTransposed ? h2t : nullptr;
    launch_gemm(g, 64, kEpiTanhBf16, 1, s);
    // output: out = h2 W3 + b3
    GemmArgs o{};
    o.A = h2; o.lda = kMlpHidden; o.Bt = w3t; o.ldb = kMlpHidden;
    o.M = rows; o.N = kMlpOut; o.K = kMlpHidden; o.kSlice = kMlpHidden;
    o.bias = dParams + offsets[5]; o.outF = out; o.ldc = kMlpOut;
    launch_gemm(o, 16, kEpiStoreF32, 1, s);
}

// Weight gradient dW (k x n) = A^T dY from transposed activations aT (k x rows)
// and transposed output gradients yT (n x rows), split over kSplits slices of
// the rows and summed in order.
void Mlp::weightGrad(const __bf16* aT, int k, const __bf16* yT, int nOut, float* dst) {
    GemmArgs g{};
    g.A = aT; g.lda = rows; g.Bt = yT; g.ldb = rows;
    g.M = k; g.N = nOut; g.K = rows;
    const int splits = kSplits;
    g.kSlice = ((rows + splits - 1) / splits + 31) / 32 * 32;  // trailing slices may be empty
    g.outF = splitBuf; g.ldc = nOut;
    launch_gemm(g, nOut % 64 == 0 ? 64 : 16, kEpiSplitK, splits, stream);
    hipLaunchKernelGGL(sum_slices_kernel, dim3(blocks_for(k * nOut)), dim3(256), 0, stream, splitBuf, splits,
                       k * nOut, dst);
    check(hipGetLastError(), "sum slices");
}

void Mlp::backward() {
    // dH2 = (dOut W3^T) * (1 - h2^2)
    GemmArgs g{};
    g.A = dOut; g.lda = 32; g.Bt = w3rm; g.ldb = 32; g.M = rows; g.N = kMlpHidden; g.K = 32; g.kSlice = 32;
    g.outB = dH2; g.outT = dH2t; g.ldc = kMlpHidden; g.ldt = rows; g.act = h2;
    launch_gemm(g, 64, kEpiDtanhBf16, 1, stream);
    // dH1 = (dH2 W2^T) * (1 - h1^2)
    g.A = dH2; g.lda = kMlpHidden; g.Bt = w2rm; g.ldb = kMlpHidden; g.K = kMlpHidden; g.kSlice = kMlpHidden;
    g.outB = dH1; g.outT = dH1t; g.act = h1;
    launch_gemm(g, 64, kEpiDtanhBf16, 1, stream);
    // dA0 = dH1 W1^T (FP32; only the embedding columns are used)
    GemmArgs d{};
    d.A = dH1; d.lda = kMlpHidden; d.Bt = w1rm; d.ldb = kMlpHidden; d.M = rows; d.N = kMlpIn; d.K = kMlpHidden;
    d.kSlice = kMlpHidden; d.outF = dA0; d.ldc = kMlpIn;
    launch_gemm(d, 16, kEpiStoreF32, 1, stream);
    // weight and bias gradients
    weightGrad(a0t, kMlpIn, dH1t, kMlpHidden, dGrads + offsets[0]);
    weightGrad(h1t, kMlpHidden, dH2t, kMlpHidden, dGrads + offsets[2]);
    weightGrad(h2t, kMlpHidden, dOutT, kMlpOut, dGrads + offsets[4]);
    hipLaunchKernelGGL(row_sum_bf16_kernel, dim3(kMlpHidden), dim3(256), 0, stream, dH1t, rows, rows,
                       dGrads + offsets[1]);
    hipLaunchKernelGGL(row_sum_bf16_kernel, dim3(kMlpHidden), dim3(256), 0, stream, dH2t, rows, rows,
                       dGrads + offsets[3]);
    hipLaunchKernelGGL(row_sum_bf16_kernel, dim3(kMlpOut), dim3(256), 0, stream, dOutT, rows, rows,
                       dGrads + offsets[5]);
    hipLaunchKernelGGL(embed_grad_kernel, dim3(blocks_for(n * kMlpEmbed)), dim3(256), 0, stream, dA0, n, p,
                       dGrads + offsets[6]);
    check(hipGetLastError(), "backward kernels");
}

// Per-dimension scales sx, sg (p each); per-location centres xc, gc (n x p
// each, host or device pointers as `centresOnDevice` says).
void Mlp::setNormalisation(const float* xc, const float* gc, const float* sx, const float* sg,
                           bool centresOnDevice) {
    for (int c = 0; c < kMlpMaxP; ++c) {
        if (c < p && !(sx[c] > 0.f && sg[c] > 0.f)) throw Error(MDS_HIP_ERR_ARGUMENT, "scales must be > 0");
        norm[c] = c < p ? 1.f / sx[c] : 1.f;
        norm[kMlpMaxP + c] = c < p ? 1.f / sg[c] : 1.f;
        norm[2 * kMlpMaxP + c] = c < p ? sg[c] : 1.f;
    }
    check(hipMemcpyAsync(dNorm, norm.data(), norm.size() * sizeof(float), hipMemcpyHostToDevice, stream),
          "hipMemcpy(norm)");
    const size_t bytes = (size_t)n * p * sizeof(float);
    const hipMemcpyKind kind = centresOnDevice ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    check(hipMemcpyAsync(dCenter, xc, bytes, kind, stream), "hipMemcpy(centres)");
    check(hipMemcpyAsync(dCenter + (size_t)n * p, gc, bytes, kind, stream), "hipMemcpy(centres)");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

// One forward + backward on a snapshot (x, g both n x p FP32): fills the
// parameter gradients dGrads and the per-row losses.
void Mlp::lossAndGradient(const float* x, const float* g) {
    hipLaunchKernelGGL((assemble_input_kernel<float>), dim3(blocks_for(rows)), dim3(256), 0, stream, x, p, n, p,
                       dNorm, dCenter, dParams + offsets[6], a0, a0t, rows);
    forward(true, stream);
    hipLaunchKernelGGL(loss_grad_kernel, dim3(blocks_for(rows)), dim3(256), 0, stream, out, g, n, p, dNorm,
                       dCenter + (size_t)n * p, dOut, dOutT, rows, lossPart);
    check(hipGetLastError(), "loss kernel");
    backward();
}

void Mlp::trainStep(const float* x, const float* g, float lr) {
    lossAndGradient(x, g);
    ++adamStep;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f / (1.f - std::pow(b1, (float)adamStep));
    const float c2 = 1.f / (1.f - std::pow(b2, (float)adamStep));
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(paramCount)), dim3(256), 0, stream, dParams, dGrads, dM, dV,
                       paramCount, lr, b1, b2, c1, c2, 1e-8f);
    check(hipGetLastError(), "adam kernel");
    convertWeights();
}

double Mlp::loss() {
    std::vector<float> h(rows);
    check(hipMemcpyAsync(h.data(), lossPart, rows * sizeof(float), hipMemcpyDeviceToHost, stream), "hipMemcpy");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += h[i];
    return s / ((double)n * p);
}

template <typename XT>
void Mlp::predictInto(const XT* x, int xStride, double* grad, hipStream_t s) {
    hipLaunchKernelGGL((assemble_input_kernel<XT>), dim3(blocks_for(rows)), dim3(256), 0, s, x, xStride, n, p, dNorm,
                       dCenter, dParams + offsets[6], a0, nullptr, rows);
    forward(false, s);
    hipLaunchKernelGGL(denormalize_kernel, dim3(blocks_for(n * p)), dim3(256), 0, s, out, n, p, dNorm,
                       dCenter + (size_t)n * p, grad);
    check(hipGetLastError(), "predict kernels");
}

template void Mlp::predictInto<double>(const double*, int, double*, hipStream_t);
template void Mlp::predictInto<float>(const float*, int, double*, hipStream_t);

void launch_snapshot(const Engine& e, const double* grad, float* sx, float* sg, const double* slotPtr,
                     int capacity, hipStream_t s) {
    const int count = e.N * e.P;
    if (e.fp32)
        hipLaunchKernelGGL((snapshot_kernel<float>), dim3(blocks_for(count)), dim3(256), 0, s,
                           (const float*)e.dX[e.cur], e.kdim, grad, e.N, e.P, sx, sg, slotPtr, capacity);
    else
        hipLaunchKernelGGL((snapshot_kernel<double>), dim3(blocks_for(count)), dim3(256), 0, s,
                           (const double*)e.dX[e.cur], e.kdim, grad, e.N, e.P, sx, sg, slotPtr, capacity);
    hipLaunchKernelGGL(snapshot_advance_kernel, dim3(1), dim3(64), 0, s, (double*)slotPtr);
    check(hipGetLastError(), "snapshot kernels");
}

SnapshotStore::SnapshotStore(int n_, int p_, int capacity_, int device_)
    : n(n_), p(p_), capacity(capacity_), device(device_) {
    if (capacity < 1) throw Error(MDS_HIP_ERR_ARGUMENT, "snapshot capacity must be >= 1");
    const size_t bytes = (size_t)capacity * n * p * sizeof(float);
    try {
        check(hipMalloc((void**)&sx, bytes), "hipMalloc(snapshots x)");
        check(hipMalloc((void**)&sg, bytes), "hipMalloc(snapshots g)");
        check(hipMalloc((void**)&dSlot, 2 * sizeof(double)), "hipMalloc(slot)");
        check(hipMemset(dSlot, 0, 2 * sizeof(double)), "hipMemset(slot)");
    } catch (...) {
        (void)hipFree(sx);
        (void)hipFree(sg);
        (void)hipFree(dSlot);
        throw;
    }
}

SnapshotStore::~SnapshotStore() {
    (void)hipSetDevice(device);
    (void)hipFree(sx);
    (void)hipFree(sg);
    (void)hipFree(dSlot);
}

int SnapshotStore::count() {
    double v = 0.0;
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    check(hipMemcpy(&v, dSlot, sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy(slot)");
    return std::min((int)v, capacity);
}

// Trains on every stored snapshot `epochs` times in a shuffled order, Adam
// with a cosine learning-rate decay to lr / 20.  Inputs and targets are
// centred per location on their means over all snapshots and scaled per
// dimension by the root mean square deviation.  Returns the mean squared
// error (normalised targets) over up to 32 snapshots after training.
double Surrogate::train(int epochs, double lr) {
    const int count = store.count();
    if (count < 1) throw Error(MDS_HIP_ERR_STATE, "no snapshots to train on");
    if (epochs < 1 || !(lr > 0)) throw Error(MDS_HIP_ERR_ARGUMENT, "epochs >= 1 and learning rate > 0 required");
    const int n = store.n, p = store.p;
    const size_t np = (size_t)n * p;
    float *centres = nullptr, *ssd = nullptr;
    std::vector<float> hssd(2 * np);
    try {
        check(hipMalloc((void**)&centres, 2 * np * sizeof(float)), "hipMalloc(centres)");
        check(hipMalloc((void**)&ssd, 2 * np * sizeof(float)), "hipMalloc(moments)");
        hipLaunchKernelGGL(snapshot_moments_kernel, dim3(blocks_for((int64_t)np)), dim3(256), 0, mlp.stream, store.sx,
                           count, (int64_t)np, centres, ssd);
        hipLaunchKernelGGL(snapshot_moments_kernel, dim3(blocks_for((int64_t)np)), dim3(256), 0, mlp.stream, store.sg,
                           count, (int64_t)np, centres + np, ssd + np);
        check(hipGetLastError(), "snapshot moments");
        check(hipMemcpyAsync(hssd.data(), ssd, 2 * np * sizeof(float), hipMemcpyDeviceToHost, mlp.stream),
              "hipMemcpy");
        check(hipStreamSynchronize(mlp.stream), "hipStreamSynchronize");
        std::vector<double> qx(p, 0.0), qg(p, 0.0);
        for (size_t k = 0; k < np; ++k) {
            qx[k % p] += hssd[k];
            qg[k % p] += hssd[np + k];
        }
        std::vector<float> sx(p), sg(p);
        for (int c = 0; c < p; ++c) {
            sx[c] = (float)std::sqrt(std::max(qx[c] / ((double)n * count), 1e-30));
            sg[c] = (float)std::sqrt(std::max(qg[c] / ((double)n * count), 1e-30));
        }
        mlp.setNormalisation(centres, centres + np, sx.data(), sg.data(), true);
    } catch (...) {
        (void)hipFree(centres);
        (void)hipFree(ssd);
        throw;
    }
    (void)hipFree(centres);
    (void)hipFree(ssd);
    mdsgen::Stream rng(seed, 91);
    std::vector<int> order(count);
    const long long total = (long long)epochs * count;
    long long step = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int k = 0; k < count; ++k) order[k] = k;
        for (int k = count - 1; k > 0; --k) std::swap(order[k], order[(int)(rng.uniform() * (k + 1))]);
        for (int k = 0; k < count; ++k, ++step) {
            const double frac = (double)step / std::max(1LL, total - 1);
            const float rate = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * frac))));
            mlp.trainStep(store.sx + (size_t)order[k] * np, store.sg + (size_t)order[k] * np, rate);
        }
    }
    double l = 0.0;
    const int evals = std::min(count, 32);
    for (int q = 0; q < evals; ++q) {
        const int s = (int)((int64_t)q * count / evals);
        mlp.lossAndGradient(store.sx + (size_t)s * np, store.sg + (size_t)s * np);
        l += mlp.loss();
    }
    return l / evals;
}

}  // namespace mdship

static thread_local std::string g_surrogate_error = "no error";

template <typename F>
static int sguard(mds_hip_surrogate_handle h, F&& f) {
    if (!h || !h->s) {
        g_surrogate_error = "null handle";
        return MDS_HIP_ERR_ARGUMENT;
    }
    try {
        (void)hipSetDevice(h->s->mlp.device);
        f(*h->s);
        return MDS_HIP_OK;
    } catch (const mdship::Error& err) {
        h->s->lastError = err.what();
        return err.code;
    } catch (const std::exception& err) {
        h->s->lastError = err.what();
        return MDS_HIP_ERR_DEVICE;
    }
}

extern "C" {

int mds_hip_surrogate_create(int locationCount, int embeddingDimension, int snapshotCapacity,
                             unsigned long long seed, int device, mds_hip_surrogate_handle* out) {
    if (!out) {
        g_surrogate_error = "null output handle";
        return MDS_HIP_ERR_ARGUMENT;
    }
    *out = nullptr;
    try {
        int devices = 0;
        mdship::check(hipGetDeviceCount(&devices), "hipGetDeviceCount");
        if (devices <= 0) throw mdship::Error(MDS_HIP_ERR_DEVICE, "no HIP device");
        const int dev = (device < 0 || device >= devices) ? 0 : device;
        *out = new mds_hip_surrogate{
            new mdship::Surrogate(locationCount, embeddingDimension, snapshotCapacity, seed, dev)};
        return MDS_HIP_OK;
    } catch (const mdship::Error& err) {
        g_surrogate_error = err.what();
        return err.code;
    } catch (const std::exception& err) {
        g_surrogate_error = err.what();
        return MDS_HIP_ERR_DEVICE;
    }
}

int mds_hip_surrogate_destroy(mds_hip_surrogate_handle h) {
    if (!h) return MDS_HIP_ERR_ARGUMENT;
    delete h->s;
    delete h;
    return MDS_HIP_OK;
}

const char* mds_hip_surrogate_last_error(mds_hip_surrogate_handle h) {
    if (!h || !h->s) return g_surrogate_error.c_str();
    return h->s->lastError.c_str();
}

int mds_hip_surrogate_snapshot_count(mds_hip_surrogate_handle h, int* count) {
    return sguard(h, [&](mdship::Surrogate& s) {
        if (!count) throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "null output pointer");
        *count = s.store.count();
    });
}

int mds_hip_surrogate_add_snapshot(mds_hip_surrogate_handle h, const double* x, const double* g, size_t length) {
    return sguard(h, [&](mdship::Surrogate& s) {
        const size_t np = (size_t)s.store.n * s.store.p;
        if (!x || !g || length != np) throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "snapshot length must be N*P");
        const int slot = s.store.count();
        if (slot >= s.store.capacity) throw mdship::Error(MDS_HIP_ERR_STATE, "snapshot store full");
        std::vector<float> fx(np), fg(np);
        for (size_t k = 0; k < np; ++k) {
            fx[k] = (float)x[k];
            fg[k] = (float)g[k];
        }
        mdship::check(hipMemcpy(s.store.sx + (size_t)slot * np, fx.data(), np * sizeof(float), hipMemcpyHostToDevice),
                      "hipMemcpy");
        mdship::check(hipMemcpy(s.store.sg + (size_t)slot * np, fg.data(), np * sizeof(float), hipMemcpyHostToDevice),
                      "hipMemcpy");
        const double next = slot + 1;
        mdship::check(hipMemcpy(s.store.dSlot, &next, sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    });
}

int mds_hip_surrogate_train(mds_hip_surrogate_handle h, int epochs, double learningRate, double* finalLoss) {
    return sguard(h, [&](mdship::Surrogate& s) {
        const double l = s.train(epochs, learningRate);
        if (finalLoss) *finalLoss = l;
    });
}

int mds_hip_surrogate_predict(mds_hip_surrogate_handle h, const double* x, size_t length, double* grad) {
    return sguard(h, [&](mdship::Surrogate& s) {
        const size_t np = (size_t)s.mlp.n * s.mlp.p;
        if (!x || !grad || length != np) throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "length must be N*P");
        double *dx = nullptr, *dg = nullptr;
        try {
            mdship::check(hipMalloc((void**)&dx, np * sizeof(double)), "hipMalloc");
            mdship::check(hipMalloc((void**)&dg, np * sizeof(double)), "hipMalloc");
            mdship::check(hipMemcpy(dx, x, np * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
            s.mlp.predictInto<double>(dx, s.mlp.p, dg, s.mlp.stream);
            mdship::check(hipStreamSynchronize(s.mlp.stream), "hipStreamSynchronize");
            mdship::check(hipMemcpy(grad, dg, np * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
        } catch (...) {
            (void)hipFree(dx);
            (void)hipFree(dg);
            throw;
        }
        (void)hipFree(dx);
        (void)hipFree(dg);
    });
}

int mds_hip_surrogate_parameter_count(mds_hip_surrogate_handle h, long long* count) {
    return sguard(h, [&](mdship::Surrogate& s) {
        if (!count) throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "null output pointer");
        *count = s.mlp.paramCount;
    });
}

int mds_hip_surrogate_get_parameters(mds_hip_surrogate_handle h, float* out, size_t count) {
    return sguard(h, [&](mdship::Surrogate& s) {
        if (!out || count != (size_t)s.mlp.paramCount) throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "wrong length");
        mdship::check(hipStreamSynchronize(s.mlp.stream), "hipStreamSynchronize");
        mdship::check(hipMemcpy(out, s.mlp.dParams, count * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

int mds_hip_surrogate_set_parameters(mds_hip_surrogate_handle h, const float* in, size_t count) {
    return sguard(h, [&](mdship::Surrogate& s) {
        if (!in || count != (size_t)s.mlp.paramCount) throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "wrong length");
        mdship::check(hipMemcpy(s.mlp.dParams, in, count * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        s.mlp.convertWeights();
        mdship::check(hipStreamSynchronize(s.mlp.stream), "hipStreamSynchronize");
    });
}

int mds_hip_surrogate_set_normalisation(mds_hip_surrogate_handle h, const float* xCentre, const float* gCentre,
                                        size_t centreLength, const float* sigmaX, const float* sigmaG,
                                        size_t scaleLength) {
    return sguard(h, [&](mdship::Surrogate& s) {
        if (!xCentre || !gCentre || !sigmaX || !sigmaG || centreLength != (size_t)s.mlp.n * s.mlp.p ||
            scaleLength != (size_t)s.mlp.p)
            throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "centres must hold N*P values and scales P");
        s.mlp.setNormalisation(xCentre, gCentre, sigmaX, sigmaG, false);
    });
}

int mds_hip_surrogate_loss_gradient(mds_hip_surrogate_handle h, const double* x, const double* g, size_t length,
                                    float* paramGrad, size_t count, double* loss) {
    return sguard(h, [&](mdship::Surrogate& s) {
        const size_t np = (size_t)s.mlp.n * s.mlp.p;
        if (!x || !g || length != np) throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "length must be N*P");
        if (paramGrad && count != (size_t)s.mlp.paramCount) throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "wrong length");
        std::vector<float> fx(np), fg(np);
        for (size_t k = 0; k < np; ++k) {
            fx[k] = (float)x[k];
            fg[k] = (float)g[k];
        }
        float *dx = nullptr, *dg = nullptr;
        try {
            mdship::check(hipMalloc((void**)&dx, np * sizeof(float)), "hipMalloc");
            mdship::check(hipMalloc((void**)&dg, np * sizeof(float)), "hipMalloc");
            mdship::check(hipMemcpy(dx, fx.data(), np * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
            mdship::check(hipMemcpy(dg, fg.data(), np * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
            s.mlp.lossAndGradient(dx, dg);
            const double l = s.mlp.loss();
            if (loss) *loss = l;
            if (paramGrad)
                mdship::check(hipMemcpy(paramGrad, s.mlp.dGrads, count * sizeof(float), hipMemcpyDeviceToHost),
                              "hipMemcpy");
        } catch (...) {
            (void)hipFree(dx);
            (void)hipFree(dg);
            throw;
        }
        (void)hipFree(dx);
        (void)hipFree(dg);
    });
}

}  // extern "C"
