// Batched evaluation of many independent MDS instances of uneven N and mixed
// P in one persistent launch (no counterpart in the reference, whose
// instances are independent objects, src/jni/...NativeMDSSingleton.cpp:9-21).
//
// Work list.  Every dataset's packed tiles are cut into items of
// `tilesPerItem` consecutive tiles; the items of all datasets form one list,
// ordered heaviest first (larger embedding dimension first, datasets
// interleaved).  A persistent grid of waves walks the list with a fixed
// stride (wave w takes items w, w + W, ...), so the schedule is
// deterministic and every wave reaches the end of the list.  Each tile
// writes its likelihood partial and gradient slots exactly as in the
// single-instance engine; a batched reduce launch then adds them per
// dataset in a fixed order in FP64.
//
// Tuning.  wavesPerBlock, blocksPerCU, tilesPerItem and the strategy (one
// launch over all items, or one launch per embedding-dimension class, which
// lowers register pressure for small P) are chosen by an autotuner that
// times every candidate on the actual batch and caches the winner in a text
// file keyed by the batch signature.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "mds_hip.h"
#include "mds_hmc.hpp"
#include "mds_hmc_device.hpp"
#include "mds_host.hpp"
#include "mds_internal.hpp"
#include "mds_tile.hpp"
#include "mds_tune_cache.hpp"

namespace mdship {

struct BatchDataset {
    const void* delta;      // packed tiles (Real), planes as in the engine
    const void* x;          // [nb * 64][kdim] Real
    const uint32_t* tiles;  // tile codes (I << 16) | J
    void* part;             // [nb][nb][kdim][64], Real
    double* grad;           // [n][p]
    int64_t planeStride;
    double tau;
    const double* devTau;   // if set, the precision is read here (a sampler's parameter block)
    double* kickMom;        // a sampler's momenta [n][p] and parameter block, for reduces that
    const double* kickParams;  // kick each gradient row they write (null otherwise)
    int nb, n, p, kdim;
    int truncated, symmetric;
    int tileSumBase, numTiles;
};

struct BatchItem {
    uint32_t ds;
    uint32_t tile;   // first tile of the run
    uint32_t count;  // tiles in the run
    uint32_t pad;
};

struct ReduceItem {
    int ds;
    int K;  // >= 0: gradient rows of row block K; -1: likelihood sum
};

constexpr int kBatchMaxThreads = 512;

// FLAGS >= 0: every item of the launch has truncation = FLAGS & 1 and
// symmetric observations = FLAGS & 2 (one walk compiled, fewer registers);
// FLAGS = -1: the walk is chosen per item.  The walks are the same either
// way, so are the results.
template <typename Real, int P, int FLAGS = -1>
__device__ __forceinline__ void batch_tile(const BatchDataset& d, uint32_t t, double* __restrict__ tileSums,
                                           const Real* lnPhiTab, Real* stage) {
    const uint32_t code = d.tiles[t];
    const int I = (int)(code >> 16), J = (int)(code & 0xffffu);
    const Real* base = (const Real*)d.delta + (int64_t)t * kTileElems;
    const Real* x = (const Real*)d.x;
    const PairConst<Real> pc = make_pair_const<Real>(d.devTau ? *d.devTau : d.tau, lnPhiTab);
    double* sum = tileSums + d.tileSumBase + t;
    Real* part = (Real*)d.part;
    // the previous tile's reads of the staged columns precede the new writes
    wave_lds_sync();
    stage_columns<Real, P>(stage, x, J);
    wave_lds_sync();
    if constexpr (FLAGS >= 0) {
        tile_walk<Real, P, (FLAGS & 1) != 0, (FLAGS & 2) != 0, true, kGroup>(base, d.planeStride, x, stage, I, J,
                                                                             d.nb, pc, sum, part);
    } else if (d.truncated) {
        if (d.symmetric)
            tile_walk<Real, P, true, true, true, kGroup>(base, d.planeStride, x, stage, I, J, d.nb, pc, sum, part);
        else
            tile_walk<Real, P, true, false, true, kGroup>(base, d.planeStride, x, stage, I, J, d.nb, pc, sum,
                                                          part);
    } else {
        if (d.symmetric)
            tile_walk<Real, P, false, true, true, kGroup>(base, d.planeStride, x, stage, I, J, d.nb, pc, sum,
                                                          part);
        else
            tile_walk<Real, P, false, false, true, kGroup>(base, d.planeStride, x, stage, I, J, d.nb, pc, sum,
                                                           part);
    }
}

template <typename Real>
__device__ __forceinline__ void batch_tile_any(const BatchDataset& d, uint32_t t, double* __restrict__ tileSums,
                                               const Real* lnPhiTab, Real* stage) {
    switch (d.kdim) {
        case 1: batch_tile<Real, 1>(d, t, tileSums, lnPhiTab, stage); break;
        case 2: batch_tile<Real, 2>(d, t, tileSums, lnPhiTab, stage); break;
        case 3: batch_tile<Real, 3>(d, t, tileSums, lnPhiTab, stage); break;
        case 4: batch_tile<Real, 4>(d, t, tileSums, lnPhiTab, stage); break;
        case 5: batch_tile<Real, 5>(d, t, tileSums, lnPhiTab, stage); break;
        case 6: batch_tile<Real, 6>(d, t, tileSums, lnPhiTab, stage); break;
        case 8: batch_tile<Real, 8>(d, t, tileSums, lnPhiTab, stage); break;
        case 10: batch_tile<Real, 10>(d, t, tileSums, lnPhiTab, stage); break;
        default: break;  // the host only admits kernel dimensions <= 10 here
    }
}

// KDIM = 0: items of any embedding dimension (switch per item); KDIM > 0:
// every item has that kernel dimension.
// Staged column elements per wavefront of a launch (see ColumnStage).
template <typename Real, int KDIM>
constexpr int batch_stage_elems() {
    return KDIM == 0 ? ColumnStage<Real, 1>::kMaxElems : ColumnStage<Real, KDIM>::kElems;
}

template <typename Real, int KDIM, int FLAGS = -1>
__global__ __launch_bounds__(kBatchMaxThreads) void batch_tile_kernel(const BatchDataset* __restrict__ ds,
                                                                       const BatchItem* __restrict__ items,
                                                                       int numItems,
                                                                       double* __restrict__ tileSums) {
    __shared__ __align__(16) Real lnPhiTab[TableOf<Real>::kCount];  // truncated datasets read it
    // staged columns: batch_stage_elems values per wavefront, sized by the launch
    extern __shared__ __align__(16) unsigned char batchColStage[];
    copy_lnphi_table(lnPhiTab);
    __syncthreads();
    Real* stage = (Real*)batchColStage + (threadIdx.x / kWave) * batch_stage_elems<Real, KDIM>();
    const int wavesPerBlock = blockDim.x / kWave;
    const int stride = gridDim.x * wavesPerBlock;
    for (int w = __builtin_amdgcn_readfirstlane(blockIdx.x * wavesPerBlock + threadIdx.x / kWave); w < numItems;
         w += stride) {
        const BatchItem it = items[w];
        const BatchDataset d = ds[it.ds];
        for (uint32_t c = 0; c < it.count; ++c) {
            if constexpr (KDIM == 0)
                batch_tile_any<Real>(d, it.tile + c, tileSums, lnPhiTab, stage);
            else
                batch_tile<Real, KDIM, FLAGS>(d, it.tile + c, tileSums, lnPhiTab, stage);
        }
    }
}

constexpr int kBatchReduceWaves = 8;

// Gradient rows of row block K of one dataset: each of the 8 wavefronts adds
// the column blocks J = w mod 8 for every coordinate at once (the loads of
// kAhead column blocks of all KD coordinates in flight together), then the
// wavefront partials are combined per coordinate in wavefront order.  Per
// coordinate the additions are those of the single-instance reduce, in the
// same order, so the rows are bitwise the same.  KICK >= 0: each row's
// thread then applies the sampler's kick of that form to its row, as the
// chains' kick kernel would (kick_row: the same bits, one launch fewer per
// leapfrog step; no reduce block reads locations).
template <typename PartT, int KD, int KICK>
__device__ __forceinline__ void batch_reduce_rows(const BatchDataset& d, int K, double* red) {
    constexpr int kAhead = 4;
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    double acc[KD];
#pragma unroll
    for (int p = 0; p < KD; ++p) acc[p] = 0.0;
    const int64_t stride = (int64_t)KD * kTile;  // between column blocks J
    const PartT* base = (const PartT*)d.part + (int64_t)K * d.nb * stride + lane;
    int J = wave;
    for (; J + (kAhead - 1) * kBatchReduceWaves < d.nb; J += kAhead * kBatchReduceWaves) {
        double v[kAhead][KD];
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            const PartT* slot = base + (J + u * kBatchReduceWaves) * stride;
#pragma unroll
            for (int p = 0; p < KD; ++p) v[u][p] = (double)slot[p * kTile];
        }
#pragma unroll
        for (int u = 0; u < kAhead; ++u)
#pragma unroll
            for (int p = 0; p < KD; ++p) acc[p] += v[u][p];
    }
    for (; J < d.nb; J += kBatchReduceWaves) {
        const PartT* slot = base + J * stride;
#pragma unroll
        for (int p = 0; p < KD; ++p) acc[p] += (double)slot[p * kTile];
    }
    const int row = K * kTile + lane;
#pragma unroll
    for (int p = 0; p < KD; ++p) {
        red[tid] = acc[p];
        __syncthreads();
        if (wave == 0) {
            double v = 0.0;
            for (int w = 0; w < kBatchReduceWaves; ++w) v += red[w * kWave + lane];
            if (row < d.n && p < d.p) d.grad[(int64_t)row * d.p + p] = v;
        }
        __syncthreads();
    }
    if constexpr (KICK >= 0) {
        if (wave == 0) kick_row<PartT, KICK>((PartT*)d.x, d.kickMom, d.grad, d.kickParams, d.n, d.p, KD, row);
    }
}

template <typename PartT, int KICK>
__global__ __launch_bounds__(kBatchReduceWaves * kWave) void batch_reduce_kernel(
    const BatchDataset* __restrict__ ds, const ReduceItem* __restrict__ items,
    const double* __restrict__ tileSums, double* __restrict__ sums) {
    __shared__ double red[kBatchReduceWaves * kWave];
    const ReduceItem item = items[blockIdx.x];
    const BatchDataset d = ds[item.ds];
    const int tid = threadIdx.x;
    if (item.K < 0) {
        // loads issued kAhead iterations at a time, added in the plain loop's order
        constexpr int kAhead = 4;
        constexpr int kThreads = kBatchReduceWaves * kWave;
        double acc = 0.0;
        const double* ts = tileSums + d.tileSumBase;
        int k = tid;
        for (; k + (kAhead - 1) * kThreads < d.numTiles; k += kAhead * kThreads) {
            double v[kAhead];
#pragma unroll
            for (int u = 0; u < kAhead; ++u) v[u] = ts[k + u * kThreads];
#pragma unroll
            for (int u = 0; u < kAhead; ++u) acc += v[u];
        }
        for (; k < d.numTiles; k += kThreads) acc += ts[k];
        red[tid] = acc;
        __syncthreads();
        for (int s = kBatchReduceWaves * kWave / 2; s > 0; s >>= 1) {
            if (tid < s) red[tid] += red[tid + s];
            __syncthreads();
        }
        if (tid == 0) sums[item.ds] = 0.5 * red[0];
        return;
    }
    switch (d.kdim) {  // uniform over the block
        case 1: batch_reduce_rows<PartT, 1, KICK>(d, item.K, red); break;
        case 2: batch_reduce_rows<PartT, 2, KICK>(d, item.K, red); break;
        case 3: batch_reduce_rows<PartT, 3, KICK>(d, item.K, red); break;
        case 4: batch_reduce_rows<PartT, 4, KICK>(d, item.K, red); break;
        case 5: batch_reduce_rows<PartT, 5, KICK>(d, item.K, red); break;
        case 6: batch_reduce_rows<PartT, 6, KICK>(d, item.K, red); break;
        case 8: batch_reduce_rows<PartT, 8, KICK>(d, item.K, red); break;
        case 10: batch_reduce_rows<PartT, 10, KICK>(d, item.K, red); break;
        default: break;  // the host only admits kernel dimensions <= 10 here
    }
}

// ---------------------------------------------------------------- host

struct BatchConfig {
    int wavesPerBlock = 4;
    int blocksPerCU = 8;
    int tilesPerItem = 1;
    int strategy = 0;  // 0: one launch; 1: one launch per kernel-dimension class
};

struct BatchEntry {
    int P = 0, N = 0, kdim = 0, nb = 0, numTiles = 0;
    bool symmetric = true, dataSet = false, precisionSet = false;
    double tau = 0.0;
    const double* devTau = nullptr;  // set by a sampler for the length of its run
    double* kickMom = nullptr;       // likewise: the chain's momenta and parameter
    const double* kickParams = nullptr;  // block, for reduces that kick (enqueue)
    void* dDelta = nullptr;          // owned packed observations (null while shared)
    size_t deltaBytes = 0;
    int shareSource = -1;            // >= 0: reads that entry's packed observations
    void* dX = nullptr;
    void* dPart = nullptr;  // gradient partials in the batch's real type
    double* dGrad = nullptr;
    uint32_t* dTiles = nullptr;
    std::vector<uint32_t> tiles;
    int tileSumBase = 0;
};

struct Batch {
    int device = 0;
    bool fp32 = false, truncated = false;
    int cuCount = 0;
    size_t realSize = 8;
    hipStream_t stream = nullptr;
    // One launch per dimension class runs its classes side by side: the
    // first on `stream`, the others on side streams forked from it and
    // joined at the end, each class followed on its stream by the reduce of
    // its own datasets, so one class's tail and reduce overlap the other
    // classes' work (the results do not depend on the order: every tile
    // writes its own slots, every dataset is reduced by one launch).
    static constexpr int kSideStreams = 3;
    hipStream_t side[kSideStreams] = {};
    hipEvent_t forkEvent = nullptr;
    hipEvent_t joinEvent[kSideStreams] = {};
    std::vector<BatchEntry> e;
    double* dTileSums = nullptr;
    double* dSums = nullptr;
    double* hSums = nullptr;
    BatchDataset* dDesc = nullptr;
    BatchItem* dItems = nullptr;
    ReduceItem* dReduce = nullptr;
    int numReduce = 0;
    // launch plan: (kdim class or 0, first item, item count, first reduce
    // item, reduce item count, the flags its datasets share or -1); each
    // launch's reduce follows it on its stream
    struct Launch {
        int kdim, first, count, redFirst, redCount, flags;
    };
    std::vector<Launch> plan;
    BatchConfig cfg;
    bool planDirty = true;
    bool timing = false;
    std::vector<hipEvent_t> pending;  // pairs: start, end
    double timedMs = 0.0;
    long long timedCount = 0;
    std::string lastError;

    Batch(int count, const int* dims, const int* ns, long flags, int dev) {
        if (count < 1) throw Error(MDS_HIP_ERR_ARGUMENT, "batch needs at least one dataset");
        int devices = 0;
        check(hipGetDeviceCount(&devices), "hipGetDeviceCount");
        if (devices <= 0) throw Error(MDS_HIP_ERR_DEVICE, "no HIP device");
        device = (dev < 0 || dev >= devices) ? 0 : dev;
        check(hipSetDevice(device), "hipSetDevice");
        check(hipDeviceGetAttribute(&cuCount, hipDeviceAttributeMultiprocessorCount, device), "CU count");
        fp32 = (flags & MDS_HIP_FLAG_FLOAT) != 0;
        truncated = (flags & MDS_HIP_FLAG_LEFT_TRUNCATION) != 0;
        realSize = fp32 ? 4 : 8;
        e.resize(count);
        int sumBase = 0;
        for (int k = 0; k < count; ++k) {
            BatchEntry& b = e[k];
            b.P = dims[k];
            b.N = ns[k];
            if (b.P < 1 || b.P > 10)
                throw Error(MDS_HIP_ERR_ARGUMENT, "batched embedding dimension must be in [1, 10]");
            if (b.N < 1) throw Error(MDS_HIP_ERR_ARGUMENT, "location count must be >= 1");
            b.kdim = kernel_dim(b.P);
            b.nb = (b.N + kTile - 1) / kTile;
            if (b.nb > 0xffff) throw Error(MDS_HIP_ERR_ARGUMENT, "location count too large");
            b.tiles = make_tiles(b.nb);
            b.numTiles = (int)b.tiles.size();
            b.tileSumBase = sumBase;
            sumBase += b.numTiles;
        }
        try {
            check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
            for (int k = 0; k < kSideStreams; ++k) {
                check(hipStreamCreateWithFlags(&side[k], hipStreamNonBlocking), "hipStreamCreate");
                check(hipEventCreateWithFlags(&joinEvent[k], hipEventDisableTiming), "hipEventCreate");
            }
            check(hipEventCreateWithFlags(&forkEvent, hipEventDisableTiming), "hipEventCreate");
            for (BatchEntry& b : e) {
                const size_t xBytes = (size_t)b.nb * kTile * b.kdim * realSize;
                check(hipMalloc(&b.dX, xBytes), "hipMalloc(locations)");
                check(hipMemsetAsync(b.dX, 0, xBytes, stream), "hipMemset(locations)");
                check(hipMalloc(&b.dPart, (size_t)b.nb * b.nb * b.kdim * kTile * realSize), "hipMalloc(partials)");
                check(hipMalloc((void**)&b.dGrad, (size_t)b.N * b.P * sizeof(double)), "hipMalloc(gradient)");
                check(hipMalloc((void**)&b.dTiles, b.tiles.size() * sizeof(uint32_t)), "hipMalloc(tiles)");
                check(hipMemcpyAsync(b.dTiles, b.tiles.data(), b.tiles.size() * sizeof(uint32_t),
                                     hipMemcpyHostToDevice, stream),
                      "hipMemcpy(tiles)");
            }
            check(hipMalloc((void**)&dTileSums, (size_t)sumBase * sizeof(double)), "hipMalloc(tile sums)");
            check(hipMalloc((void**)&dSums, e.size() * sizeof(double)), "hipMalloc(sums)");
            check(hipHostMalloc((void**)&hSums, e.size() * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
            check(hipMalloc((void**)&dDesc, e.size() * sizeof(BatchDataset)), "hipMalloc(descriptors)");
            check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        } catch (...) {
            release();
            throw;
        }
    }

    ~Batch() { release(); }

    void release() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t ev : pending) (void)hipEventDestroy(ev);
        pending.clear();
        for (BatchEntry& b : e) {
            (void)hipFree(b.dDelta);
            (void)hipFree(b.dX);
            (void)hipFree(b.dPart);
            (void)hipFree(b.dGrad);
            (void)hipFree(b.dTiles);
            b.dDelta = b.dX = b.dPart = nullptr;
            b.dGrad = nullptr;
            b.dTiles = nullptr;
        }
        (void)hipFree(dTileSums);
        (void)hipFree(dSums);
        (void)hipHostFree(hSums);
        (void)hipFree(dDesc);
        (void)hipFree(dItems);
        (void)hipFree(dReduce);
        for (int k = 0; k < kSideStreams; ++k) {
            if (side[k]) {
                (void)hipStreamSynchronize(side[k]);
                (void)hipStreamDestroy(side[k]);
            }
            if (joinEvent[k]) (void)hipEventDestroy(joinEvent[k]);
            side[k] = nullptr;
            joinEvent[k] = nullptr;
        }
        if (forkEvent) (void)hipEventDestroy(forkEvent);
        forkEvent = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        dTileSums = dSums = hSums = nullptr;
        dDesc = nullptr;
        dItems = nullptr;
        dReduce = nullptr;
        stream = nullptr;
    }

    BatchEntry& at(int index) {
        if (index < 0 || index >= (int)e.size()) throw Error(MDS_HIP_ERR_ARGUMENT, "dataset index out of range");
        return e[index];
    }

    // Entry `index` reads the packed observations of entry `source` (same
    // N; P may differ), e.g. several chains of one dataset: no copy.  Later
    // data for `source` are seen by `index` too; data set on `index` itself
    // end the sharing.
    void sharePairwiseData(int index, int source) {
        BatchEntry& b = at(index);
        const BatchEntry& src = at(source);
        if (index == source) throw Error(MDS_HIP_ERR_ARGUMENT, "an entry cannot share its own data");
        if (src.shareSource >= 0) throw Error(MDS_HIP_ERR_ARGUMENT, "the source entry shares another's data");
        if (!src.dataSet) throw Error(MDS_HIP_ERR_STATE, "the source entry has no data");
        if (src.N != b.N) throw Error(MDS_HIP_ERR_ARGUMENT, "shared data need the same location count");
        for (const BatchEntry& o : e)
            if (o.shareSource == index) throw Error(MDS_HIP_ERR_ARGUMENT, "other entries share this entry's data");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        check(hipFree(b.dDelta), "hipFree");
        b.dDelta = nullptr;
        b.deltaBytes = 0;
        b.shareSource = source;
        b.dataSet = true;
        planDirty = true;
    }

    void setPairwiseData(int index, const double* data, size_t length) {
        BatchEntry& b = at(index);
        if (!data || length != (size_t)b.N * b.N)
            throw Error(MDS_HIP_ERR_ARGUMENT, "setPairwiseData: length must be N*N");
        b.shareSource = -1;
        const size_t planeElems = (size_t)b.numTiles * kTileElems;
        std::unique_ptr<char[]> host(new char[2 * planeElems * realSize]);
        const bool sym = fp32 ? pack_tiles<float>(data, b.N, b.tiles, (float*)host.get(),
                                                  (float*)host.get() + planeElems)
                              : pack_tiles<double>(data, b.N, b.tiles, (double*)host.get(),
                                                   (double*)host.get() + planeElems);
        const size_t bytes = (sym ? 1 : 2) * planeElems * realSize;
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        if (bytes != b.deltaBytes) {
            check(hipFree(b.dDelta), "hipFree");
            b.dDelta = nullptr;
            b.deltaBytes = 0;
            check(hipMalloc(&b.dDelta, bytes), "hipMalloc(observations)");
            b.deltaBytes = bytes;
        }
        check(hipMemcpy(b.dDelta, host.get(), bytes, hipMemcpyHostToDevice), "hipMemcpy(observations)");
        b.symmetric = sym;
        b.dataSet = true;
        planDirty = true;
    }

    void updateLocations(int index, const double* loc, size_t length) {
        BatchEntry& b = at(index);
        if (!loc || length != (size_t)b.N * b.P)
            throw Error(MDS_HIP_ERR_ARGUMENT, "updateLocations: length must be N*P");
        const size_t bytes = (size_t)b.N * b.kdim * realSize;
        std::vector<char> staging(bytes);
        for (int r = 0; r < b.N; ++r)
            for (int k = 0; k < b.kdim; ++k) {
                const double v = k < b.P ? loc[(size_t)r * b.P + k] : 0.0;
                if (fp32)
                    ((float*)staging.data())[(size_t)r * b.kdim + k] = (float)v;
                else
                    ((double*)staging.data())[(size_t)r * b.kdim + k] = v;
            }
        check(hipMemcpyAsync(b.dX, staging.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpy(locations)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    void setParameters(int index, const double* data, size_t length) {
        BatchEntry& b = at(index);
        if (!data || length != 1) throw Error(MDS_HIP_ERR_ARGUMENT, "setParameters: expects the precision");
        if (!std::isfinite(data[0]) || data[0] <= 0.0)
            throw Error(MDS_HIP_ERR_ARGUMENT, "setParameters: precision must be finite and > 0");
        b.tau = data[0];
        b.precisionSet = true;
        planDirty = true;
    }

    std::string signature() const {
        std::vector<std::pair<int, int>> keys;
        for (const BatchEntry& b : e) keys.push_back({b.kdim, b.nb});
        std::sort(keys.begin(), keys.end());
        std::ostringstream s;
        s << (fp32 ? "f32" : "f64") << (truncated ? "-T" : "-N") << "-cu" << cuCount;
        for (auto& k : keys) s << "-" << k.first << "x" << k.second;
        return s.str();
    }

    void buildPlan() {
        for (size_t k = 0; k < e.size(); ++k) {
            if (!e[k].dataSet) throw Error(MDS_HIP_ERR_STATE, "pairwise data not set for dataset " + std::to_string(k));
            if (!e[k].precisionSet) throw Error(MDS_HIP_ERR_STATE, "precision not set for dataset " + std::to_string(k));
        }
        std::vector<BatchDataset> desc(e.size());
        for (size_t k = 0; k < e.size(); ++k) {
            const BatchEntry& b = e[k];
            const BatchEntry& data = b.shareSource >= 0 ? e[b.shareSource] : b;
            BatchDataset& d = desc[k];
            d.delta = data.dDelta;
            d.x = b.dX;
            d.tiles = b.dTiles;
            d.part = b.dPart;
            d.grad = b.dGrad;
            d.planeStride = (int64_t)b.numTiles * kTileElems;
            d.tau = b.tau;
            d.devTau = b.devTau;
            d.kickMom = b.kickMom;
            d.kickParams = b.kickParams;
            d.nb = b.nb;
            d.n = b.N;
            d.p = b.P;
            d.kdim = b.kdim;
            d.truncated = truncated ? 1 : 0;
            d.symmetric = data.symmetric ? 1 : 0;
            d.tileSumBase = b.tileSumBase;
            d.numTiles = b.numTiles;
        }
        // items: per class (kdim, heaviest first), datasets interleaved run by
        // run; reduce items: the total and the row blocks of each dataset, in
        // class order, so that a class's reduce is one contiguous range
        std::map<int, std::vector<int>, std::greater<int>> classes;
        for (size_t k = 0; k < e.size(); ++k) classes[e[k].kdim].push_back((int)k);
        std::vector<BatchItem> items;
        std::vector<ReduceItem> red;
        plan.clear();
        const int tpi = std::max(1, cfg.tilesPerItem);
        for (auto& cls : classes) {
            const int first = (int)items.size();
            const int redFirst = (int)red.size();
            std::vector<int> cursor(cls.second.size(), 0);
            bool more = true;
            while (more) {
                more = false;
                for (size_t q = 0; q < cls.second.size(); ++q) {
                    const BatchEntry& b = e[cls.second[q]];
                    if (cursor[q] >= b.numTiles) continue;
                    BatchItem it;
                    it.ds = (uint32_t)cls.second[q];
                    it.tile = (uint32_t)cursor[q];
                    it.count = (uint32_t)std::min(tpi, b.numTiles - cursor[q]);
                    it.pad = 0;
                    items.push_back(it);
                    cursor[q] += it.count;
                    more = true;
                }
            }
            for (int k : cls.second) {
                red.push_back({k, -1});
                for (int K = 0; K < e[k].nb; ++K) red.push_back({k, K});
            }
            // the walk variant every dataset of the class takes, if they agree
            int flags = -2;
            for (int k : cls.second) {
                const int f = desc[k].truncated | (desc[k].symmetric << 1);
                flags = flags == -2 || flags == f ? f : -1;
            }
            plan.push_back({cls.first, first, (int)items.size() - first, redFirst, (int)red.size() - redFirst, flags});
        }
        if (cfg.strategy == 0) plan = {{0, 0, (int)items.size(), 0, (int)red.size(), -1}};
        numReduce = (int)red.size();
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        (void)hipFree(dItems);
        (void)hipFree(dReduce);
        dItems = nullptr;
        dReduce = nullptr;
        check(hipMalloc((void**)&dItems, items.size() * sizeof(BatchItem)), "hipMalloc(items)");
        check(hipMalloc((void**)&dReduce, red.size() * sizeof(ReduceItem)), "hipMalloc(reduce items)");
        check(hipMemcpy(dItems, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemcpy(dReduce, red.data(), red.size() * sizeof(ReduceItem), hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemcpy(dDesc, desc.data(), desc.size() * sizeof(BatchDataset), hipMemcpyHostToDevice), "hipMemcpy");
        planDirty = false;
    }

    template <typename Real, int KDIM, int FLAGS>
    void launchTilesAs(const BatchItem* items, int count, hipStream_t stream) {
        const int wpb = cfg.wavesPerBlock;
        const int want = (count + wpb - 1) / wpb;
        const int grid = std::max(1, std::min(want, cuCount * cfg.blocksPerCU));
        const size_t stageBytes = (size_t)wpb * batch_stage_elems<Real, KDIM>() * sizeof(Real);
        hipLaunchKernelGGL((batch_tile_kernel<Real, KDIM, FLAGS>), dim3(grid), dim3(wpb * kWave), stageBytes, stream,
                           dDesc, items, count, dTileSums);
    }

    // flags: the walk variant every item of the launch takes, or -1
    template <typename Real, int KDIM>
    void launchTiles(const BatchItem* items, int count, hipStream_t stream, int flags) {
        if constexpr (KDIM > 0) {
            switch (flags) {
                case 0: launchTilesAs<Real, KDIM, 0>(items, count, stream); return;
                case 1: launchTilesAs<Real, KDIM, 1>(items, count, stream); return;
                case 2: launchTilesAs<Real, KDIM, 2>(items, count, stream); return;
                case 3: launchTilesAs<Real, KDIM, 3>(items, count, stream); return;
                default: break;
            }
        }
        launchTilesAs<Real, KDIM, -1>(items, count, stream);
    }

    template <typename Real>
    void launchClass(int kdim, const BatchItem* items, int count, hipStream_t stream, int flags) {
        switch (kdim) {
            case 0: launchTiles<Real, 0>(items, count, stream, flags); break;
            case 1: launchTiles<Real, 1>(items, count, stream, flags); break;
            case 2: launchTiles<Real, 2>(items, count, stream, flags); break;
            case 3: launchTiles<Real, 3>(items, count, stream, flags); break;
            case 4: launchTiles<Real, 4>(items, count, stream, flags); break;
            case 5: launchTiles<Real, 5>(items, count, stream, flags); break;
            case 6: launchTiles<Real, 6>(items, count, stream, flags); break;
            case 8: launchTiles<Real, 8>(items, count, stream, flags); break;
            case 10: launchTiles<Real, 10>(items, count, stream, flags); break;
            default: throw Error(MDS_HIP_ERR_ARGUMENT, "unsupported kernel dimension");
        }
        check(hipGetLastError(), "batch tile kernel");
    }

    template <int KICK>
    void launchReduceKick(const ReduceItem* items, int count, hipStream_t s) {
        if (fp32)
            hipLaunchKernelGGL((batch_reduce_kernel<float, KICK>), dim3(count), dim3(kBatchReduceWaves * kWave), 0,
                               s, dDesc, items, dTileSums, dSums);
        else
            hipLaunchKernelGGL((batch_reduce_kernel<double, KICK>), dim3(count), dim3(kBatchReduceWaves * kWave), 0,
                               s, dDesc, items, dTileSums, dSums);
    }

    void launchReduce(const ReduceItem* items, int count, hipStream_t s, int kickForm) {
        switch (kickForm) {
            case -1: launchReduceKick<-1>(items, count, s); break;
            case kKick: launchReduceKick<kKick>(items, count, s); break;
            case kKick2Drift: launchReduceKick<kKick2Drift>(items, count, s); break;
            default: throw Error(MDS_HIP_ERR_ARGUMENT, "unsupported kick form");
        }
        check(hipGetLastError(), "batch reduce kernel");
    }

    // Every entry's likelihood and gradient; kickForm >= 0 (a sampler's
    // leapfrog step, with every entry's kickMom and kickParams set): the
    // reduces also apply that kick to every gradient row.
    void enqueue(int kickForm = -1) {
        if (planDirty) buildPlan();
        int used = 0;  // side streams forked
        if (plan.size() > 1) check(hipEventRecord(forkEvent, stream), "hipEventRecord(fork)");
        for (size_t c = 0; c < plan.size(); ++c) {
            const Launch& l = plan[c];
            if (l.count == 0) continue;
            hipStream_t s = stream;
            if (c > 0) {
                const int k = (int)((c - 1) % kSideStreams);
                s = side[k];
                if (k >= used) {
                    check(hipStreamWaitEvent(s, forkEvent, 0), "hipStreamWaitEvent(fork)");
                    used = k + 1;
                }
            }
            if (fp32)
                launchClass<float>(l.kdim, dItems + l.first, l.count, s, l.flags);
            else
                launchClass<double>(l.kdim, dItems + l.first, l.count, s, l.flags);
            // the class's datasets reduce while other classes still walk
            if (l.redCount > 0) launchReduce(dReduce + l.redFirst, l.redCount, s, kickForm);
        }
        for (int k = 0; k < used; ++k) {
            check(hipEventRecord(joinEvent[k], side[k]), "hipEventRecord(join)");
            check(hipStreamWaitEvent(stream, joinEvent[k], 0), "hipStreamWaitEvent(join)");
        }
    }

    void evaluate() {
        hipEvent_t a = nullptr, b = nullptr;
        if (timing) {
            check(hipEventCreate(&a), "hipEventCreate");
            check(hipEventCreate(&b), "hipEventCreate");
            check(hipEventRecord(a, stream), "hipEventRecord");
        }
        enqueue();
        if (timing) {
            check(hipEventRecord(b, stream), "hipEventRecord");
            pending.push_back(a);
            pending.push_back(b);
        }
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    double timeConfig(int reps) {
        hipEvent_t a, b;
        check(hipEventCreate(&a), "hipEventCreate");
        check(hipEventCreate(&b), "hipEventCreate");
        enqueue();  // warm-up (also builds the plan)
        check(hipEventRecord(a, stream), "hipEventRecord");
        for (int r = 0; r < reps; ++r) enqueue();
        check(hipEventRecord(b, stream), "hipEventRecord");
        check(hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return ms / reps;
    }

    // Returns the chosen configuration's mean time per evaluation (ms).
    double autotune(const char* cachePath, int force) {
        const std::string sig = signature();
        if (cachePath && !force) {
            TuneRecord r;
            if (tune_cache_find(cachePath, sig, 4, r)) {
                cfg = {(int)r.fields[0], (int)r.fields[1], (int)r.fields[2], (int)r.fields[3]};
                planDirty = true;
                return r.ms;
            }
        }
        BatchConfig best = cfg;
        double bestMs = 1e300;
        for (int strategy : {0, 1})
            for (int wpb : {1, 2, 4, 8})
                for (int bpc : {1, 2, 4, 8})
                    for (int tpi : {1, 2, 4}) {
                        if (wpb * bpc < 4 || wpb * bpc > 32) continue;
                        cfg = {wpb, bpc, tpi, strategy};
                        planDirty = true;
                        const double ms = timeConfig(3);
                        if (ms < bestMs) {
                            bestMs = ms;
                            best = cfg;
                        }
                    }
        cfg = best;
        planDirty = true;
        if (cachePath) {
            TuneRecord r;
            r.fields = {cfg.wavesPerBlock, cfg.blocksPerCU, cfg.tilesPerItem, cfg.strategy};
            r.ms = bestMs;
            (void)tune_cache_store(cachePath, sig, r);
        }
        return bestMs;
    }

    void getSums(double* out, size_t length) {
        if (!out || length != e.size()) throw Error(MDS_HIP_ERR_ARGUMENT, "sums: length must be the dataset count");
        check(hipMemcpyAsync(hSums, dSums, e.size() * sizeof(double), hipMemcpyDeviceToHost, stream), "hipMemcpy");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        for (size_t k = 0; k < e.size(); ++k)
            out[k] = truncated ? hSums[k] : 0.5 * e[k].tau * hSums[k];
    }

    void getGradient(int index, double* out, size_t length) {
        BatchEntry& b = at(index);
        if (!out || length != (size_t)b.N * b.P) throw Error(MDS_HIP_ERR_ARGUMENT, "gradient: length must be N*P");
        check(hipMemcpyAsync(out, b.dGrad, length * sizeof(double), hipMemcpyDeviceToHost, stream), "hipMemcpy");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    void getTiming(double* ms, long long* count) {
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        for (size_t k = 0; k + 1 < pending.size(); k += 2) {
            float t = 0.f;
            check(hipEventElapsedTime(&t, pending[k], pending[k + 1]), "hipEventElapsedTime");
            timedMs += t;
            ++timedCount;
            (void)hipEventDestroy(pending[k]);
            (void)hipEventDestroy(pending[k + 1]);
        }
        pending.clear();
        if (ms) *ms = timedMs;
        if (count) *count = timedCount;
    }
};

// ---------------------------------------------------------------- sampler
//
// Many chains in one launch per leapfrog step: the HMC sampler of
// mds_hmc.hip (R/Untitled.R:477-729, no tree) run over every entry of a
// batch at once, one chain per entry -- several chains of one dataset
// (entries sharing its observations) or chains of different datasets of
// their own N and P.  Per leapfrog step one batched tile launch and one
// batched reduce evaluate every chain's likelihood and gradient and one
// kick launch moves every chain; the trajectory of L steps is captured once
// as a hipGraph and replayed per iteration.  Each chain keeps its own
// precision, step size and trait precision in a device parameter block
// (the tile kernel reads its precision there), its own random streams
// (seed + chain index, as a single-chain run with that seed) and its own
// Metropolis tests in FP64 on the host, in the single-chain sampler's
// order, so chain k gives the bits of mds_hip_hmc_run with seed + k.

struct ChainDesc {
    void* x;                // the entry's locations, Real [nb * 64][kdim]
    void* xSave;            // their copy at the start of an iteration
    int64_t xCount;         // nb * 64 * kdim
    double* mom;            // [n][p]
    double* grad;           // the entry's gradient [n][p]
    double* gradSave;       // its copy
    const double* params;   // kHmcParamCount values
    double* rows;           // [kTraceRows][p], pinned host memory
    double* gram;           // [p][p], pinned host memory
    TraceRows traceRows;
    uint64_t seed;
    int n, p, kdim;
};

constexpr int kChainThreads = 256;

// The start of an iteration for every chain in one launch: the momenta
// (momentum_pair) and the copies of the locations and gradient that a
// rejected move restores.
template <typename Real>
__global__ __launch_bounds__(kChainThreads) void chains_start_kernel(const ChainDesc* __restrict__ cd,
                                                                     uint64_t iteration) {
    const ChainDesc& d = cd[blockIdx.y];
    const int64_t t = (int64_t)blockIdx.x * kChainThreads + threadIdx.x;
    momentum_pair(d.mom, d.n * d.p, d.seed, iteration, (int)t);
    if (t < d.xCount) ((Real*)d.xSave)[t] = ((const Real*)d.x)[t];
    if (t < (int64_t)d.n * d.p) d.gradSave[t] = d.grad[t];
}

// Chains chain0 + y for the set bits y of mask (up to kMaskChains per launch):
// what & 1 the locations, what & 2 the gradient, copied to their saves or
// (restore) back from them.
constexpr int kMaskChains = 256;
struct ChainMask {
    uint32_t w[kMaskChains / 32];
};

template <typename Real>
__global__ __launch_bounds__(kChainThreads) void chains_copy_kernel(const ChainDesc* __restrict__ cd, ChainMask mask,
                                                                    int chain0, int what, int restore) {
    if (!((mask.w[blockIdx.y >> 5] >> (blockIdx.y & 31)) & 1u)) return;  // the same for the whole block
    const ChainDesc& d = cd[chain0 + blockIdx.y];
    const int64_t t = (int64_t)blockIdx.x * kChainThreads + threadIdx.x;
    if ((what & 1) && t < d.xCount) {
        Real* x = (Real*)d.x;
        Real* xs = (Real*)d.xSave;
        if (restore)
            x[t] = xs[t];
        else
            xs[t] = x[t];
    }
    if ((what & 2) && t < (int64_t)d.n * d.p) {
        if (restore)
            d.grad[t] = d.gradSave[t];
        else
            d.gradSave[t] = d.grad[t];
    }
}

template <typename Real, int FORM>
__global__ __launch_bounds__(kChainThreads) void chains_kick_kernel(const ChainDesc* __restrict__ cd) {
    const ChainDesc& d = cd[blockIdx.y];
    kick_row<Real, FORM>((Real*)d.x, d.mom, d.grad, d.params, d.n, d.p, d.kdim, blockIdx.x * kChainThreads + threadIdx.x);
}

// Chain c's energies into out[2 c], out[2 c + 1]; with sumsIn, its last
// pass's sum into sumsOut[c] as well.
template <typename Real>
__global__ __launch_bounds__(kEnergyThreads) void chains_energy_kernel(const ChainDesc* __restrict__ cd,
                                                                       double* __restrict__ out,
                                                                       const double* __restrict__ sumsIn,
                                                                       double* __restrict__ sumsOut) {
    __shared__ double red[2][kEnergyThreads];
    const ChainDesc& d = cd[blockIdx.x];
    energy_block<Real>((const Real*)d.x, d.mom, d.params, d.n, d.p, d.kdim, out + 2 * blockIdx.x, red);
    if (sumsIn && threadIdx.x == 0) sumsOut[blockIdx.x] = sumsIn[blockIdx.x];
}

template <typename Real>
__global__ void chains_rows_kernel(const ChainDesc* __restrict__ cd) {
    const ChainDesc& d = cd[blockIdx.x];
    gather_rows<Real>((const Real*)d.x, d.kdim, d.p, d.traceRows, d.rows);
}

template <typename Real>
__global__ __launch_bounds__(kGramThreads) void chains_gram_kernel(const ChainDesc* __restrict__ cd) {
    __shared__ double red[kGramThreads];
    const ChainDesc& d = cd[blockIdx.y];
    if ((int)blockIdx.x >= d.p * d.p) return;  // the same for the whole block
    gram_entry<Real>((const Real*)d.x, d.n, d.kdim, d.p, (int)blockIdx.x / d.p, (int)blockIdx.x % d.p, d.gram, red);
}

struct BatchHmc {
    struct Chain {
        int n = 0, p = 0;
        double tau = 0.0, eps = 0.0, U = 0.0, logLik = 0.0, quad = 0.0;
        std::vector<double> trait;
        mdsgen::Stream rng;
        double daMu = 0.0, daH = 0.0, daLogEpsBar = 0.0;
        long long accepted = 0, proposedTau = 0, acceptedTau = 0, evaluations = 1;
        size_t xBytes = 0, momOffset = 0, xSaveOffset = 0, locOffset = 0;
        bool proposing = false;
        double savedTau = 0.0;
        PrecisionWidth width;
        explicit Chain(uint64_t seed) : rng(seed, 77) {}
    };

    Batch& b;
    mds_hip_hmc_options opt;
    int C = 0, maxRows = 0, maxP = 0;
    std::vector<Chain> ch;
    double* dMom = nullptr;
    double* dGradSave = nullptr;
    char* dXSave = nullptr;
    double* dParams = nullptr;   // [C][kHmcParamCount]
    ChainDesc* dChains = nullptr;
    double* hParams = nullptr;   // pinned, copied to dParams
    // Written by the kernels themselves into pinned coherent memory and read
    // after a stream wait (k*: the same memory's device addresses):
    // energies [C][2] at the start and the end of a trajectory, each chain's
    // last sum [C], trace rows [C][kTraceRows * kMaxDim], X^T X [C][kMaxDim^2].
    double* hEnergy0 = nullptr;
    double* hEnergy = nullptr;
    double* hSums = nullptr;
    double* hRows = nullptr;
    double* hGram = nullptr;
    double* kEnergy0 = nullptr;
    double* kEnergy = nullptr;
    double* kSums = nullptr;
    double* kRows = nullptr;
    double* kGram = nullptr;
    int64_t maxWork = 0;         // threads per chain of the start and copy kernels
    hipEvent_t uploadDone = nullptr;
    bool uploadPending = false;
    bool paramsStale = false;    // a precision changed on the host since the last upload
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    BatchConfig savedCfg;

    BatchHmc(Batch& batch, const mds_hip_hmc_options& options) : b(batch), opt(options), savedCfg(batch.cfg) {
        if (opt.iterations < 1 || opt.leapfrogSteps < 1 || opt.burnIn < 0 || opt.burnIn >= opt.iterations)
            throw Error(MDS_HIP_ERR_ARGUMENT, "invalid iteration counts");
        if (!(opt.trajectory > 0.0)) throw Error(MDS_HIP_ERR_ARGUMENT, "trajectory must be > 0");
        if (opt.traitInvWeight < 1) throw Error(MDS_HIP_ERR_ARGUMENT, "traitInvWeight must be >= 1");
        if (opt.useSurrogate || opt.collectSnapshots || opt.surrogate)
            throw Error(MDS_HIP_ERR_ARGUMENT, "the batched sampler runs exact gradients only (no surrogate)");
        C = (int)b.e.size();
        size_t momCount = 0, xSaveBytes = 0, locCount = 0;
        for (int c = 0; c < C; ++c) {
            const BatchEntry& en = b.e[c];
            if (!en.dataSet) throw Error(MDS_HIP_ERR_STATE, "pairwise data not set for dataset " + std::to_string(c));
            if (!en.precisionSet) throw Error(MDS_HIP_ERR_STATE, "precision not set for dataset " + std::to_string(c));
            Chain k(opt.seed + (uint64_t)c);
            k.n = en.N;
            k.p = en.P;
            k.tau = en.tau;
            k.trait.assign((size_t)k.p * k.p, 0.0);
            for (int d = 0; d < k.p; ++d) k.trait[d * k.p + d] = 1.0;  // solve(traitVcv) with traitVcv = I_P
            k.eps = opt.stepSize > 0.0 ? opt.stepSize : opt.trajectory / opt.leapfrogSteps;
            k.daMu = std::log(10.0 * k.eps);
            k.width.w = opt.precisionProposalWidth;
            k.width.adapt = opt.adaptPrecisionWidth != 0;
            k.xBytes = (size_t)en.nb * kTile * en.kdim * b.realSize;
            k.momOffset = momCount;
            k.xSaveOffset = xSaveBytes;
            k.locOffset = locCount;
            momCount += (size_t)k.n * k.p;
            xSaveBytes += (k.xBytes + 255) / 256 * 256;
            locCount += (size_t)k.n * k.p;
            maxRows = std::max(maxRows, k.n);
            maxP = std::max(maxP, k.p);
            ch.push_back(std::move(k));
        }
        try {
            check(hipMalloc((void**)&dMom, momCount * sizeof(double)), "hipMalloc(momentum)");
            check(hipMalloc((void**)&dGradSave, momCount * sizeof(double)), "hipMalloc(gradient copies)");
            check(hipMalloc((void**)&dXSave, xSaveBytes), "hipMalloc(location copies)");
            check(hipMalloc((void**)&dParams, (size_t)C * kHmcParamCount * sizeof(double)), "hipMalloc(parameters)");
            check(hipMalloc((void**)&dChains, (size_t)C * sizeof(ChainDesc)), "hipMalloc(chains)");
            check(hipHostMalloc((void**)&hParams, (size_t)C * kHmcParamCount * sizeof(double), hipHostMallocDefault),
                  "hipHostMalloc");
            const auto pinned = [](double** h, double** k, size_t count) {
                check(hipHostMalloc((void**)h, count * sizeof(double), hipHostMallocCoherent), "hipHostMalloc");
                check(hipHostGetDevicePointer((void**)k, *h, 0), "hipHostGetDevicePointer");
            };
            pinned(&hEnergy0, &kEnergy0, (size_t)C * 2);
            pinned(&hEnergy, &kEnergy, (size_t)C * 2);
            pinned(&hSums, &kSums, (size_t)C);
            pinned(&hRows, &kRows, (size_t)C * kTraceRows * kMaxDim);
            pinned(&hGram, &kGram, (size_t)C * kMaxDim * kMaxDim);
            check(hipEventCreateWithFlags(&uploadDone, hipEventDisableTiming), "hipEventCreate");
            std::memset(hParams, 0, (size_t)C * kHmcParamCount * sizeof(double));
            std::vector<ChainDesc> desc(C);
            for (int c = 0; c < C; ++c) {
                ChainDesc& d = desc[c];
                d.x = b.e[c].dX;
                d.xSave = dXSave + ch[c].xSaveOffset;
                d.xCount = (int64_t)(ch[c].xBytes / b.realSize);
                d.mom = dMom + ch[c].momOffset;
                d.grad = b.e[c].dGrad;
                d.gradSave = dGradSave + ch[c].momOffset;
                d.params = dParams + (size_t)c * kHmcParamCount;
                d.rows = kRows + (size_t)c * kTraceRows * kMaxDim;
                d.gram = kGram + (size_t)c * kMaxDim * kMaxDim;
                // the momentum pairs are fewer than the n * p gradient values
                maxWork = std::max(maxWork, std::max<int64_t>(d.xCount, (int64_t)ch[c].n * ch[c].p));
                d.traceRows = trace_rows(ch[c].n);
                d.seed = opt.seed + (uint64_t)c;
                d.n = ch[c].n;
                d.p = ch[c].p;
                d.kdim = b.e[c].kdim;
            }
            check(hipMemcpy(dChains, desc.data(), (size_t)C * sizeof(ChainDesc), hipMemcpyHostToDevice),
                  "hipMemcpy(chains)");
            // one tile launch over every chain, on the batch's stream (no side
            // streams: the trajectory is captured there): the kernel of the
            // chains' one embedding-dimension class when they share it (its
            // registers are that class's alone), else the kernel for every
            // class; the precision read from each chain's parameter block
            bool oneClass = true;
            for (int c = 1; c < C; ++c) oneClass = oneClass && b.e[c].kdim == b.e[0].kdim;
            b.cfg.strategy = oneClass ? 1 : 0;
            for (int c = 0; c < C; ++c) {
                b.e[c].devTau = dParams + (size_t)c * kHmcParamCount + kHmcTau;
                b.e[c].kickMom = dMom + ch[c].momOffset;
                b.e[c].kickParams = dParams + (size_t)c * kHmcParamCount;
            }
            b.planDirty = true;
            b.buildPlan();
        } catch (...) {
            release();
            throw;
        }
    }

    ~BatchHmc() { release(); }

    void release() {
        if (b.stream) (void)hipStreamSynchronize(b.stream);
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
        if (graph) (void)hipGraphDestroy(graph);
        graphExec = nullptr;
        graph = nullptr;
        (void)hipFree(dMom);
        (void)hipFree(dGradSave);
        (void)hipFree(dXSave);
        (void)hipFree(dParams);
        (void)hipFree(dChains);
        (void)hipHostFree(hParams);
        (void)hipHostFree(hEnergy0);
        (void)hipHostFree(hEnergy);
        (void)hipHostFree(hSums);
        (void)hipHostFree(hRows);
        (void)hipHostFree(hGram);
        if (uploadDone) (void)hipEventDestroy(uploadDone);
        dMom = dGradSave = dParams = nullptr;
        dXSave = nullptr;
        dChains = nullptr;
        hParams = hEnergy0 = hEnergy = hSums = hRows = hGram = nullptr;
        kEnergy0 = kEnergy = kSums = kRows = kGram = nullptr;
        uploadDone = nullptr;
        // the batch evaluates with the host precisions again
        for (BatchEntry& en : b.e) {
            en.devTau = nullptr;
            en.kickMom = nullptr;
            en.kickParams = nullptr;
        }
        b.cfg = savedCfg;
        b.planDirty = true;
    }

    // The pinned block is rewritten only after the previous upload from it has run.
    void uploadParams() {
        if (uploadPending) check(hipEventSynchronize(uploadDone), "hipEventSynchronize(parameters)");
        paramsStale = false;
        for (int c = 0; c < C; ++c) {
            double* h = hParams + (size_t)c * kHmcParamCount;
            h[kHmcTau] = ch[c].tau;
            h[kHmcEps] = ch[c].eps;
            for (int k = 0; k < ch[c].p * ch[c].p; ++k) h[kHmcTrait + k] = ch[c].trait[k];
        }
        check(hipMemcpyAsync(dParams, hParams, (size_t)C * kHmcParamCount * sizeof(double), hipMemcpyHostToDevice,
                             b.stream),
              "hipMemcpy(parameters)");
        check(hipEventRecord(uploadDone, b.stream), "hipEventRecord(parameters)");
        uploadPending = true;
    }

    template <int FORM>
    void kick() {
        const dim3 grid((maxRows + kChainThreads - 1) / kChainThreads, C);
        if (b.fp32)
            hipLaunchKernelGGL((chains_kick_kernel<float, FORM>), grid, dim3(kChainThreads), 0, b.stream, dChains);
        else
            hipLaunchKernelGGL((chains_kick_kernel<double, FORM>), grid, dim3(kChainThreads), 0, b.stream, dChains);
        check(hipGetLastError(), "chains kick kernel");
    }

    // Every chain's energies into out (kEnergy0 or kEnergy); withSums: the
    // last pass's sums into kSums as well.
    void energies(double* out, bool withSums) {
        const double* in = withSums ? b.dSums : nullptr;
        if (b.fp32)
            hipLaunchKernelGGL(chains_energy_kernel<float>, dim3(C), dim3(kEnergyThreads), 0, b.stream, dChains, out,
                               in, kSums);
        else
            hipLaunchKernelGGL(chains_energy_kernel<double>, dim3(C), dim3(kEnergyThreads), 0, b.stream, dChains, out,
                               in, kSums);
        check(hipGetLastError(), "chains energy kernel");
    }

    void startChains(int it) {
        const dim3 grid((unsigned)((maxWork + kChainThreads - 1) / kChainThreads), C);
        if (b.fp32)
            hipLaunchKernelGGL(chains_start_kernel<float>, grid, dim3(kChainThreads), 0, b.stream, dChains,
                               (uint64_t)it);
        else
            hipLaunchKernelGGL(chains_start_kernel<double>, grid, dim3(kChainThreads), 0, b.stream, dChains,
                               (uint64_t)it);
        check(hipGetLastError(), "chains start kernel");
    }

    // Copies of the selected chains (what: 1 locations, 2 gradient, 3
    // both) to their saves, or back from them; one launch per kMaskChains
    // chains that has a selected one.
    void copyChains(const std::vector<char>& sel, int what, bool restore) {
        for (int c0 = 0; c0 < C; c0 += kMaskChains) {
            const int count = std::min(kMaskChains, C - c0);
            ChainMask mask{};
            bool any = false;
            for (int y = 0; y < count; ++y)
                if (sel[c0 + y]) {
                    mask.w[y >> 5] |= 1u << (y & 31);
                    any = true;
                }
            if (!any) continue;
            const dim3 grid((unsigned)((maxWork + kChainThreads - 1) / kChainThreads), count);
            if (b.fp32)
                hipLaunchKernelGGL(chains_copy_kernel<float>, grid, dim3(kChainThreads), 0, b.stream, dChains, mask,
                                   c0, what, restore ? 1 : 0);
            else
                hipLaunchKernelGGL(chains_copy_kernel<double>, grid, dim3(kChainThreads), 0, b.stream, dChains, mask,
                                   c0, what, restore ? 1 : 0);
            check(hipGetLastError(), "chains copy kernel");
        }
    }

    // Every chain's likelihood and gradient at its current locations.
    void evaluateAll() { b.enqueue(); }

    // The reduces kick every chain's rows right after writing them (the
    // kick_row arithmetic of the kick kernel: the same bits).
    void trajectory() {
        kick<kKickDrift>();
        for (int s = 0; s < opt.leapfrogSteps; ++s) b.enqueue(s + 1 < opt.leapfrogSteps ? kKick2Drift : kKick);
    }

    // The whole trajectory of every chain as one graph; every exit ends the capture.
    void buildGraph() {
        check(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        hipGraph_t g = nullptr;
        try {
            trajectory();
        } catch (...) {
            if (hipStreamEndCapture(b.stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            throw;
        }
        check(hipStreamEndCapture(b.stream, &g), "hipStreamEndCapture");
        graph = g;
        check(hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }

    double logPrior(const Chain& k, double quad) const { return matrix_normal_log_prior(quad, k.n, k.p, k.trait); }
    double logLik(const Chain& k, double rawSum, double t) const {
        return mds_log_likelihood(rawSum, t, k.n, b.truncated);
    }

    void readSums() {
        check(hipMemcpyAsync(hSums, b.dSums, (size_t)C * sizeof(double), hipMemcpyDeviceToHost, b.stream),
              "hipMemcpy(sums)");
    }


    void run(const double* x0, double* traces, double* finalX, mds_hip_hmc_stats* stats) {
        using clock = std::chrono::steady_clock;
        const auto tStart = clock::now();
        const int kept = opt.iterations - opt.burnIn;
        for (int c = 0; c < C; ++c) b.updateLocations(c, x0 + ch[c].locOffset, (size_t)ch[c].n * ch[c].p);
        uploadParams();
        evaluateAll();
        size_t momTotal = 0;
        for (const Chain& k : ch) momTotal += (size_t)k.n * k.p;
        check(hipMemsetAsync(dMom, 0, momTotal * sizeof(double), b.stream), "hipMemset(momentum)");
        energies(kEnergy, true);
        check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
        for (int c = 0; c < C; ++c) {
            Chain& k = ch[c];
            k.logLik = logLik(k, hSums[c], k.tau);
            k.quad = hEnergy[2 * c + 1];
            k.U = -(k.logLik + logPrior(k, k.quad));
        }
        if (opt.useGraph) buildGraph();
        const double daGamma = 0.05, daT0 = 10.0, daKappa = 0.75;
        auto tSampling = clock::now();
        std::vector<char> sel(C);
        for (int it = 0; it < opt.iterations; ++it) {
            if (it == opt.burnIn) {
                tSampling = clock::now();
                if (opt.adaptStepSize && opt.burnIn > 0) {
                    for (Chain& k : ch) k.eps = std::exp(k.daLogEpsBar);
                    uploadParams();
                }
            }
            // the momenta and the copies a rejection restores, then the
            // starting kinetic energies (read after the trajectory's wait)
            startChains(it);
            energies(kEnergy0, false);
            if (graphExec)
                check(hipGraphLaunch(graphExec, b.stream), "hipGraphLaunch");
            else
                trajectory();
            energies(kEnergy, true);
            check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
            bool adapted = false;
            std::fill(sel.begin(), sel.end(), 0);
            for (int c = 0; c < C; ++c) {
                Chain& k = ch[c];
                const double K0 = hEnergy0[2 * c], K1 = hEnergy[2 * c];
                const double logLik1 = logLik(k, hSums[c], k.tau);
                const double quad1 = hEnergy[2 * c + 1];
                const double U1 = -(logLik1 + logPrior(k, quad1));
                const double ratio = (k.U + K0) - (U1 + K1);  // R/Untitled.R:630-637
                const double u = k.rng.uniform();
                const bool acc = std::isfinite(ratio) && ratio > std::min(0.0, std::log(u));
                if (acc) {
                    k.U = U1;
                    k.logLik = logLik1;
                    k.quad = quad1;
                    ++k.accepted;
                } else {
                    sel[c] = 1;  // locations and gradient restored below
                }
                k.evaluations += opt.leapfrogSteps;
                if (opt.adaptStepSize && it < opt.burnIn) {
                    // dual averaging (Hoffman & Gelman 2014, sec. 3.2), per chain
                    const double alpha = std::isfinite(ratio) ? std::min(1.0, std::exp(std::min(ratio, 0.0))) : 0.0;
                    const double m = it + 1.0;
                    k.daH = (1.0 - 1.0 / (m + daT0)) * k.daH + (opt.targetAccept - alpha) / (m + daT0);
                    const double logEps = k.daMu - std::sqrt(m) / daGamma * k.daH;
                    const double eta = std::pow(m, -daKappa);
                    k.daLogEpsBar = eta * logEps + (1.0 - eta) * k.daLogEpsBar;
                    k.eps = std::exp(logEps);
                    adapted = true;
                }
            }
            copyChains(sel, 3, true);  // the rejected chains' locations and gradients
            if (adapted) uploadParams();

            // trace rows and X^T X of the kept locations (the precision step
            // does not move them), written into pinned memory and read after
            // the next wait
            const bool traitStep = opt.learnTraitPrecision && it < opt.iterations - 1 &&
                                   ((it + 1) % opt.traitInvWeight == 0);
            const bool recordIt = it >= opt.burnIn && traces;
            if (traitStep) {
                if (b.fp32)
                    hipLaunchKernelGGL(chains_gram_kernel<float>, dim3(maxP * maxP, C), dim3(kGramThreads), 0,
                                       b.stream, dChains);
                else
                    hipLaunchKernelGGL(chains_gram_kernel<double>, dim3(maxP * maxP, C), dim3(kGramThreads), 0,
                                       b.stream, dChains);
                check(hipGetLastError(), "chains gram kernel");
            }
            if (recordIt) {
                if (b.fp32)
                    hipLaunchKernelGGL(chains_rows_kernel<float>, dim3(C), dim3(kTraceRows * kMaxDim), 0, b.stream,
                                       dChains);
                else
                    hipLaunchKernelGGL(chains_rows_kernel<double>, dim3(C), dim3(kTraceRows * kMaxDim), 0, b.stream,
                                       dChains);
                check(hipGetLastError(), "chains rows kernel");
            }

            if (opt.learnPrecision) {
                if (b.truncated) {
                    // random-walk Metropolis on each chain's precision at its current locations
                    bool any = false;
                    for (int c = 0; c < C; ++c) {
                        Chain& k = ch[c];
                        const double lo = k.tau - k.width.w, hi = k.tau + k.width.w;
                        const double tauStar = std::fabs(lo + (hi - lo) * k.rng.uniform());
                        ++k.proposedTau;
                        k.proposing = tauStar > 0.0;
                        sel[c] = k.proposing ? 1 : 0;
                        if (k.proposing) {
                            k.savedTau = k.tau;
                            k.tau = tauStar;
                            any = true;
                        }
                    }
                    if (any) {
                        copyChains(sel, 2, false);  // the proposing chains' gradients
                        uploadParams();
                        evaluateAll();
                        readSums();
                        check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
                        bool restored = false;
                        std::fill(sel.begin(), sel.end(), 0);
                        for (int c = 0; c < C; ++c) {
                            Chain& k = ch[c];
                            if (!k.proposing) continue;
                            ++k.evaluations;
                            const double logLikStar = logLik(k, hSums[c], k.tau);
                            const double r = logLikStar - k.logLik;  // the prior does not involve tau
                            const bool accTau = std::isfinite(r) && r > std::min(0.0, std::log(k.rng.uniform()));
                            k.width.update(accTau, it < opt.burnIn);
                            if (accTau) {
                                // the pass at the new precision left its gradient in place
                                k.logLik = logLikStar;
                                k.U = -(k.logLik + logPrior(k, k.quad));
                                ++k.acceptedTau;
                                ++k.evaluations;
                            } else {
                                k.tau = k.savedTau;
                                sel[c] = 1;
                                restored = true;
                            }
                        }
                        // the rejected chains' gradients back; their precisions
                        // go with the iteration's last upload
                        if (restored) {
                            copyChains(sel, 2, true);
                            paramsStale = true;
                        }
                    }
                } else {
                    // Gibbs: tau | X ~ Gamma(1 + N(N-1)/4, C / 2) with C the unscaled
                    // sum of squared residuals (the raw sum at the current X)
                    evaluateAll();
                    readSums();
                    check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
                    std::vector<double> raw(hSums, hSums + C);
                    for (int c = 0; c < C; ++c) {
                        Chain& k = ch[c];
                        const double pairs = (double)k.n * (k.n - 1) / 2.0;
                        k.tau = gamma_draw(k.rng, 1.0 + 0.5 * pairs) / (0.5 * raw[c]);
                        ++k.proposedTau;
                        ++k.acceptedTau;
                        k.evaluations += 2;
                    }
                    uploadParams();
                    evaluateAll();
                    for (int c = 0; c < C; ++c) {
                        Chain& k = ch[c];
                        k.logLik = logLik(k, raw[c], k.tau);
                        k.U = -(k.logLik + logPrior(k, k.quad));
                    }
                }
            }

            if (traitStep || recordIt) check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
            if (traitStep) {
                for (int c = 0; c < C; ++c) {
                    Chain& k = ch[c];
                    const double* gram = hGram + (size_t)c * kMaxDim * kMaxDim;
                    std::vector<double> m((size_t)k.p * k.p, 0.0);  // T0 + X^T X, T0 = I_P
                    for (int d = 0; d < k.p; ++d) m[d * k.p + d] = 1.0;
                    for (int q = 0; q < k.p * k.p; ++q) m[q] += gram[q];
                    const double df = (opt.priorDf > 0 ? opt.priorDf : k.p) + k.n;
                    k.trait = wishart_draw(k.rng, df, spd_inverse(m, k.p), k.p);
                    k.quad = 0.0;  // sum_i x_i^T T x_i = sum_ab T_ab (X^T X)_ab
                    for (int q = 0; q < k.p * k.p; ++q) k.quad += k.trait[q] * gram[q];
                    k.U = -(k.logLik + logPrior(k, k.quad));
                }
                uploadParams();
            }
            if (paramsStale) uploadParams();
            if (recordIt)
                for (int c = 0; c < C; ++c) {
                    const Chain& k = ch[c];
                    double* row = traces + ((size_t)c * kept + (it - opt.burnIn)) * kHmcTraceColumns;
                    write_trace_row(row, hRows + (size_t)c * kTraceRows * kMaxDim, k.p, -k.U, k.tau, k.trait);
                }
        }
        check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
        const auto tEnd = clock::now();
        if (finalX) {
            for (int c = 0; c < C; ++c) {
                const Chain& k = ch[c];
                const int kd = b.e[c].kdim;
                std::vector<char> host((size_t)k.n * kd * b.realSize);
                check(hipMemcpy(host.data(), b.e[c].dX, host.size(), hipMemcpyDeviceToHost), "hipMemcpy(locations)");
                for (int i = 0; i < k.n; ++i)
                    for (int d = 0; d < k.p; ++d)
                        finalX[k.locOffset + (size_t)i * k.p + d] =
                            b.fp32 ? (double)((const float*)host.data())[(size_t)i * kd + d]
                                   : ((const double*)host.data())[(size_t)i * kd + d];
            }
        }
        for (int c = 0; c < C; ++c) b.e[c].tau = ch[c].tau;  // the batch keeps each chain's final precision
        if (stats)
            for (int c = 0; c < C; ++c) {
                const Chain& k = ch[c];
                mds_hip_hmc_stats& st = stats[c];
                std::memset(&st, 0, sizeof(st));
                st.acceptRate = (double)k.accepted / opt.iterations;
                st.precisionAcceptRate = k.proposedTau ? (double)k.acceptedTau / k.proposedTau : 0.0;
                st.samplingSeconds = std::chrono::duration<double>(tEnd - tSampling).count();
                st.totalSeconds = std::chrono::duration<double>(tEnd - tStart).count();
                st.evaluations = (double)k.evaluations;
                st.finalPrecision = k.tau;
                st.stepSize = k.eps;
                st.finalPotential = k.U;
                st.exactIterations = opt.iterations;
                st.precisionProposalWidth = k.width.w;
            }
    }
};

}  // namespace mdship

struct mds_hip_batch {
    mdship::Batch* b;
};

static thread_local std::string g_batch_error = "no error";

template <typename F>
static int bguard(mds_hip_batch_handle h, F&& f) {
    if (!h || !h->b) {
        g_batch_error = "null handle";
        return MDS_HIP_ERR_ARGUMENT;
    }
    try {
        (void)hipSetDevice(h->b->device);
        f(*h->b);
        return MDS_HIP_OK;
    } catch (const mdship::Error& err) {
        h->b->lastError = err.what();
        return err.code;
    } catch (const std::exception& err) {
        h->b->lastError = err.what();
        return MDS_HIP_ERR_DEVICE;
    }
}

extern "C" {

int mds_hip_batch_create(int count, const int* embeddingDimensions, const int* locationCounts, long flags,
                         int device, mds_hip_batch_handle* out) {
    if (!out || !embeddingDimensions || !locationCounts) {
        g_batch_error = "null argument";
        return MDS_HIP_ERR_ARGUMENT;
    }
    *out = nullptr;
    try {
        *out = new mds_hip_batch{new mdship::Batch(count, embeddingDimensions, locationCounts, flags, device)};
        return MDS_HIP_OK;
    } catch (const mdship::Error& err) {
        g_batch_error = err.what();
        return err.code;
    } catch (const std::exception& err) {
        g_batch_error = err.what();
        return MDS_HIP_ERR_DEVICE;
    }
}

int mds_hip_batch_destroy(mds_hip_batch_handle h) {
    if (!h) return MDS_HIP_ERR_ARGUMENT;
    delete h->b;
    delete h;
    return MDS_HIP_OK;
}

const char* mds_hip_batch_last_error(mds_hip_batch_handle h) {
    if (!h || !h->b) return g_batch_error.c_str();
    return h->b->lastError.c_str();
}

int mds_hip_batch_set_pairwise_data(mds_hip_batch_handle h, int index, const double* data, size_t length) {
    return bguard(h, [&](mdship::Batch& b) { b.setPairwiseData(index, data, length); });
}

int mds_hip_batch_update_locations(mds_hip_batch_handle h, int index, const double* locations, size_t length) {
    return bguard(h, [&](mdship::Batch& b) { b.updateLocations(index, locations, length); });
}

int mds_hip_batch_set_parameters(mds_hip_batch_handle h, int index, const double* data, size_t length) {
    return bguard(h, [&](mdship::Batch& b) { b.setParameters(index, data, length); });
}

int mds_hip_batch_evaluate(mds_hip_batch_handle h) {
    return bguard(h, [&](mdship::Batch& b) { b.evaluate(); });
}

int mds_hip_batch_get_sums_of_increments(mds_hip_batch_handle h, double* out, size_t length) {
    return bguard(h, [&](mdship::Batch& b) { b.getSums(out, length); });
}

int mds_hip_batch_get_gradient(mds_hip_batch_handle h, int index, double* out, size_t length) {
    return bguard(h, [&](mdship::Batch& b) { b.getGradient(index, out, length); });
}

int mds_hip_batch_autotune(mds_hip_batch_handle h, const char* cachePath, int force, double* bestMs) {
    return bguard(h, [&](mdship::Batch& b) {
        const double ms = b.autotune(cachePath, force);
        if (bestMs) *bestMs = ms;
    });
}

int mds_hip_batch_set_config(mds_hip_batch_handle h, int wavesPerBlock, int blocksPerCU, int tilesPerItem,
                             int strategy) {
    return bguard(h, [&](mdship::Batch& b) {
        if (wavesPerBlock < 1 || wavesPerBlock > 8 || blocksPerCU < 1 || blocksPerCU > 64 || tilesPerItem < 1 ||
            tilesPerItem > 64 || (strategy != 0 && strategy != 1))
            throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "invalid batch configuration");
        b.cfg = {wavesPerBlock, blocksPerCU, tilesPerItem, strategy};
        b.planDirty = true;
    });
}

int mds_hip_batch_get_config(mds_hip_batch_handle h, int* wavesPerBlock, int* blocksPerCU, int* tilesPerItem,
                             int* strategy) {
    return bguard(h, [&](mdship::Batch& b) {
        if (wavesPerBlock) *wavesPerBlock = b.cfg.wavesPerBlock;
        if (blocksPerCU) *blocksPerCU = b.cfg.blocksPerCU;
        if (tilesPerItem) *tilesPerItem = b.cfg.tilesPerItem;
        if (strategy) *strategy = b.cfg.strategy;
    });
}

int mds_hip_batch_share_pairwise_data(mds_hip_batch_handle h, int index, int sourceIndex) {
    return bguard(h, [&](mdship::Batch& b) { b.sharePairwiseData(index, sourceIndex); });
}

int mds_hip_batch_hmc_run(mds_hip_batch_handle h, const mds_hip_hmc_options* options, const double* initialLocations,
                          size_t length, double* traces, size_t traceLength, double* finalLocations,
                          size_t finalLength, mds_hip_hmc_stats* stats, int statsCount) {
    return bguard(h, [&](mdship::Batch& b) {
        size_t locations = 0;
        for (const mdship::BatchEntry& en : b.e) locations += (size_t)en.N * en.P;
        if (!options || !initialLocations || length != locations)
            throw mdship::Error(MDS_HIP_ERR_ARGUMENT,
                                "batch hmc: options and the concatenated N_k x P_k initial locations required");
        if (traces && traceLength != b.e.size() * (size_t)(options->iterations - options->burnIn) *
                                         MDS_HIP_HMC_TRACE_COLUMNS)
            throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "batch hmc: trace length must be count * (iterations - burnIn) "
                                                      "* columns");
        if (finalLocations && finalLength != locations)
            throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "batch hmc: final locations length must be sum N_k P_k");
        if (stats && statsCount != (int)b.e.size())
            throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "batch hmc: one stats entry per dataset");
        mdship::BatchHmc hmc(b, *options);
        hmc.run(initialLocations, traces, finalLocations, stats);
    });
}

int mds_hip_batch_set_timing(mds_hip_batch_handle h, int enable) {
    return bguard(h, [&](mdship::Batch& b) { b.timing = enable != 0; });
}

int mds_hip_batch_get_timing(mds_hip_batch_handle h, double* ms, long long* count) {
    return bguard(h, [&](mdship::Batch& b) { b.getTiming(ms, count); });
}

}  // extern "C"
