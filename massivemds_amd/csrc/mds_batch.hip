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
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "mds_hip.h"
#include "mds_host.hpp"
#include "mds_internal.hpp"
#include "mds_tile.hpp"
#include "mds_tune_cache.hpp"

namespace mdship {

struct BatchDataset {
    const void* delta;      // packed tiles (Real), planes as in the engine
    const void* x;          // [nb * 64][kdim] Real
    const uint32_t* tiles;  // tile codes (I << 16) | J
    double* part;           // [nb][nb][kdim][64]
    double* grad;           // [n][p]
    int64_t planeStride;
    double tau;
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

template <typename Real, int P>
__device__ __forceinline__ void batch_tile(const BatchDataset& d, uint32_t t, double* __restrict__ tileSums,
                                           const Real* lnPhiTab, Real* stage) {
    const uint32_t code = d.tiles[t];
    const int I = (int)(code >> 16), J = (int)(code & 0xffffu);
    const Real* base = (const Real*)d.delta + (int64_t)t * kTileElems;
    const Real* x = (const Real*)d.x;
    const PairConst<Real> pc = make_pair_const<Real>(d.tau, lnPhiTab);
    double* sum = tileSums + d.tileSumBase + t;
    // the previous tile's reads of the staged columns precede the new writes
    wave_lds_sync();
    stage_columns<Real, P>(stage, x, J);
    wave_lds_sync();
    if (d.truncated) {
        if (d.symmetric)
            tile_walk<Real, P, true, true, true, kGroup>(base, d.planeStride, x, stage, I, J, d.nb, pc, sum, d.part);
        else
            tile_walk<Real, P, true, false, true, kGroup>(base, d.planeStride, x, stage, I, J, d.nb, pc, sum,
                                                          d.part);
    } else {
        if (d.symmetric)
            tile_walk<Real, P, false, true, true, kGroup>(base, d.planeStride, x, stage, I, J, d.nb, pc, sum,
                                                          d.part);
        else
            tile_walk<Real, P, false, false, true, kGroup>(base, d.planeStride, x, stage, I, J, d.nb, pc, sum,
                                                           d.part);
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
template <typename Real, int KDIM>
__global__ __launch_bounds__(kBatchMaxThreads) void batch_tile_kernel(const BatchDataset* __restrict__ ds,
                                                                       const BatchItem* __restrict__ items,
                                                                       int numItems,
                                                                       double* __restrict__ tileSums) {
    __shared__ __align__(16) Real lnPhiTab[TableOf<Real>::kCount];  // truncated datasets read it
    __shared__ __align__(16) Real colStage[kBatchMaxThreads / kWave][ColumnStage<Real, 1>::kMaxElems];
    copy_lnphi_table(lnPhiTab);
    __syncthreads();
    Real* stage = colStage[threadIdx.x / kWave];
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
                batch_tile<Real, KDIM>(d, it.tile + c, tileSums, lnPhiTab, stage);
        }
    }
}

constexpr int kBatchReduceWaves = 8;

// Gradient rows of row block K of one dataset: each of the 8 wavefronts adds
// the column blocks J = w mod 8 for every coordinate at once (the loads of
// kAhead column blocks of all KD coordinates in flight together), then the
// wavefront partials are combined per coordinate in wavefront order.  Per
// coordinate the additions are those of the single-instance reduce, in the
// same order, so the rows are bitwise the same.
template <int KD>
__device__ __forceinline__ void batch_reduce_rows(const BatchDataset& d, int K, double* red) {
    constexpr int kAhead = 4;
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    double acc[KD];
#pragma unroll
    for (int p = 0; p < KD; ++p) acc[p] = 0.0;
    const int64_t stride = (int64_t)KD * kTile;  // between column blocks J
    const double* base = d.part + (int64_t)K * d.nb * stride + lane;
    int J = wave;
    for (; J + (kAhead - 1) * kBatchReduceWaves < d.nb; J += kAhead * kBatchReduceWaves) {
        double v[kAhead][KD];
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            const double* slot = base + (J + u * kBatchReduceWaves) * stride;
#pragma unroll
            for (int p = 0; p < KD; ++p) v[u][p] = slot[p * kTile];
        }
#pragma unroll
        for (int u = 0; u < kAhead; ++u)
#pragma unroll
            for (int p = 0; p < KD; ++p) acc[p] += v[u][p];
    }
    for (; J < d.nb; J += kBatchReduceWaves) {
        const double* slot = base + J * stride;
#pragma unroll
        for (int p = 0; p < KD; ++p) acc[p] += slot[p * kTile];
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
}

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
        case 1: batch_reduce_rows<1>(d, item.K, red); break;
        case 2: batch_reduce_rows<2>(d, item.K, red); break;
        case 3: batch_reduce_rows<3>(d, item.K, red); break;
        case 4: batch_reduce_rows<4>(d, item.K, red); break;
        case 5: batch_reduce_rows<5>(d, item.K, red); break;
        case 6: batch_reduce_rows<6>(d, item.K, red); break;
        case 8: batch_reduce_rows<8>(d, item.K, red); break;
        case 10: batch_reduce_rows<10>(d, item.K, red); break;
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
    void* dDelta = nullptr;
    size_t deltaBytes = 0;
    void* dX = nullptr;
    double* dPart = nullptr;
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
    // joined before the reduce, so one class's tail overlaps the next
    // class's work (the results do not depend on the order: every tile
    // writes its own slots).
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
    // launch plan: (kdim class or 0, first item, item count)
    struct Launch {
        int kdim, first, count;
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
                check(hipMalloc((void**)&b.dPart, (size_t)b.nb * b.nb * b.kdim * kTile * sizeof(double)),
                      "hipMalloc(partials)");
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
            b.dDelta = b.dX = nullptr;
            b.dPart = b.dGrad = nullptr;
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

    void setPairwiseData(int index, const double* data, size_t length) {
        BatchEntry& b = at(index);
        if (!data || length != (size_t)b.N * b.N)
            throw Error(MDS_HIP_ERR_ARGUMENT, "setPairwiseData: length must be N*N");
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
            BatchDataset& d = desc[k];
            d.delta = b.dDelta;
            d.x = b.dX;
            d.tiles = b.dTiles;
            d.part = b.dPart;
            d.grad = b.dGrad;
            d.planeStride = (int64_t)b.numTiles * kTileElems;
            d.tau = b.tau;
            d.nb = b.nb;
            d.n = b.N;
            d.p = b.P;
            d.kdim = b.kdim;
            d.truncated = truncated ? 1 : 0;
            d.symmetric = b.symmetric ? 1 : 0;
            d.tileSumBase = b.tileSumBase;
            d.numTiles = b.numTiles;
        }
        // items: per class (kdim, heaviest first), datasets interleaved run by run
        std::map<int, std::vector<int>, std::greater<int>> classes;
        for (size_t k = 0; k < e.size(); ++k) classes[e[k].kdim].push_back((int)k);
        std::vector<BatchItem> items;
        plan.clear();
        const int tpi = std::max(1, cfg.tilesPerItem);
        for (auto& cls : classes) {
            const int first = (int)items.size();
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
            plan.push_back({cls.first, first, (int)items.size() - first});
        }
        if (cfg.strategy == 0) plan = {{0, 0, (int)items.size()}};
        std::vector<ReduceItem> red;
        for (size_t k = 0; k < e.size(); ++k) {
            red.push_back({(int)k, -1});
            for (int K = 0; K < e[k].nb; ++K) red.push_back({(int)k, K});
        }
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

    template <typename Real>
    void launchClass(int kdim, const BatchItem* items, int count, hipStream_t stream) {
        const int wpb = cfg.wavesPerBlock;
        const int want = (count + wpb - 1) / wpb;
        const int grid = std::max(1, std::min(want, cuCount * cfg.blocksPerCU));
        const dim3 g(grid), b(wpb * kWave);
        switch (kdim) {
            case 0: hipLaunchKernelGGL((batch_tile_kernel<Real, 0>), g, b, 0, stream, dDesc, items, count, dTileSums); break;
            case 1: hipLaunchKernelGGL((batch_tile_kernel<Real, 1>), g, b, 0, stream, dDesc, items, count, dTileSums); break;
            case 2: hipLaunchKernelGGL((batch_tile_kernel<Real, 2>), g, b, 0, stream, dDesc, items, count, dTileSums); break;
            case 3: hipLaunchKernelGGL((batch_tile_kernel<Real, 3>), g, b, 0, stream, dDesc, items, count, dTileSums); break;
            case 4: hipLaunchKernelGGL((batch_tile_kernel<Real, 4>), g, b, 0, stream, dDesc, items, count, dTileSums); break;
            case 5: hipLaunchKernelGGL((batch_tile_kernel<Real, 5>), g, b, 0, stream, dDesc, items, count, dTileSums); break;
            case 6: hipLaunchKernelGGL((batch_tile_kernel<Real, 6>), g, b, 0, stream, dDesc, items, count, dTileSums); break;
            case 8: hipLaunchKernelGGL((batch_tile_kernel<Real, 8>), g, b, 0, stream, dDesc, items, count, dTileSums); break;
            case 10: hipLaunchKernelGGL((batch_tile_kernel<Real, 10>), g, b, 0, stream, dDesc, items, count, dTileSums); break;
            default: throw Error(MDS_HIP_ERR_ARGUMENT, "unsupported kernel dimension");
        }
        check(hipGetLastError(), "batch tile kernel");
    }

    void enqueue() {
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
                launchClass<float>(l.kdim, dItems + l.first, l.count, s);
            else
                launchClass<double>(l.kdim, dItems + l.first, l.count, s);
        }
        for (int k = 0; k < used; ++k) {
            check(hipEventRecord(joinEvent[k], side[k]), "hipEventRecord(join)");
            check(hipStreamWaitEvent(stream, joinEvent[k], 0), "hipStreamWaitEvent(join)");
        }
        hipLaunchKernelGGL(batch_reduce_kernel, dim3(numReduce), dim3(kBatchReduceWaves * kWave), 0, stream, dDesc,
                           dReduce, dTileSums, dSums);
        check(hipGetLastError(), "batch reduce kernel");
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

int mds_hip_batch_set_timing(mds_hip_batch_handle h, int enable) {
    return bguard(h, [&](mdship::Batch& b) { b.timing = enable != 0; });
}

int mds_hip_batch_get_timing(mds_hip_batch_handle h, double* ms, long long* count) {
    return bguard(h, [&](mdship::Batch& b) { b.getTiming(ms, count); });
}

}  // extern "C"
