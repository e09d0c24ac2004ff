// The single-instance engine (state, caches and device buffers) behind the C
// ABI; shared by the C ABI in mds_engine.hip and the samplers in mds_hmc.hip.
//
// State and caching follow the reference CPU engine
// (src/NewMultiDimensionalScaling.hpp:134-302):
//   * updateLocations invalidates the cached sum of increments (:165);
//   * setParameters invalidates it only with truncation (:262-274), and the
//     non-truncated sum is returned as 0.5 * precision * cached (:198-202);
//   * storeState copies the locations, sum and precision (:205-217);
//     restoreState restores sum and precision, marks the sum known and swaps
//     the location buffers (:234-255);
//   * setPairwiseData does not touch the cached sum (:257-260);
//   * a single-location update (updateLocations(i, ...)) of a state whose
//     sum is known is followed by an O(N) update of the sum
//     (mds_row.hip), the path the reference sketches and disables
//     (:146-157, :563-616); repeated updates of the same row chain from the
//     same base, a second row forces a full recomputation (:150-153), and so
//     do data changes, makeDirty, and every kMaxRowUpdates-th consecutive
//     incremental update (bounding the accumulated rounding);
//   * every other evaluation is a full O(N^2) pass.
// In addition the engine keeps the device gradient of the last fused
// evaluation and reuses it while the state is unchanged (the reference
// recomputes it; the values are identical), and a gradient call fills the
// sum cache when the reference would have recomputed the same value.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "mds_hip.h"
#include "mds_host.hpp"
#include "mds_internal.hpp"

namespace mdship {

// An HMC kick applied by the reduce to each gradient row it writes (see
// ReduceLaunch): form -1 for none.
struct ReduceKick {
    int form = -1;
    double* mom = nullptr;
    const double* params = nullptr;
};

// Where one pass (Engine::evaluate) sends its results besides dGrad and dSum.
struct PassOut {
    double tau = 0.0;             // > 0: at this precision instead of the engine's (no devTau)
    double* sum = nullptr;        // the sum here instead of dSum, which keeps its value
    double* sumHost = nullptr;    // a second copy of the sum (pinned host memory)
    double* gradHost = nullptr;   // a second copy of the gradient rows (pinned host memory)
};

struct Engine {
    int P = 0, N = 0, kdim = 0, device = 0;
    long flags = 0;
    bool fp32 = false, truncated = false;
    int nb = 0, npad = 0, numTiles = 0;
    size_t realSize = 8;

    hipStream_t stream = nullptr;
    void* dDelta = nullptr;
    size_t deltaBytes = 0;
    bool symmetric = true;
    bool dataSet = false;
    void* dX[2] = {nullptr, nullptr};
    int cur = 0;
    // storeState keeps the stored locations in dX[1 - cur], lazily: it only
    // sets storePending (the stored state is still dX[cur] itself); a write
    // of every location then goes to the other buffer and flips cur (no
    // copy), and a write of part of dX[cur] copies it to dX[1 - cur] first.
    bool storePending = false;
    uint32_t* dTiles = nullptr;
    double* dTileSums = nullptr;
    void* dPart = nullptr;     // gradient partials in the engine's real type
    double* dGrad = nullptr;
    double* dSum = nullptr;    // dGrad + N * P: the sum follows the gradient, as in hGrad
    double* hGrad = nullptr;   // pinned, coherent (the reduce kernel can write it directly)
    double* kGrad = nullptr;   // its device address
    double* hSum = nullptr;    // pinned, coherent: the reduce kernel can write it directly
    double* kSum = nullptr;    // hSum's device address
    void* hXBuf[2] = {nullptr, nullptr};  // pinned staging for location uploads, used in turn
    hipEvent_t hXDone[2] = {nullptr, nullptr};  // the last upload from each has run
    bool hXBusy[2] = {false, false};
    int hXNext = 0;

    bool precisionSet = false;
    double precision = 0.0, storedPrecision = 0.0;
    double sumOfIncrements = 0.0, storedSumOfIncrements = 0.0;
    bool sumKnown = false;
    bool devSumFresh = false;  // dSum matches the current state
    bool gradKnown = false;    // dGrad matches the current state

    // Sum-then-gradient prediction.  Callers often ask for the gradient
    // right after the sum of the same state (the R Potential, BEAST's HMC),
    // which as two API calls would run a likelihood pass and then the fused
    // pass.  The engine keeps a running rate of gradient requests per
    // evaluated state (weight 1/8 per state); above kPredictRate the sum runs
    // the fused kernels and reads the gradient back with it.  The fused pass
    // costs ~1.3 likelihood passes, so predicting pays once more than ~0.3
    // of the states have their gradient asked for.  Both kernels compute the
    // sum in the same operations and order: the sum is bitwise the same
    // either way, only the work per call changes.
    enum class StateEval { None, SumOnly, Fused };
    static constexpr double kPredictRate = 0.3;
    double gradRequestRate = 0.0;
    StateEval stateEval = StateEval::None;  // how the current state's sum was evaluated
    bool gradRequested = false;             // the current state's gradient was asked for
    bool hostGradFresh = false;             // hGrad holds the current state's gradient
    // The gradient of the state storeState saved, when it was on the host
    // then: restoreState brings it back with the locations and precision, so
    // a gradient request after a rejected move runs no pass (the values are
    // those a recomputation gives).  New observations, a new precision or
    // makeDirty drop it.
    // Copy on write: storeState only marks the current buffer as shared with
    // the stored state; the next write of a gradient moves to the other
    // buffer first.
    double* hGradStored = nullptr;  // pinned, coherent
    double* kGradStored = nullptr;  // its device address
    bool storedGradFresh = false;
    bool storedSharesCurrent = false;  // the stored gradient is the one in hGrad

    void swap_host_gradients() {
        std::swap(hGrad, hGradStored);
        std::swap(kGrad, kGradStored);
    }

    // hGrad is about to be overwritten: keep a stored gradient that shares it
    void own_host_gradient() {
        if (storedSharesCurrent) {
            swap_host_gradients();
            storedSharesCurrent = false;
        }
    }

    // single-row updates: the known sum of the state before row pendingRow
    // moved is baseSum
    static constexpr int kMaxRowUpdates = 1024;
    bool sumFromCurrentData = false;  // the known sum was computed on the current observations
    // storeState saved a sum that was known and computed on the current
    // observations; new observations, makeDirty and a truncated precision
    // change clear it.  restoreState always restores the stored sum as the
    // known sum (src/NewMultiDimensionalScaling.hpp:234-236), but only a
    // valid one may serve as the base of a later O(N) single-row update.
    bool storedSumValid = false;
    int pendingRow = -1;
    double baseSum = 0.0;
    int rowUpdatesInARow = 0;
    double* hRowPartial = nullptr;  // pinned, coherent: the row kernel's per-workgroup partials
    unsigned int* dRowDone = nullptr;  // the row kernel's completion counter
    // A single-row updateLocations does not touch the device: the location
    // array keeps the row's previous value, which the row kernel reads
    // before its last workgroup writes the new one (no copies per move).
    // Any other use of the locations first writes the pending row
    // (flush_row_write), which also ends the O(N) path for it.
    bool rowWritePending = false;
    int rowWriteIndex = -1;
    double rowWriteValues[kMaxDim] = {};  // in the engine's real type, widened
    long long fullEvaluations = 0, rowUpdates = 0;

    bool timing = false;
    std::vector<hipEvent_t> eventPool;
    std::vector<hipEvent_t> pending;  // triples: start, after tiles, after reduce
    double tileMsTotal = 0.0, evalMsTotal = 0.0;
    long long timedCount = 0;

    std::string lastError;

    Engine(int p, int n, long f, int dev) : P(p), N(n), flags(f) {
        if (p < 1 || p > kMaxDim)
            throw Error(MDS_HIP_ERR_ARGUMENT,
                        "embedding dimension must be in [1, " + std::to_string(kMaxDim) + "]");
        if (n < 1) throw Error(MDS_HIP_ERR_ARGUMENT, "location count must be >= 1");
        int count = 0;
        check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (count <= 0) throw Error(MDS_HIP_ERR_DEVICE, "no HIP device");
        device = (dev < 0 || dev >= count) ? 0 : dev;
        check(hipSetDevice(device), "hipSetDevice");
        fp32 = (flags & MDS_HIP_FLAG_FLOAT) != 0;
        truncated = (flags & MDS_HIP_FLAG_LEFT_TRUNCATION) != 0;
        realSize = fp32 ? sizeof(float) : sizeof(double);
        kdim = kernel_dim(P);
        nb = (N + kTile - 1) / kTile;
        if (nb > 0xffff) throw Error(MDS_HIP_ERR_ARGUMENT, "location count too large");
        npad = nb * kTile;
        const std::vector<uint32_t> tiles = make_tiles(nb);
        numTiles = (int)tiles.size();
        try {
            check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
            const size_t xBytes = (size_t)npad * kdim * realSize;
            for (int k = 0; k < 2; ++k) {
                check(hipMalloc(&dX[k], xBytes), "hipMalloc(locations)");
                check(hipMemsetAsync(dX[k], 0, xBytes, stream), "hipMemset(locations)");
            }
            check(hipMalloc((void**)&dTiles, tiles.size() * sizeof(uint32_t)), "hipMalloc(tiles)");
            check(hipMemcpyAsync(dTiles, tiles.data(), tiles.size() * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, stream),
                  "hipMemcpy(tiles)");
            check(hipMalloc((void**)&dTileSums, tiles.size() * sizeof(double)), "hipMalloc(sums)");
            check(hipMalloc(&dPart, (size_t)nb * nb * kdim * kTile * realSize), "hipMalloc(partials)");
            check(hipMalloc((void**)&dGrad, ((size_t)N * P + 1) * sizeof(double)), "hipMalloc(gradient)");
            dSum = dGrad + (size_t)N * P;
            // the host copies hold the sum after the gradient as well
            check(hipHostMalloc((void**)&hGrad, ((size_t)N * P + 1) * sizeof(double), hipHostMallocCoherent),
                  "hipHostMalloc(gradient)");
            check(hipHostMalloc((void**)&hGradStored, ((size_t)N * P + 1) * sizeof(double), hipHostMallocCoherent),
                  "hipHostMalloc(stored gradient)");
            check(hipHostGetDevicePointer((void**)&kGrad, hGrad, 0), "hipHostGetDevicePointer(gradient)");
            check(hipHostGetDevicePointer((void**)&kGradStored, hGradStored, 0),
                  "hipHostGetDevicePointer(stored gradient)");
            check(hipHostMalloc((void**)&hSum, sizeof(double), hipHostMallocCoherent), "hipHostMalloc(sum)");
            check(hipHostGetDevicePointer((void**)&kSum, hSum, 0), "hipHostGetDevicePointer(sum)");
            for (int k = 0; k < 2; ++k) {
                check(hipHostMalloc(&hXBuf[k], (size_t)npad * kdim * realSize, hipHostMallocDefault),
                      "hipHostMalloc(locations)");
                check(hipEventCreateWithFlags(&hXDone[k], hipEventDisableTiming), "hipEventCreate(locations)");
                std::memset(hXBuf[k], 0, (size_t)npad * kdim * realSize);
            }
            // written by the row kernel itself (coherent pinned memory, no copy)
            check(hipHostMalloc((void**)&hRowPartial, row_delta_blocks(N) * sizeof(double), hipHostMallocCoherent),
                  "hipHostMalloc(row sums)");
            check(hipMalloc((void**)&dRowDone, sizeof(unsigned int)), "hipMalloc(row counter)");
            check(hipMemsetAsync(dRowDone, 0, sizeof(unsigned int), stream), "hipMemset(row counter)");
            check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        } catch (...) {
            release();
            throw;
        }
    }

    ~Engine() { release(); }

    void release() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t e : eventPool) (void)hipEventDestroy(e);
        for (hipEvent_t e : pending) (void)hipEventDestroy(e);
        eventPool.clear();
        pending.clear();
        (void)hipFree(dDelta);
        (void)hipFree(dX[0]);
        (void)hipFree(dX[1]);
        (void)hipFree(dTiles);
        (void)hipFree(dTileSums);
        (void)hipFree(dPart);
        (void)hipFree(dGrad);  // dSum lies within
        (void)hipHostFree(hGrad);
        (void)hipHostFree(hGradStored);
        (void)hipHostFree(hSum);
        for (int k = 0; k < 2; ++k) {
            (void)hipHostFree(hXBuf[k]);
            if (hXDone[k]) (void)hipEventDestroy(hXDone[k]);
            hXBuf[k] = nullptr;
            hXDone[k] = nullptr;
            hXBusy[k] = false;
        }
        (void)hipHostFree(hRowPartial);
        (void)hipFree(dRowDone);
        dRowDone = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        hRowPartial = nullptr;
        dDelta = dX[0] = dX[1] = nullptr;
        dTiles = nullptr;
        dTileSums = dGrad = dSum = hGrad = hGradStored = hSum = kSum = kGrad = kGradStored = nullptr;
        dPart = nullptr;
        stream = nullptr;
    }

    void invalidate_device_results() {
        if (stateEval != StateEval::None)
            gradRequestRate += 0.125 * ((gradRequested ? 1.0 : 0.0) - gradRequestRate);
        stateEval = StateEval::None;
        gradRequested = false;
        hostGradFresh = false;
        devSumFresh = false;
        gradKnown = false;
    }

    // src/NewMultiDimensionalScaling.hpp:134-166
    void updateLocations(int index, const double* loc, size_t length) {
        if (!loc) throw Error(MDS_HIP_ERR_ARGUMENT, "null location pointer");
        size_t rows, row0;
        if (index == -1) {
            if (length != (size_t)N * P)
                throw Error(MDS_HIP_ERR_ARGUMENT, "updateLocations(-1): length must be N*P");
            rows = N;
            row0 = 0;
        } else {
            if (index < 0 || index >= N)
                throw Error(MDS_HIP_ERR_ARGUMENT, "updateLocations: index out of range");
            if (length != (size_t)P)
                throw Error(MDS_HIP_ERR_ARGUMENT, "updateLocations(i): length must be P");
            rows = 1;
            row0 = index;
        }
        if (index >= 0) {
            if (rowWritePending && rowWriteIndex != index) flush_row_write();  // another row moved
            if (sumKnown && sumFromCurrentData) {
                // a single-row update from the known sum: the device row keeps the base value
                pendingRow = index;
                baseSum = sumOfIncrements;
            } else if (index != pendingRow) {
                pendingRow = -1;  // several rows moved: full recomputation (:150-153)
            }
            for (int k = 0; k < P; ++k) rowWriteValues[k] = fp32 ? (double)(float)loc[k] : loc[k];
            rowWritePending = true;
            rowWriteIndex = index;
        } else {
            // Stage as [N][kdim] in the engine's real type, padding with zeros,
            // in the staging buffer whose last upload has run (no wait for
            // this one: the next call stages in the other buffer).
            rowWritePending = false;  // every row is overwritten
            pendingRow = -1;
            const int k = hXNext;
            if (hXBusy[k]) check(hipEventSynchronize(hXDone[k]), "hipEventSynchronize(locations)");
            hXBusy[k] = false;
            if (fp32) {
                float* st = (float*)hXBuf[k];
                for (size_t r = 0; r < rows; ++r)
                    for (int c = 0; c < kdim; ++c) st[r * kdim + c] = c < P ? (float)loc[r * P + c] : 0.0f;
            } else {
                double* st = (double*)hXBuf[k];
                for (size_t r = 0; r < rows; ++r)
                    for (int c = 0; c < kdim; ++c) st[r * kdim + c] = c < P ? loc[r * P + c] : 0.0;
            }
            check(hipMemcpyAsync((char*)full_write_target() + row0 * kdim * realSize, hXBuf[k],
                                 rows * kdim * realSize, hipMemcpyHostToDevice, stream),
                  "hipMemcpy(locations)");
            check(hipEventRecord(hXDone[k], stream), "hipEventRecord(locations)");
            hXBusy[k] = true;
            hXNext = 1 - k;
        }
        sumKnown = false;
        invalidate_device_results();
    }

    // The stored state moves out of dX[cur] before a write of part of it.
    void resolve_store() {
        if (!storePending) return;
        check(hipMemcpyAsync(dX[1 - cur], dX[cur], (size_t)npad * kdim * realSize, hipMemcpyDeviceToDevice, stream),
              "hipMemcpy(store)");
        storePending = false;
    }

    // The buffer a write of every location goes to: with a store pending the
    // other one, so that dX[cur] stays as the stored state without a copy.
    void* full_write_target() {
        if (storePending) {
            cur = 1 - cur;
            storePending = false;
        }
        return dX[cur];
    }

    // Writes the pending single-row update into the location array; the
    // O(N) path for it ends (its previous location is gone).
    void flush_row_write() {
        if (!rowWritePending) return;
        resolve_store();
        RowWrite w;
        w.x = dX[cur];
        for (int k = 0; k < P; ++k) w.values[k] = rowWriteValues[k];
        w.row = rowWriteIndex;
        w.kdim = kdim;
        w.p = P;
        w.fp32 = fp32;
        check(launch_row_write(w, stream), "row write kernel");
        rowWritePending = false;
        pendingRow = -1;
    }

    // The raw sum after the pending single-row update: base + half the
    // row's change, the block partials added in order.
    void applyRowUpdate() {
        resolve_store();
        RowDeltaLaunch a;
        a.delta = dDelta;
        a.x = dX[cur];
        for (int k = 0; k < P; ++k) a.xNew[k] = rowWriteValues[k];
        a.row = pendingRow;
        a.n = N;
        a.nb = nb;
        a.kdim = kdim;
        a.p = P;
        a.done = dRowDone;
        a.planeStride = (int64_t)numTiles * kTileElems;
        a.tau = precision;
        a.truncated = truncated;
        a.symmetric = symmetric;
        a.fp32 = fp32;
        a.partial = hRowPartial;
        check(launch_row_delta(a, stream), "row update kernel");
        rowWritePending = false;  // the kernel's last workgroup writes the row
        const int blocks = row_delta_blocks(N);
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        double change = 0.0;
        for (int b = 0; b < blocks; ++b) change += hRowPartial[b];
        sumOfIncrements = baseSum + 0.5 * change;
        ++rowUpdates;
        ++rowUpdatesInARow;
    }

    // A full evaluation's sum became the known sum.
    void sumFromFullEvaluation() {
        sumFromCurrentData = true;
        pendingRow = -1;
        rowUpdatesInARow = 0;
    }

    void setPairwiseData(const double* data, size_t length) {
        if (!data) throw Error(MDS_HIP_ERR_ARGUMENT, "null data pointer");
        if (length != (size_t)N * N)
            throw Error(MDS_HIP_ERR_ARGUMENT, "setPairwiseData: length must be N*N");
        const std::vector<uint32_t> tiles = make_tiles(nb);
        const size_t planeElems = (size_t)numTiles * kTileElems;
        std::unique_ptr<char[]> host(new char[2 * planeElems * realSize]);
        bool sym;
        if (fp32)
            sym = pack_tiles<float>(data, N, tiles, (float*)host.get(),
                                    (float*)host.get() + planeElems);
        else
            sym = pack_tiles<double>(data, N, tiles, (double*)host.get(),
                                     (double*)host.get() + planeElems);
        const size_t bytes = (sym ? 1 : 2) * planeElems * realSize;
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        if (bytes != deltaBytes) {
            check(hipFree(dDelta), "hipFree(observations)");
            dDelta = nullptr;
            deltaBytes = 0;
            check(hipMalloc(&dDelta, bytes), "hipMalloc(observations)");
            deltaBytes = bytes;
        }
        check(hipMemcpy(dDelta, host.get(), bytes, hipMemcpyHostToDevice), "hipMemcpy(observations)");
        symmetric = sym;
        dataSet = true;
        sumFromCurrentData = false;  // the cached sum stays (:257-260) but is no base for row updates
        storedSumValid = false;
        pendingRow = -1;
        storedGradFresh = false;
        storedSharesCurrent = false;
        invalidate_device_results();
    }

    // src/NewMultiDimensionalScaling.hpp:262-274
    void setParameters(const double* data, size_t length) {
        if (!data || length != 1)
            throw Error(MDS_HIP_ERR_ARGUMENT, "setParameters: expects exactly the precision");
        if (!std::isfinite(data[0]) || data[0] <= 0.0)
            throw Error(MDS_HIP_ERR_ARGUMENT, "setParameters: precision must be finite and > 0");
        precision = data[0];
        precisionSet = true;
        if (truncated) {
            sumKnown = false;
            pendingRow = -1;  // the base sum was for another precision
            storedSumValid = false;
        }
        storedGradFresh = false;
        storedSharesCurrent = false;
        invalidate_device_results();
    }

    void makeDirty() {  // :276-279
        sumKnown = false;
        pendingRow = -1;
        storedSumValid = false;
        storedGradFresh = false;
        storedSharesCurrent = false;
        invalidate_device_results();
    }

    void storeState() {  // :205-217
        flush_row_write();
        storedSumOfIncrements = sumOfIncrements;
        storedSumValid = sumKnown && sumFromCurrentData;
        storedGradFresh = hostGradFresh;
        storedSharesCurrent = hostGradFresh;
        storePending = true;  // the copy is made only if dX[cur] is then written in part
        storedPrecision = precision;
    }

    void restoreState() {  // :234-255
        // a pending single-row write belongs to the state given up: dropped
        rowWritePending = false;
        sumOfIncrements = storedSumOfIncrements;
        sumKnown = true;
        sumFromCurrentData = storedSumValid;
        pendingRow = -1;
        precision = storedPrecision;
        if (!storePending) cur = 1 - cur;  // else nothing was written since the store: dX[cur] is it
        // the current locations are the stored ones again, and stay stored
        // (a second restore keeps them, as the reference's copy does)
        storePending = true;
        invalidate_device_results();
        if (storedGradFresh) {
            if (!storedSharesCurrent) swap_host_gradients();
            hostGradFresh = true;
        }
        storedGradFresh = false;
        storedSharesCurrent = false;
    }

    hipEvent_t take_event() {
        if (eventPool.empty()) {
            hipEvent_t e;
            check(hipEventCreate(&e), "hipEventCreate");
            return e;
        }
        hipEvent_t e = eventPool.back();
        eventPool.pop_back();
        return e;
    }

    // The tile kernel of one evaluation at the current device locations into
    // the given tile-sum and partial buffers.  devTau: when non-null, the
    // kernels read the precision from this device address instead of the
    // engine's value (for captured graphs whose precision changes between
    // replays); otherwise they take tau.
    void launch_tile_pass(bool gradient, const double* devTau, double* tileSums, void* part, double tau) {
        TileLaunch a;
        a.delta = dDelta;
        a.x = dX[cur];
        a.tiles = dTiles;
        a.numTiles = numTiles;
        a.nb = nb;
        a.kdim = kdim;
        a.planeStride = (int64_t)numTiles * kTileElems;
        a.tau = tau;
        a.truncated = truncated;
        a.symmetric = symmetric;
        a.gradient = gradient;
        a.tileSums = tileSums;
        a.part = part;
        a.devTau = devTau;
        check(fp32 ? launch_tiles_f32(a, stream) : launch_tiles_f64(a, stream), "tile kernel");
    }

    // The reduce kernel: the gradient rows into grad (when gradient) and
    // (0.5 * the sum of the tile sums) * sumScale into *sumOut; with a kick,
    // each row then moves its momentum (and location) at the current
    // locations.
    void launch_reduce_pass(bool gradient, const double* tileSums, const void* part, double* grad, double* sumOut,
                            double sumScale, const ReduceKick& kick = ReduceKick(), double* sumHost = nullptr,
                            double* gradHost = nullptr) {
        ReduceLaunch r;
        r.part = part;
        r.fp32 = fp32;
        r.tileSums = tileSums;
        r.numTiles = numTiles;
        r.nb = nb;
        r.kdim = kdim;
        r.n = N;
        r.p = P;
        r.gradient = gradient;
        r.grad = grad;
        r.gradHost = gradHost;
        r.sumOut = sumOut;
        r.sumHost = sumHost;
        r.sumScale = sumScale;
        r.kickForm = kick.form;
        r.kickX = dX[cur];
        r.kickMom = kick.mom;
        r.kickParams = kick.params;
        check(launch_reduce(r, stream), "reduce kernel");
    }

    // One fused pass at the current locations (out: see PassOut).
    void evaluate(bool gradient, const double* devTau = nullptr, const ReduceKick& kick = ReduceKick(),
                  const PassOut& out = PassOut()) {
        if (!dataSet) throw Error(MDS_HIP_ERR_STATE, "pairwise data not set");
        if (!precisionSet) throw Error(MDS_HIP_ERR_STATE, "precision not set");
        flush_row_write();
        hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
        if (timing) {
            e0 = take_event();
            e1 = take_event();
            e2 = take_event();
            check(hipEventRecord(e0, stream), "hipEventRecord");
        }
        launch_tile_pass(gradient, devTau, dTileSums, dPart, out.tau > 0.0 ? out.tau : precision);
        if (timing) check(hipEventRecord(e1, stream), "hipEventRecord");
        launch_reduce_pass(gradient, dTileSums, dPart, dGrad, out.sum ? out.sum : dSum, 1.0, kick, out.sumHost,
                           out.gradHost);
        if (timing) {
            check(hipEventRecord(e2, stream), "hipEventRecord");
            pending.push_back(e0);
            pending.push_back(e1);
            pending.push_back(e2);
        }
        const bool ownTau = out.tau <= 0.0;
        if (!out.sum) devSumFresh = ownTau;  // dSum rewritten
        if (gradient) gradKnown = ownTau;    // dGrad rewritten
        hostGradFresh = false;
        ++fullEvaluations;
    }

    double getSumOfIncrements() {  // :193-203
        if (!sumKnown) {
            if (pendingRow >= 0 && !devSumFresh && rowUpdatesInARow < kMaxRowUpdates && rowWritePending &&
                rowWriteIndex == pendingRow) {
                applyRowUpdate();
                pendingRow = -1;
            } else {
                bool gradOnHost = false, sumOnHost = false;
                if (!devSumFresh) {
                    const bool predictGradient = gradRequestRate > kPredictRate;
                    // the reduce kernel writes the sum (into hSum), or with a
                    // predicted gradient the gradient and the sum after it
                    // (into hGrad), itself: no copy
                    PassOut out;
                    if (predictGradient) {
                        own_host_gradient();
                        out.gradHost = kGrad;
                        out.sumHost = kGrad + (size_t)N * P;
                        gradOnHost = true;
                    } else {
                        out.sumHost = kSum;
                        sumOnHost = true;
                    }
                    evaluate(predictGradient, nullptr, ReduceKick(), out);
                    stateEval = predictGradient ? StateEval::Fused : StateEval::SumOnly;
                }
                if (!gradOnHost && !sumOnHost)
                    check(hipMemcpyAsync(hSum, dSum, sizeof(double), hipMemcpyDeviceToHost, stream),
                          "hipMemcpy(sum)");
                check(hipStreamSynchronize(stream), "hipStreamSynchronize");
                sumOfIncrements = gradOnHost ? hGrad[(size_t)N * P] : *hSum;
                sumFromFullEvaluation();
                if (gradOnHost) hostGradFresh = true;
            }
            sumKnown = true;
        }
        return truncated ? sumOfIncrements : 0.5 * precision * sumOfIncrements;
    }

    void getLogLikelihoodGradient(double* result, size_t length) {  // :281-302
        if (!result) throw Error(MDS_HIP_ERR_ARGUMENT, "null gradient pointer");
        if (length != (size_t)N * P)
            throw Error(MDS_HIP_ERR_ARGUMENT, "getLogLikelihoodGradient: length must be N*P");
        gradRequested = true;
        if (hostGradFresh) {  // read back with the sum, or restored with the state
            std::memcpy(result, hGrad, length * sizeof(double));
            return;
        }
        const bool fillSum = !sumKnown;
        own_host_gradient();
        // the gradient, and the sum stored after it when it is wanted too:
        // written into hGrad by the pass itself, or copied from the device
        // when the pass ran before
        if (!gradKnown) {
            PassOut out;
            out.gradHost = kGrad;
            if (fillSum) out.sumHost = kGrad + length;
            evaluate(true, nullptr, ReduceKick(), out);
            if (stateEval == StateEval::None) stateEval = StateEval::Fused;
        } else {
            check(hipMemcpyAsync(hGrad, dGrad, (length + (fillSum ? 1 : 0)) * sizeof(double),
                                 hipMemcpyDeviceToHost, stream),
                  "hipMemcpy(gradient)");
        }
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        std::memcpy(result, hGrad, length * sizeof(double));
        hostGradFresh = true;
        if (fillSum) {
            sumOfIncrements = hGrad[length];
            sumKnown = true;
            sumFromFullEvaluation();
        }
    }

    void getTiming(double* tileMs, double* evalMs, long long* count) {
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        for (size_t k = 0; k + 2 < pending.size(); k += 3) {
            float a = 0.f, b = 0.f;
            check(hipEventElapsedTime(&a, pending[k], pending[k + 1]), "hipEventElapsedTime");
            check(hipEventElapsedTime(&b, pending[k], pending[k + 2]), "hipEventElapsedTime");
            tileMsTotal += a;
            evalMsTotal += b;
            ++timedCount;
        }
        for (hipEvent_t e : pending) eventPool.push_back(e);
        pending.clear();
        if (tileMs) *tileMs = tileMsTotal;
        if (evalMs) *evalMs = evalMsTotal;
        if (count) *count = timedCount;
    }
};

}  // namespace mdship

struct mds_hip_engine {
    mdship::Engine* e;
};
