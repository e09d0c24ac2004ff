// Allocation-free evaluation (mds_hip_evaluate_device_noalloc): the
// evaluation of mds_hip_evaluate_async, at device-resident locations, into
// caller-owned device outputs, with the tile sums and gradient partials in a
// static device array of this file instead of the instance's own buffers.
// For callers that may not allocate device memory at evaluation time (tight
// serving loops, code that forbids hipMalloc): nothing here calls hipMalloc,
// hipFree or an allocating library routine.
//
// The array lives in the library's code object; every GPU the library runs
// on holds its own copy until the process ends, at its own address, which
// hipGetSymbolAddress gives once per device with that device current.  Host
// code never passes the array's name to a kernel (on the host it is only a
// stand-in for the symbol), always the resolved address.
//
// Every call on one GPU shares that GPU's array and leaves the last call's
// contents in it.  A call returns after its kernels have finished, so calls
// made one after another never overlap; calls made at the same time from
// different threads race and must be serialised by the caller.  Nothing in
// the array needs clearing: each evaluation writes every tile sum and every
// gradient slot it reads.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "mds_engine.hpp"

namespace mdship {

// 64 MiB: the partials of P <= 2 instances up to N = 16,320 (FP64 slots
// [nb][nb][2][64]; the FP32 engine's slots take half), or of any P up to the
// N whose slots fit; the likelihood alone needs 8 bytes per tile.
constexpr size_t kScratchBytes = (size_t)64 << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_address[kMaxScratchDevices] = {};

// The device address of `device`'s copy of the array; `device` is current.
void* scratch_address(int device) {
    if (device < 0 || device >= kMaxScratchDevices)
        throw Error(MDS_HIP_ERR_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_address[device]) {
        void* ptr = nullptr;
        check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) throw Error(MDS_HIP_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_address[device] = ptr;
    }
    return g_scratch_address[device];
}

// Refuses a pointer unless [ptr, ptr + bytes) lies inside one device
// allocation of `device` (a host pointer given here would otherwise reach a
// kernel).
void require_device_range(const void* ptr, size_t bytes, int device, const char* what) {
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess ||
        (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged) || attr.device != device) {
        (void)hipGetLastError();
        throw Error(MDS_HIP_ERR_ARGUMENT, std::string(what) + " is not device memory of the instance's device");
    }
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)ptr) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(MDS_HIP_ERR_ARGUMENT, std::string(what) + ": allocation not found");
    }
    if ((const char*)ptr < (const char*)base || (size_t)((const char*)ptr - (const char*)base) + bytes > size)
        throw Error(MDS_HIP_ERR_ARGUMENT, std::string(what) + " is shorter than " + std::to_string(bytes) + " bytes");
}

constexpr int kPackThreads = 256;

// Caller locations (N x P row-major FP64) into the engine's layout
// ([npad][kdim] in its real type, padding coordinates 0; padding rows are
// never written and stay 0).  The conversion to float rounds to nearest, as
// the host staging of updateLocations does.
template <typename Real>
__global__ __launch_bounds__(kPackThreads) void pack_locations_kernel(const double* __restrict__ src,
                                                                      Real* __restrict__ dst, int n, int p, int kdim) {
    const int64_t t = (int64_t)blockIdx.x * kPackThreads + threadIdx.x;
    if (t >= (int64_t)n * kdim) return;
    const int64_t i = t / kdim;
    const int k = (int)(t - i * kdim);
    dst[t] = k < p ? (Real)src[i * p + k] : (Real)0;
}

void evaluate_noalloc(Engine& e, const double* locations, double* sumOut, double* gradientOut) {
    if (!sumOut) throw Error(MDS_HIP_ERR_ARGUMENT, "null sum output");
    const bool gradient = gradientOut != nullptr;
    const size_t tileBytes = (size_t)e.numTiles * sizeof(double);
    const size_t partOffset = (tileBytes + 255) / 256 * 256;
    const size_t partBytes = (size_t)e.nb * e.nb * e.kdim * kTile * e.realSize;
    const size_t need = gradient ? partOffset + partBytes : tileBytes;
    if (need > kScratchBytes)
        throw Error(MDS_HIP_ERR_ARGUMENT, "evaluate_device_noalloc: this instance's partials need " +
                                              std::to_string(need) + " bytes of scratch, more than the " +
                                              std::to_string(kScratchBytes) + " available");
    if (!e.dataSet) throw Error(MDS_HIP_ERR_STATE, "pairwise data not set");
    if (!e.precisionSet) throw Error(MDS_HIP_ERR_STATE, "precision not set");
    const size_t rowMajorBytes = (size_t)e.N * e.P * sizeof(double);
    require_device_range(sumOut, sizeof(double), e.device, "sum output");
    if (gradientOut) require_device_range(gradientOut, rowMajorBytes, e.device, "gradient output");
    if (locations) require_device_range(locations, rowMajorBytes, e.device, "locations");
    unsigned char* base = (unsigned char*)scratch_address(e.device);
    double* tileSums = (double*)base;
    void* part = base + partOffset;
    if (locations) {
        // the new locations replace the current ones, and a pending row with them
        e.rowWritePending = false;
        e.pendingRow = -1;
        const int64_t count = (int64_t)e.N * e.kdim;
        const int blocks = (int)((count + kPackThreads - 1) / kPackThreads);
        void* target = e.full_write_target();  // a pending stored state stays where it is
        if (e.fp32)
            hipLaunchKernelGGL(pack_locations_kernel<float>, dim3(blocks), dim3(kPackThreads), 0, e.stream, locations,
                               (float*)target, e.N, e.P, e.kdim);
        else
            hipLaunchKernelGGL(pack_locations_kernel<double>, dim3(blocks), dim3(kPackThreads), 0, e.stream,
                               locations, (double*)target, e.N, e.P, e.kdim);
        check(hipGetLastError(), "location pack kernel");
        e.sumKnown = false;
        e.invalidate_device_results();
    } else {
        e.flush_row_write();
    }
    e.launch_tile_pass(gradient, nullptr, tileSums, part, e.precision);
    e.launch_reduce_pass(gradient, tileSums, part, gradientOut, sumOut,
                         e.truncated ? 1.0 : 0.5 * e.precision);  // getSumOfIncrements (:198-202)
    ++e.fullEvaluations;
    // the array is free for the next call (on any instance) once this returns
    check(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
}

}  // namespace
}  // namespace mdship

extern "C" int mds_hip_evaluate_device_noalloc(mds_hip_handle h, const double* locations, double* sumOut,
                                               double* gradientOut) {
    if (!h || !h->e) return MDS_HIP_ERR_ARGUMENT;
    mdship::Engine& e = *h->e;
    try {
        mdship::check(hipSetDevice(e.device), "hipSetDevice");
        mdship::evaluate_noalloc(e, locations, sumOut, gradientOut);
        return MDS_HIP_OK;
    } catch (const mdship::Error& err) {
        e.lastError = err.what();
        return err.code;
    } catch (const std::exception& err) {
        e.lastError = err.what();
        return MDS_HIP_ERR_DEVICE;
    }
}
