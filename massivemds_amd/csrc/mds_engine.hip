// libmds_hip.so: the C ABI declared in include/mds_hip.h over the engine in
// mds_engine.hpp.
#include <hip/hip_runtime.h>

#include <string>

#include "mds_engine.hpp"

static thread_local std::string g_create_error = "no error";

template <typename F>
static int guarded(mds_hip_handle h, F&& f) {
    if (!h || !h->e) {
        g_create_error = "null handle";
        return MDS_HIP_ERR_ARGUMENT;
    }
    try {
        (void)hipSetDevice(h->e->device);
        f(*h->e);
        return MDS_HIP_OK;
    } catch (const mdship::Error& err) {
        h->e->lastError = err.what();
        return err.code;
    } catch (const std::exception& err) {
        h->e->lastError = err.what();
        return MDS_HIP_ERR_DEVICE;
    }
}

extern "C" {

int mds_hip_max_dimension(void) { return mdship::kMaxDim; }

int mds_hip_device_count(int* count) {
    if (!count) return MDS_HIP_ERR_ARGUMENT;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return MDS_HIP_OK;
}

int mds_hip_create(int embeddingDimension, int locationCount, long flags, int device,
                   mds_hip_handle* out) {
    if (!out) {
        g_create_error = "null output handle";
        return MDS_HIP_ERR_ARGUMENT;
    }
    *out = nullptr;
    try {
        auto* e = new mdship::Engine(embeddingDimension, locationCount, flags, device);
        *out = new mds_hip_engine{e};
        return MDS_HIP_OK;
    } catch (const mdship::Error& err) {
        g_create_error = err.what();
        return err.code;
    } catch (const std::exception& err) {
        g_create_error = err.what();
        return MDS_HIP_ERR_DEVICE;
    }
}

int mds_hip_destroy(mds_hip_handle h) {
    if (!h) return MDS_HIP_ERR_ARGUMENT;
    delete h->e;
    delete h;
    return MDS_HIP_OK;
}

const char* mds_hip_last_error(mds_hip_handle h) {
    if (!h || !h->e) return g_create_error.c_str();
    return h->e->lastError.c_str();
}

int mds_hip_update_locations(mds_hip_handle h, int locationIndex, const double* location,
                             size_t length) {
    return guarded(h, [&](mdship::Engine& e) { e.updateLocations(locationIndex, location, length); });
}

int mds_hip_get_sum_of_increments(mds_hip_handle h, double* out) {
    return guarded(h, [&](mdship::Engine& e) {
        if (!out) throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "null output pointer");
        *out = e.getSumOfIncrements();
    });
}

int mds_hip_get_log_likelihood_gradient(mds_hip_handle h, double* result, size_t length) {
    return guarded(h, [&](mdship::Engine& e) { e.getLogLikelihoodGradient(result, length); });
}

int mds_hip_store_state(mds_hip_handle h) {
    return guarded(h, [&](mdship::Engine& e) { e.storeState(); });
}

int mds_hip_restore_state(mds_hip_handle h) {
    return guarded(h, [&](mdship::Engine& e) { e.restoreState(); });
}

int mds_hip_accept_state(mds_hip_handle h) {
    // No per-pair increments are cached, so there is nothing to commit
    // (src/NewMultiDimensionalScaling.hpp:226-232 only copies increments).
    return guarded(h, [&](mdship::Engine&) {});
}

int mds_hip_set_pairwise_data(mds_hip_handle h, const double* data, size_t length) {
    return guarded(h, [&](mdship::Engine& e) { e.setPairwiseData(data, length); });
}

int mds_hip_set_parameters(mds_hip_handle h, const double* data, size_t length) {
    return guarded(h, [&](mdship::Engine& e) { e.setParameters(data, length); });
}

int mds_hip_make_dirty(mds_hip_handle h) {
    return guarded(h, [&](mdship::Engine& e) { e.makeDirty(); });
}

int mds_hip_get_internal_dimension(mds_hip_handle h, int* out) {
    return guarded(h, [&](mdship::Engine& e) {
        if (!out) throw mdship::Error(MDS_HIP_ERR_ARGUMENT, "null output pointer");
        *out = e.P;
    });
}

int mds_hip_evaluate_async(mds_hip_handle h, int computeGradient) {
    return guarded(h, [&](mdship::Engine& e) { e.evaluate(computeGradient != 0); });
}

int mds_hip_synchronize(mds_hip_handle h) {
    return guarded(h, [&](mdship::Engine& e) {
        mdship::check(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
    });
}

int mds_hip_set_timing(mds_hip_handle h, int enable) {
    return guarded(h, [&](mdship::Engine& e) { e.timing = enable != 0; });
}

int mds_hip_get_timing(mds_hip_handle h, double* tileMs, double* evalMs, long long* count) {
    return guarded(h, [&](mdship::Engine& e) { e.getTiming(tileMs, evalMs, count); });
}

int mds_hip_get_counters(mds_hip_handle h, long long* fullEvaluations, long long* rowUpdates) {
    return guarded(h, [&](mdship::Engine& e) {
        if (fullEvaluations) *fullEvaluations = e.fullEvaluations;
        if (rowUpdates) *rowUpdates = e.rowUpdates;
    });
}

int mds_hip_get_layout(mds_hip_handle h, int* symmetric, long long* observationBytes,
                       long long* tiles) {
    return guarded(h, [&](mdship::Engine& e) {
        if (symmetric) *symmetric = e.symmetric ? 1 : 0;
        if (observationBytes) *observationBytes = (long long)e.deltaBytes;
        if (tiles) *tiles = e.numTiles;
    });
}

}  // extern "C"
