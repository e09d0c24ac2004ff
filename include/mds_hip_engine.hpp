// mds_hip_engine.hpp -- the HIP engine as a mds::AbstractMultiDimensionalScaling.
//
// Include after the declaration of mds::AbstractMultiDimensionalScaling
// (either the reference's src/AbstractMultiDimensionalScaling.hpp or this
// project's mds_interface.hpp) and link libmds_hip.so.  The class forwards
// each virtual to the C ABI in mds_hip.h; a failed call throws
// std::runtime_error with the engine's message (or calls Rcpp::stop under
// RBUILD), where the reference engines assert or exit.
//
// constructHipMultiDimensionalScalingDouble / ...Float take the place of
// constructOpenCLMultiDimensionalScalingDouble / ...Float
// (src/backend/opencl/instantiate.cpp:14-52) in the factory's GPU branch
// (src/factory.cpp:49-56, 75-81).
#ifndef MDS_HIP_ENGINE_HPP
#define MDS_HIP_ENGINE_HPP

#include <memory>
#include <stdexcept>
#include <string>

#include "mds_hip.h"

#ifdef RBUILD
#include <Rcpp.h>
#endif

namespace mds {

class HipMultiDimensionalScaling : public AbstractMultiDimensionalScaling {
public:
    HipMultiDimensionalScaling(int embeddingDimension, int locationCount, long flags, int device)
        : AbstractMultiDimensionalScaling(embeddingDimension, locationCount, flags) {
        if (flags & (1L << 5)) isLeftTruncated = true;  // LEFT_TRUNCATION
        const int st = mds_hip_create(embeddingDimension, locationCount, flags, device, &handle);
        if (st != MDS_HIP_OK) fail(st, mds_hip_last_error(nullptr));
    }

    ~HipMultiDimensionalScaling() override {
        if (handle) mds_hip_destroy(handle);
    }

    HipMultiDimensionalScaling(const HipMultiDimensionalScaling&) = delete;
    HipMultiDimensionalScaling& operator=(const HipMultiDimensionalScaling&) = delete;

    void updateLocations(int locationIndex, double* location, size_t length) override {
        check(mds_hip_update_locations(handle, locationIndex, location, length));
        sumOfIncrementsKnown = false;
    }

    double getSumOfIncrements() override {
        double out = 0.0;
        check(mds_hip_get_sum_of_increments(handle, &out));
        return out;
    }

    void getLogLikelihoodGradient(double* result, size_t length) override {
        check(mds_hip_get_log_likelihood_gradient(handle, result, length));
    }

    void storeState() override { check(mds_hip_store_state(handle)); }
    void restoreState() override { check(mds_hip_restore_state(handle)); }
    void acceptState() override { check(mds_hip_accept_state(handle)); }

    void setPairwiseData(double* data, size_t length) override {
        check(mds_hip_set_pairwise_data(handle, data, length));
    }

    void setParameters(double* data, size_t length) override {
        check(mds_hip_set_parameters(handle, data, length));
    }

    void makeDirty() override { check(mds_hip_make_dirty(handle)); }

    int getInternalDimension() override {
        int out = 0;
        check(mds_hip_get_internal_dimension(handle, &out));
        return out;
    }

    // Device-resident evaluation (no host copy), for samplers and timing.
    void evaluateAsync(bool gradient) { check(mds_hip_evaluate_async(handle, gradient ? 1 : 0)); }
    void synchronize() { check(mds_hip_synchronize(handle)); }
    mds_hip_handle nativeHandle() const { return handle; }

private:
    mds_hip_handle handle = nullptr;

    [[noreturn]] static void fail(int status, const char* message) {
        const std::string text = std::string("HIP MDS engine (status ") + std::to_string(status) +
                                 "): " + (message ? message : "");
#ifdef RBUILD
        Rcpp::stop(text);
#endif
        throw std::runtime_error(text);
    }

    void check(int status) const {
        if (status != MDS_HIP_OK) fail(status, mds_hip_last_error(handle));
    }
};

inline std::shared_ptr<AbstractMultiDimensionalScaling>
constructHipMultiDimensionalScalingDouble(int embeddingDimension, int locationCount, long flags, int device) {
    return std::make_shared<HipMultiDimensionalScaling>(embeddingDimension, locationCount,
                                                        flags & ~(1L << 2), device);
}

inline std::shared_ptr<AbstractMultiDimensionalScaling>
constructHipMultiDimensionalScalingFloat(int embeddingDimension, int locationCount, long flags, int device) {
    return std::make_shared<HipMultiDimensionalScaling>(embeddingDimension, locationCount,
                                                        flags | (1L << 2), device);
}

}  // namespace mds

#endif  // MDS_HIP_ENGINE_HPP
