// mds::factory for builds against this project's interface header
// (mds_interface.hpp): the flag-driven engine selection of
// src/factory.cpp:32-118 restricted to the GPU route, which the OPENCL bit
// (or its HIP alias) selects; FLOAT picks the FP32 engine.  The reference's
// CPU engines are not part of this library, so CPU flag sets throw.
#include <stdexcept>

#include "mds_interface.hpp"
#include "mds_hip_engine.hpp"

namespace mds {

SharedPtr factory(int dim1, int dim2, long flags, int device, int threads) {
    (void)threads;
    const bool useGpu = (flags & Flags::OPENCL) || (flags & Flags::HIP);
    if (!useGpu)
        throw std::invalid_argument(
            "libmds_hip provides the GPU engine only: set mds::Flags::OPENCL (or HIP)");
    if (flags & Flags::FLOAT) return constructHipMultiDimensionalScalingFloat(dim1, dim2, flags, device);
    return constructHipMultiDimensionalScalingDouble(dim1, dim2, flags, device);
}

}  // namespace mds
