// mds_interface.hpp -- the engine interface of MassiveMDS, restated for
// builds that do not include the reference sources.
//
// Same virtual signatures, protected state and factory signature as
// mds::AbstractMultiDimensionalScaling (src/AbstractMultiDimensionalScaling.hpp:38-77)
// and the same flag bits as mds::Flags (src/flags.h:6-14), plus HIP = 1 << 10
// as an explicit alias of the GPU route.  A host that already includes the
// reference's own header uses that one instead; mds_hip_engine.hpp works
// with either.
#ifndef MDS_INTERFACE_HPP
#define MDS_INTERFACE_HPP

#include <climits>
#include <cstddef>
#include <memory>

namespace mds {

// N (N - 1) / 2 in 64-bit arithmetic, saturated to INT_MAX: the member
// stays an int as in the reference, whose int product overflows for
// N > 46,341 (src/AbstractMultiDimensionalScaling.hpp:42).
inline int observationCountOf(int locationCount) {
    const long long n = locationCount;
    const long long pairs = n * (n - 1) / 2;
    return pairs > INT_MAX ? INT_MAX : (int)pairs;
}

enum Flags {
    FLOAT = 1 << 2,
    TBB = 1 << 3,
    OPENCL = 1 << 4,
    LEFT_TRUNCATION = 1 << 5,
    SSE = 1 << 7,
    AVX = 1 << 8,
    AVX512 = 1 << 9,
    HIP = 1 << 10
};

class AbstractMultiDimensionalScaling {
public:
    AbstractMultiDimensionalScaling(int embeddingDimension, int locationCount, long flags)
        : embeddingDimension(embeddingDimension),
          locationCount(locationCount),
          observationCount(observationCountOf(locationCount)),
          flags(flags) {}

    virtual ~AbstractMultiDimensionalScaling() = default;

    virtual void updateLocations(int, double*, size_t) = 0;
    virtual double getSumOfIncrements() = 0;
    virtual void getLogLikelihoodGradient(double*, size_t) = 0;
    virtual void storeState() = 0;
    virtual void restoreState() = 0;
    virtual void acceptState() = 0;
    virtual void setPairwiseData(double*, size_t) = 0;
    virtual void setParameters(double*, size_t) = 0;
    virtual void makeDirty() = 0;
    virtual int getInternalDimension() = 0;

protected:
    int embeddingDimension;
    int locationCount;
    int observationCount;
    long flags;

    int updatedLocation = -1;
    bool incrementsKnown = false;
    bool sumOfIncrementsKnown = false;
    bool isLeftTruncated = false;
};

typedef std::shared_ptr<mds::AbstractMultiDimensionalScaling> SharedPtr;

SharedPtr factory(int dim1, int dim2, long flags, int device, int threads);

}  // namespace mds

#endif  // MDS_INTERFACE_HPP
