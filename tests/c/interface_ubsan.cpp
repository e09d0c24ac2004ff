// The restated engine interface (include/mds_interface.hpp) constructed at
// location counts whose N (N - 1) exceeds INT_MAX (N > 46,341) and whose
// N (N - 1) / 2 does too (N > 65,536), under
// UndefinedBehaviorSanitizer: the count is formed in 64-bit arithmetic and
// saturated, so no signed overflow happens.  Exit status 0 = clean.
#include <climits>
#include <cstdio>

#include "mds_interface.hpp"

namespace {

class Probe : public mds::AbstractMultiDimensionalScaling {
public:
    Probe(int p, int n) : AbstractMultiDimensionalScaling(p, n, mds::OPENCL) {}
    void updateLocations(int, double*, size_t) override {}
    double getSumOfIncrements() override { return 0.0; }
    void getLogLikelihoodGradient(double*, size_t) override {}
    void storeState() override {}
    void restoreState() override {}
    void acceptState() override {}
    void setPairwiseData(double*, size_t) override {}
    void setParameters(double*, size_t) override {}
    void makeDirty() override {}
    int getInternalDimension() override { return embeddingDimension; }
    int observations() const { return observationCount; }
};

}  // namespace

int main() {
    int failures = 0;
    const struct {
        int n;
        int expected;
    } cases[] = {{1, 0}, {2, 1}, {1000, 499500}, {46341, 1073720970}, {46342, 1073767311}, {50000, 1249975000},
                 {65536, 2147450880}, {65537, INT_MAX}, {INT_MAX, INT_MAX}};
    for (const auto& c : cases) {
        Probe probe(2, c.n);
        if (probe.observations() != c.expected) {
            std::fprintf(stderr, "N=%d: observationCount %d, expected %d\n", c.n, probe.observations(), c.expected);
            ++failures;
        }
    }
    return failures == 0 ? 0 : 1;
}
