// mds_benchmark -- command-line benchmark of the HIP engine through the
// engine interface and mds::factory, with the options and output lines of
// the reference's benchmark (src/benchmark.cpp:28-332):
//
//   --gpu k (k > 0: GPU engine on device k - 1)  --tbb t  --float  --truncation
//   --iterations n  --locations N  --dimension P  --internal  --missing
//   --sse / --avx / --avx512 (accepted; they only select CPU engines)
// plus --generator chisq|euclid10 (default chisq, the reference's z^2 data),
//      --precision tau (default 1), --seed s (default 666).
//
// Loop per iteration, as the reference: storeState, new locations for all
// rows, time getSumOfIncrements, restore with probability 0.75 otherwise
// accept, time getLogLikelihoodGradient.  Prints AvgLogLik (mean
// sumOfIncrements), AvgSumGradient (mean of gradient[1]), the two timers and
// the total, and appends a line to report.txt.  Inputs come from the
// portable generator shared with the Python side (mds_datagen.hpp).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "mds_datagen.hpp"
#include "mds_interface.hpp"

namespace {

struct Options {
    std::map<std::string, std::string> values;
    bool has(const std::string& k) const { return values.count(k) != 0; }
    long getInt(const std::string& k, long d) const { return has(k) ? std::atol(values.at(k).c_str()) : d; }
    double getDouble(const std::string& k, double d) const {
        return has(k) ? std::atof(values.at(k).c_str()) : d;
    }
    std::string get(const std::string& k, const std::string& d) const { return has(k) ? values.at(k) : d; }
};

const char* kUsage =
    "Allowed options:\n"
    "  --help                produce help message\n"
    "  --gpu arg (=0)        number of GPU on which to run\n"
    "  --tbb arg (=0)        use TBB with specified number of threads\n"
    "  --float               run in single-precision\n"
    "  --truncation          enable truncation\n"
    "  --iterations arg (=1) number of iterations\n"
    "  --locations arg (=3)  number of locations\n"
    "  --dimension arg (=2)  number of dimensions\n"
    "  --internal            use internal dimension\n"
    "  --missing             allow for missing entries\n"
    "  --sse | --avx | --avx512  hand-rolled SIMD (CPU engines only)\n"
    "  --generator arg (=chisq)  chisq | euclid10\n"
    "  --precision arg (=1)  MDS precision tau\n"
    "  --seed arg (=666)     data seed\n";

bool parse(int argc, char** argv, Options& o) {
    static const char* flagsOnly[] = {"help", "float", "truncation", "internal", "missing",
                                      "sse", "avx", "avx512"};
    static const char* withValue[] = {"gpu", "tbb", "iterations", "locations", "dimension",
                                      "generator", "precision", "seed"};
    for (int a = 1; a < argc; ++a) {
        std::string s = argv[a];
        if (s.rfind("--", 0) != 0) return false;
        s = s.substr(2);
        std::string value;
        const size_t eq = s.find('=');
        if (eq != std::string::npos) {
            value = s.substr(eq + 1);
            s = s.substr(0, eq);
        }
        bool known = false;
        for (const char* f : flagsOnly)
            if (s == f) {
                o.values[s] = "1";
                known = true;
            }
        for (const char* f : withValue)
            if (s == f) {
                if (value.empty()) {
                    if (a + 1 >= argc) return false;
                    value = argv[++a];
                }
                o.values[s] = value;
                known = true;
            }
        if (!known) return false;
    }
    return true;
}

}  // namespace

int main(int argc, char* argv[]) {
    Options opt;
    if (!parse(argc, argv, opt) || opt.has("help")) {
        std::cout << kUsage << std::endl;
        return 1;
    }
    const uint64_t seed = (uint64_t)opt.getInt("seed", 666);
    std::cout << "Loading data" << std::endl;

    const int embeddingDimension = (int)opt.getInt("dimension", 2);
    const int locationCount = (int)opt.getInt("locations", 3);
    long flags = 0;
    int deviceNumber = -1;
    int threads = 0;
    if (opt.getInt("gpu", 0) > 0) {
        std::cout << "Running on GPU" << std::endl;
        flags |= mds::Flags::OPENCL;
        deviceNumber = (int)opt.getInt("gpu", 0) - 1;
    } else {
        std::cout << "Running on CPU" << std::endl;
        threads = (int)opt.getInt("tbb", 0);
        if (threads != 0) flags |= mds::Flags::TBB;
    }
    if (opt.has("float")) {
        std::cout << "Running in single-precision" << std::endl;
        flags |= mds::Flags::FLOAT;
    } else {
        std::cout << "Running in double-precision" << std::endl;
    }
    const char* simd = "no simd";
    const int simdCount = (int)opt.has("sse") + (int)opt.has("avx") + (int)opt.has("avx512");
    if (simdCount > 1) {
        std::cerr << "Can not request more than one SIMD simultaneously" << std::endl;
        return 1;
    }
    if (opt.has("avx512")) {
        flags |= mds::Flags::AVX512;
        simd = "avx512";
    } else if (opt.has("avx")) {
        flags |= mds::Flags::AVX;
        simd = "avx";
    } else if (opt.has("sse")) {
        flags |= mds::Flags::SSE;
        simd = "sse";
    }
    if (opt.has("truncation")) {
        std::cout << "Enabling truncation" << std::endl;
        flags |= mds::Flags::LEFT_TRUNCATION;
    }
    const bool internalDimension = opt.has("internal");

    mds::SharedPtr instance;
    try {
        instance = mds::factory(embeddingDimension, locationCount, flags, deviceNumber, threads);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }

    const bool missing = opt.has("missing");
    if (missing) std::cout << "Allowing for missingness" << std::endl;
    const std::string generator = opt.get("generator", "chisq");
    std::vector<double> data;
    if (generator == "chisq") {
        data = mdsgen::chisq(locationCount, seed, missing);
    } else if (generator == "euclid10") {
        data = mdsgen::euclid10(locationCount, seed, 0.0, missing ? 0.05 : 0.0);
    } else {
        std::cerr << "unknown generator " << generator << std::endl;
        return 1;
    }
    const size_t elementCount = (size_t)locationCount * locationCount;
    instance->setPairwiseData(data.data(), elementCount);

    const int dataDimension = internalDimension ? instance->getInternalDimension() : embeddingDimension;
    const std::vector<double> initial = mdsgen::locations(locationCount, dataDimension, seed + 1, 1);
    for (int i = 0; i < locationCount; ++i) {
        std::vector<double> row(initial.begin() + (size_t)i * dataDimension,
                                initial.begin() + (size_t)(i + 1) * dataDimension);
        instance->updateLocations(i, row.data(), dataDimension);
    }

    const int gradientIndex = 1;
    double precision = opt.getDouble("precision", 1.0);
    instance->setParameters(&precision, 1);
    instance->makeDirty();
    double logLik = instance->getSumOfIncrements();
    std::vector<double> gradient((size_t)locationCount * dataDimension);
    instance->getLogLikelihoodGradient(gradient.data(), gradient.size());
    double sumGradient = gradient[gradientIndex];

    std::cout << "Starting MDS benchmark" << std::endl;
    const auto startTime = std::chrono::steady_clock::now();
    const int iterations = (int)opt.getInt("iterations", 1);
    const std::vector<double> restoreDraw = mdsgen::uniforms(seed, 6, (size_t)std::max(iterations, 1));
    double timer = 0.0, timer2 = 0.0;
    for (int itr = 0; itr < iterations; ++itr) {
        instance->storeState();
        std::vector<double> all = mdsgen::locations(locationCount, dataDimension, seed + 1, 2 + itr);
        instance->updateLocations(-1, all.data(), all.size());

        const auto t1 = std::chrono::steady_clock::now();
        logLik += instance->getSumOfIncrements();
        timer += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();

        if (restoreDraw[itr] < 0.75)
            instance->restoreState();
        else
            instance->acceptState();

        const auto t2 = std::chrono::steady_clock::now();
        instance->getLogLikelihoodGradient(gradient.data(), gradient.size());
        timer2 += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t2).count();
        sumGradient += gradient[gradientIndex];
    }
    logLik /= iterations + 1;
    sumGradient /= iterations + 1;
    const double total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - startTime).count();

    std::cout << "End MDS benchmark" << std::endl;
    std::cout.precision(17);
    std::cout << "AvgLogLik = " << logLik << std::endl;
    std::cout << "AvgSumGradient = " << sumGradient << std::endl;
    std::cout.precision(6);
    std::cout << timer << " ms" << std::endl;
    std::cout << timer2 << " ms" << std::endl;
    std::cout << total << " ms " << std::endl;

    std::ofstream outfile("report.txt", std::ios_base::app);
    outfile << deviceNumber << " " << threads << " " << simd << " " << locationCount << " "
            << embeddingDimension << " " << iterations << " " << timer << " " << timer2 << "\n";
    return 0;
}
