// Portable synthetic inputs, the C++ twin of massivemds_amd/datagen.py:
// splitmix64 streams keyed by (seed, key), uniforms ((u >> 11) + 0.5) 2^-53,
// Box-Muller normals, and the chisq / euclid10 observation generators.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace mdsgen {

inline uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Stream {
    uint64_t state;
    uint64_t index = 0;
    Stream(uint64_t seed, uint64_t key)
        : state(mix(seed * 0x100000001B3ULL + key * 0xC2B2AE3D27D4EB4FULL)) {}
    double uniform() {
        ++index;
        const uint64_t z = mix(state + index * 0x9E3779B97F4A7C15ULL);
        return ((double)(z >> 11) + 0.5) * 0x1.0p-53;
    }
};

inline std::vector<double> normals(uint64_t seed, uint64_t key, size_t count) {
    Stream s(seed, key);
    std::vector<double> out(count + 1);
    for (size_t k = 0; k < count; k += 2) {
        const double u1 = s.uniform(), u2 = s.uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double t = 2.0 * M_PI * u2;
        out[k] = r * std::cos(t);
        out[k + 1] = r * std::sin(t);
    }
    out.resize(count);
    return out;
}

inline std::vector<double> uniforms(uint64_t seed, uint64_t key, size_t count) {
    Stream s(seed, key);
    std::vector<double> out(count);
    for (auto& v : out) v = s.uniform();
    return out;
}

// delta_ij = z^2 per i < j, mirrored (src/benchmark.cpp:176-189); with
// `missing`, 25% of the pairs are NaN on both sides.
inline std::vector<double> chisq(int n, uint64_t seed, bool missing) {
    const size_t pairs = (size_t)n * (n - 1) / 2;
    const std::vector<double> z = normals(seed, 4, pairs);
    const std::vector<double> u = missing ? uniforms(seed, 5, pairs) : std::vector<double>();
    std::vector<double> d((size_t)n * n, 0.0);
    size_t k = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j, ++k) {
            double v = z[k] * z[k];
            if (missing && u[k] < 0.25) v = NAN;
            d[(size_t)i * n + j] = d[(size_t)j * n + i] = v;
        }
    return d;
}

// Distances of N(0, I_10) points, optional N(0, sd^2) noise clipped at 0 and
// a symmetric missing fraction.
inline std::vector<double> euclid10(int n, uint64_t seed, double noiseSd, double missing) {
    const int dim = 10;
    const std::vector<double> z = normals(seed, 1, (size_t)n * dim);
    const size_t pairs = (size_t)n * (n - 1) / 2;
    const std::vector<double> eps = noiseSd > 0 ? normals(seed, 2, pairs) : std::vector<double>();
    const std::vector<double> u = missing > 0 ? uniforms(seed, 3, pairs) : std::vector<double>();
    std::vector<double> d((size_t)n * n, 0.0);
    size_t k = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j, ++k) {
            double s = 0.0;
            for (int c = 0; c < dim; ++c) {
                const double t = z[(size_t)i * dim + c] - z[(size_t)j * dim + c];
                s += t * t;
            }
            double v = std::sqrt(s);
            if (noiseSd > 0) v = std::max(v + noiseSd * eps[k], 0.0);
            if (missing > 0 && u[k] < missing) v = NAN;
            d[(size_t)i * n + j] = d[(size_t)j * n + i] = v;
        }
    return d;
}

inline std::vector<double> locations(int n, int p, uint64_t seed, uint64_t key) {
    return normals(seed, key, (size_t)n * p);
}

}  // namespace mdsgen
