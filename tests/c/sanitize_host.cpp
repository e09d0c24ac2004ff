// The host side of the engine that arranges observations for the kernels
// (massivemds_amd/csrc/mds_host.hpp: the tile list and the packing into the
// skewed tile layout), run under AddressSanitizer and
// UndefinedBehaviorSanitizer with ragged N (padding rows and columns),
// missing entries and asymmetric data.  Every packed value is compared
// with the layout's definition: element k * 64 + l of tile (I, J) is
// delta(64 I + l, 64 J + (l + k) mod 64) in plane 0 and the transposed
// entry in plane 1, NaN outside the matrix; the tile list holds each tile
// of the upper block triangle once, off-diagonal tiles row-major first, at
// the index the single-row kernel computes in closed form
// (I (nb - 1) - I (I - 1) / 2 + J - I - 1; diagonal: nb (nb - 1) / 2 + I).
// Exit status 0 = clean.
#include <cmath>
#include <cstdio>
#include <vector>

#include "mds_host.hpp"

using namespace mdship;

static int failures = 0;

static int64_t closed_form_index(int I, int J, int nb) {
    if (I == J) return (int64_t)nb * (nb - 1) / 2 + I;
    return (int64_t)I * (nb - 1) - (int64_t)I * (I - 1) / 2 + (J - I - 1);
}

template <typename Real>
static bool same(Real a, double b) {
    return (std::isnan((double)a) && std::isnan(b)) || (double)a == (double)(Real)b;
}

template <typename Real>
static void check_pack(int n, bool asym) {
    const int nb = (n + kTile - 1) / kTile;
    const std::vector<uint32_t> tiles = make_tiles(nb);
    if ((int64_t)tiles.size() != (int64_t)nb * (nb + 1) / 2) {
        std::fprintf(stderr, "n=%d: %zu tiles\n", n, tiles.size());
        ++failures;
    }
    for (size_t t = 0; t < tiles.size(); ++t) {
        const int I = (int)(tiles[t] >> 16), J = (int)(tiles[t] & 0xffffu);
        if (I > J || closed_form_index(I, J, nb) != (int64_t)t) {
            std::fprintf(stderr, "n=%d: tile %zu is (%d, %d)\n", n, t, I, J);
            ++failures;
        }
    }
    std::vector<double> d((size_t)n * n);
    uint64_t s = 12345 + (uint64_t)n;
    auto uni = [&]() {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(s >> 11) * 0x1.0p-53;
    };
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            const double v = (i == j) ? 0.0 : (uni() < 0.05 ? NAN : uni());
            d[(size_t)i * n + j] = v;
            d[(size_t)j * n + i] = (asym && i != j && uni() < 0.5) ? v + 0.25 : v;
        }
    std::vector<Real> p0(tiles.size() * kTileElems), p1(tiles.size() * kTileElems);
    const bool sym = pack_tiles<Real>(d.data(), n, tiles, p0.data(), p1.data());
    bool anyAsym = false;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (!same<Real>((Real)d[(size_t)i * n + j], d[(size_t)j * n + i])) anyAsym = true;
    if (sym == anyAsym) {
        std::fprintf(stderr, "n=%d: symmetric flag %d\n", n, (int)sym);
        ++failures;
    }
    for (size_t t = 0; t < tiles.size(); ++t) {
        const int I = (int)(tiles[t] >> 16), J = (int)(tiles[t] & 0xffffu);
        for (int k = 0; k < kTile; ++k)
            for (int l = 0; l < kTile; ++l) {
                const int64_t i = (int64_t)I * kTile + l, j = (int64_t)J * kTile + ((l + k) % kTile);
                const bool in = i < n && j < n;
                const double a = in ? d[(size_t)i * n + j] : NAN, b = in ? d[(size_t)j * n + i] : NAN;
                const size_t e = t * kTileElems + (size_t)k * kTile + l;
                if (!same<Real>(p0[e], a) || !same<Real>(p1[e], b)) {
                    if (failures < 10) std::fprintf(stderr, "n=%d tile (%d,%d) k=%d l=%d\n", n, I, J, k, l);
                    ++failures;
                }
            }
    }
}

int main() {
    for (int n : {1, 2, 63, 64, 65, 130, 200}) {
        for (bool asym : {false, true}) {
            check_pack<double>(n, asym);
            check_pack<float>(n, asym);
        }
    }
    if (failures) std::fprintf(stderr, "%d failures\n", failures);
    return failures ? 1 : 0;
}
