// The autotuner's on-disk cache (massivemds_amd/csrc/mds_tune_cache.hpp):
// a forced re-tune's record is what later loads find, one line per
// signature, other signatures' lines kept, files written by appending
// (several lines for one signature) read as their last line, malformed
// lines skipped.  Exit status 0 = all checks hold.
#include <cstdio>
#include <fstream>
#include <string>

#include "mds_tune_cache.hpp"

using namespace mdship;

static int failures = 0;

static void expect(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "failed: %s\n", what);
        ++failures;
    }
}

static int count_lines(const char* path, const std::string& sig) {
    std::ifstream in(path);
    std::string line;
    int n = 0;
    while (std::getline(in, line))
        if (line.compare(0, sig.size() + 1, sig + " ") == 0) ++n;
    return n;
}

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    const char* path = argv[1];
    std::remove(path);
    TuneRecord r;
    expect(!tune_cache_find(path, "f32-T-cu256-2x8", 4, r), "missing file finds nothing");

    TuneRecord a{{4, 8, 1, 0}, 0.789};
    TuneRecord b{{2, 4, 2, 1}, 1.25};
    expect(tune_cache_store(path, "f32-T-cu256-2x8", a), "store a");
    expect(tune_cache_store(path, "f64-N-cu256-5x3", b), "store b");
    expect(tune_cache_find(path, "f32-T-cu256-2x8", 4, r) && r.fields == a.fields && r.ms == a.ms, "find a");

    // forced re-tune of the first signature
    TuneRecord a2{{8, 2, 4, 1}, 0.6543210987654321};
    expect(tune_cache_store(path, "f32-T-cu256-2x8", a2), "store a2");
    expect(tune_cache_find(path, "f32-T-cu256-2x8", 4, r) && r.fields == a2.fields && r.ms == a2.ms,
           "re-tuned record wins, time exact");
    expect(count_lines(path, "f32-T-cu256-2x8") == 1, "one line per signature");
    expect(tune_cache_find(path, "f64-N-cu256-5x3", 4, r) && r.fields == b.fields, "other signature kept");

    // a file written by appending, with a malformed line and a prefix signature
    {
        std::ofstream out(path, std::ios_base::app);
        out << "f32-T-cu256-2x8 1 1 1 0 9.0\n";
        out << "f32-T-cu256-2x8 7 7\n";
        out << "f32-T-cu256-2x8x 3 3 3 0 1.0\n";
    }
    expect(tune_cache_find(path, "f32-T-cu256-2x8", 4, r) && r.fields[0] == 1 && r.ms == 9.0,
           "last well-formed line of an appended file");
    expect(!tune_cache_find(path, "f32-T-cu256", 4, r), "no prefix match");
    std::remove(path);
    return failures == 0 ? 0 : 1;
}
