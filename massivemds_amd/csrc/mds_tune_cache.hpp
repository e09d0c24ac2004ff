// On-disk cache of tuned launch configurations (host only): one line per
// batch signature, "<signature> <field> ... <field> <ms>", the signature
// without blanks.  A store replaces the signature's line in place (or
// appends it) and rewrites the file through a temporary and a rename, so the
// file never holds two lines for one signature and a forced re-tune is what
// later loads find.  A load takes the last well-formed line of the
// signature, which also reads files written by older versions that
// appended.
#pragma once

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace mdship {

struct TuneRecord {
    std::vector<long long> fields;
    double ms = 0.0;
};

// Parses one line: signature, nfields integers, the time.
inline bool tune_cache_parse(const std::string& line, size_t nfields, std::string& sig, TuneRecord& r) {
    std::istringstream ls(line);
    if (!(ls >> sig)) return false;
    r.fields.assign(nfields, 0);
    for (size_t k = 0; k < nfields; ++k)
        if (!(ls >> r.fields[k])) return false;
    if (!(ls >> r.ms)) return false;
    std::string rest;
    return !(ls >> rest);  // nothing after the time
}

inline bool tune_cache_find(const char* path, const std::string& sig, size_t nfields, TuneRecord& out) {
    std::ifstream in(path);
    std::string line, s;
    TuneRecord r;
    bool found = false;
    while (std::getline(in, line))
        if (tune_cache_parse(line, nfields, s, r) && s == sig) {
            out = r;
            found = true;
        }
    return found;
}

// Returns false when the file could not be written (the tuned choice still
// applies to the running batch).
inline bool tune_cache_store(const char* path, const std::string& sig, const TuneRecord& r) {
    std::vector<std::string> lines;
    {
        std::ifstream in(path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string s;
            if ((ls >> s) && s == sig) continue;  // replaced below
            if (!line.empty()) lines.push_back(line);
        }
    }
    std::ostringstream rec;
    rec.precision(17);  // the time reads back exactly
    rec << sig;
    for (long long f : r.fields) rec << " " << f;
    rec << " " << r.ms;
    lines.push_back(rec.str());
    const std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream out(tmp, std::ios_base::trunc);
        if (!out) return false;
        for (const std::string& l : lines) out << l << "\n";
        if (!out.flush()) return false;
    }
    return std::rename(tmp.c_str(), path) == 0;
}

}  // namespace mdship
