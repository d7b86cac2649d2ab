// Reads MMPLD files through the renderer's C++ data source (megamol_amd/module/MMPLDSource.cpp),
// for an AddressSanitizer / UndefinedBehaviorSanitizer build: every file named on the command
// line, then every truncation of each and a copy with flipped header bytes, which must be
// rejected or read without touching memory outside the file. Prints one line per file: lists,
// particles and the sum of the first coordinates (so the views are dereferenced).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "MMPLDSource.h"

using namespace megamol;

static double touch(geocalls::MultiParticleDataCall& c, uint64_t& parts) {
    double s = 0.0;
    for (unsigned i = 0; i < c.GetParticleListCount(); ++i) {
        auto& p = c.AccessParticles(i);
        parts += p.GetCount();
        const uint8_t* v = static_cast<const uint8_t*>(p.GetVertexData());
        for (uint64_t k = 0; v && k < p.GetCount(); ++k) {
            uint8_t b[4];
            std::memcpy(b, v + k * p.GetVertexDataStride(), 4);
            s += b[0];
        }
        const uint8_t* col = static_cast<const uint8_t*>(p.GetColourData());
        for (uint64_t k = 0; col && k < p.GetCount(); ++k) s += col[k * p.GetColourDataStride()];
    }
    return s;
}

static bool read_one(const std::string& path, bool print) {
    mmhrt_plugin::MMPLDSource src;
    if (!src.Open(path)) return false;
    geocalls::MultiParticleDataCall c;
    bool ok = true;
    for (unsigned f = 0; f < src.FrameCount(); ++f) {
        c.SetFrameID(f);
        if (!src.Serve(c, 1) || !src.Serve(c, 0)) {
            ok = false;
            break;
        }
        uint64_t parts = 0;
        double s = touch(c, parts);
        if (print) std::printf("%s frame %u: %u lists, %llu particles, %.1f\n", path.c_str(), f,
                               c.GetParticleListCount(), (unsigned long long)parts, s);
    }
    return ok;
}

int main(int argc, char** argv) {
    int bad = 0;
    for (int i = 1; i < argc; ++i) {
        if (!read_one(argv[i], true)) {
            std::printf("%s: not read\n", argv[i]);
            bad = 1;
        }
        std::ifstream f(argv[i], std::ios::binary);
        std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        const std::string tmp = std::string(argv[i]) + ".cut";
        for (size_t len = 0; len < b.size(); len += 7) {
            std::ofstream(tmp, std::ios::binary).write(b.data(), (std::streamsize)len);
            read_one(tmp, false);  // rejected or read within the bytes
        }
        for (size_t at = 6; at < b.size() && at < 200; at += 3) {
            std::vector<char> c2 = b;
            c2[at] = (char)~c2[at];
            std::ofstream(tmp, std::ios::binary).write(c2.data(), (std::streamsize)c2.size());
            read_one(tmp, false);
        }
        std::remove(tmp.c_str());
    }
    return bad;
}
