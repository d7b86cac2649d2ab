/*
 * MMPLDSource — an MMPLD file as the source of a MultiParticleDataCall, the part of the reference's
 * MMPLDDataSource (plugins/moldyn/src/io/MMPLDDataSource.cpp) that a renderer sees: the file header
 * and seek table (:373-399), and per frame the particle lists as strided views into the frame
 * bytes (:61-216), with the data hash, frame count and bounding boxes of the call
 * (:443-461, 480-497). Versions 100-103. Used by the headless project runner.
 */
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "geometry_calls/MultiParticleDataCall.h"

namespace megamol::mmhrt_plugin {

class MMPLDSource {
public:
    // Reads the whole file; false (with Error()) when it is not a readable MMPLD file.
    bool Open(const std::string& path);
    // MultiParticleDataCall function 0 (GetData) or 1 (GetExtent) for the call's frame id.
    bool Serve(geocalls::MultiParticleDataCall& call, int fn);
    const std::string& Error() const { return err_; }
    unsigned int FrameCount() const { return frames_; }
    // A changed hash makes a renderer re-upload (the reference bumps it on a new file, :480-497).
    void SetDataHash(size_t h) { hash_ = h; }

private:
    bool load_frame(unsigned int idx);

    std::vector<uint8_t> file_;
    uint16_t version_ = 0;
    unsigned int frames_ = 0;
    float bbox_[6] = {0, 0, 0, 0, 0, 0};
    std::vector<uint64_t> seek_;
    std::vector<uint8_t> frame_;  // bytes of the loaded frame (the lists point into them)
    unsigned int loaded_ = ~0u;
    struct List {
        geocalls::SimpleSphericalParticles::VertexDataType vt;
        geocalls::SimpleSphericalParticles::ColourDataType ct;
        size_t vert_off, col_off;
        unsigned int stride;
        uint64_t count;
        float radius;
        unsigned char rgba[4];
        float cmin, cmax;
    };
    std::vector<List> lists_;
    size_t hash_ = 1;
    std::string err_;
};

}  // namespace megamol::mmhrt_plugin
