/*
 * MMPLDSource — see MMPLDSource.h. Byte layout (plugins/moldyn/src/io/MMPLDDataSource.cpp):
 *   header:  "MMPLD\0", u16 version (100..103), u32 frame count, 6 f32 bounding box, 6 f32 clip
 *            box, u64 seek table of frame count + 1 offsets (:373-399)
 *   frame:   [f32 timestamp, version >= 102], u32 list count, then per list: u8 vertex code, u8
 *            colour code, [f32 global radius for vertex codes 1, 3, 4], [4 u8 global colour (the
 *            alpha byte read and replaced by 255) for colour code 0], [2 f32 intensity range for
 *            colour codes 3 and 7], u64 count, [6 f32 list box, version >= 103], the interleaved
 *            records (vertex then colour bytes), [version 101: u32 cluster count, u64 byte size,
 *            cluster bytes] (:61-216). Which optional fields follow is decided by the raw codes;
 *            an unknown code reads as NONE for the record layout (:89-157).
 */
#include "MMPLDSource.h"

#include <cstring>
#include <fstream>

namespace megamol::mmhrt_plugin {

namespace {
using SSP = geocalls::SimpleSphericalParticles;

template <typename T>
bool rd(const std::vector<uint8_t>& b, size_t& p, T& v) {
    if (p + sizeof(T) > b.size()) return false;
    std::memcpy(&v, b.data() + p, sizeof(T));
    p += sizeof(T);
    return true;
}

unsigned int vert_size(int code) {
    switch (code) {
    case 1: return 12;
    case 2: return 16;
    case 3: return 6;
    case 4: return 24;
    default: return 0;
    }
}
unsigned int col_size(int code) {
    switch (code) {
    case 1: return 3;
    case 2: return 4;
    case 3: return 4;
    case 4: return 12;
    case 5: return 16;
    case 6: return 8;
    case 7: return 8;
    default: return 0;
    }
}
// file colour code -> in-memory type (file 3 is FLOAT_I, 4 FLOAT_RGB, 5 FLOAT_RGBA)
SSP::ColourDataType col_type(int code) {
    switch (code) {
    case 1: return SSP::COLDATA_UINT8_RGB;
    case 2: return SSP::COLDATA_UINT8_RGBA;
    case 3: return SSP::COLDATA_FLOAT_I;
    case 4: return SSP::COLDATA_FLOAT_RGB;
    case 5: return SSP::COLDATA_FLOAT_RGBA;
    case 6: return SSP::COLDATA_USHORT_RGBA;
    case 7: return SSP::COLDATA_DOUBLE_I;
    default: return SSP::COLDATA_NONE;
    }
}
}  // namespace

bool MMPLDSource::Open(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err_ = "cannot open " + path;
        return false;
    }
    file_.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    if (file_.size() < 60 || std::memcmp(file_.data(), "MMPLD\0", 6) != 0) {
        err_ = "MMPLD file header id wrong";
        return false;
    }
    size_t p = 6;
    uint32_t frames = 0;
    rd(file_, p, version_);
    rd(file_, p, frames);
    if (version_ < 100 || version_ > 103) {
        err_ = "MMPLD file header version wrong";
        return false;
    }
    if (frames == 0) {
        err_ = "MMPLD file does not contain any frame information";
        return false;
    }
    for (float& v : bbox_) rd(file_, p, v);
    p += 24;  // clip box
    seek_.resize(frames + 1);
    for (auto& s : seek_)
        if (!rd(file_, p, s)) {
            err_ = "Unable to read MMPLD file header";
            return false;
        }
    frames_ = frames;
    loaded_ = ~0u;
    ++hash_;
    return true;
}

bool MMPLDSource::load_frame(unsigned int idx) {
    if (idx == loaded_) return true;
    if (idx >= frames_ || seek_[idx] > seek_[idx + 1] || seek_[idx + 1] > file_.size()) {
        err_ = "frame index out of range or corrupt seek table";
        return false;
    }
    frame_.assign(file_.begin() + (ptrdiff_t)seek_[idx], file_.begin() + (ptrdiff_t)seek_[idx + 1]);
    size_t p = 0;
    if (version_ >= 102) p += 4;  // timestamp
    uint32_t plc = 0;
    if (!rd(frame_, p, plc)) return false;
    lists_.clear();
    for (uint32_t i = 0; i < plc; ++i) {
        uint8_t raw_vt = 0, raw_ct = 0;
        if (!rd(frame_, p, raw_vt) || !rd(frame_, p, raw_ct)) return false;
        const int vt = vert_size(raw_vt) ? raw_vt : 0;
        const int ct = (vt != 0 && col_size(raw_ct)) ? raw_ct : 0;
        List l{};
        l.vt = static_cast<SSP::VertexDataType>(vt);
        l.ct = col_type(ct);
        l.radius = 0.05f;
        l.rgba[0] = l.rgba[1] = l.rgba[2] = 192;
        l.rgba[3] = 255;
        l.cmin = 0.0f;
        l.cmax = 1.0f;
        if (raw_vt == 1 || raw_vt == 3 || raw_vt == 4)
            if (!rd(frame_, p, l.radius)) return false;
        if (raw_ct == 0) {
            uint8_t c[4];
            for (auto& v : c)
                if (!rd(frame_, p, v)) return false;
            l.rgba[0] = c[0];
            l.rgba[1] = c[1];
            l.rgba[2] = c[2];
            l.rgba[3] = 255;
        } else if (raw_ct == 3 || raw_ct == 7) {
            if (!rd(frame_, p, l.cmin) || !rd(frame_, p, l.cmax)) return false;
        }
        if (!rd(frame_, p, l.count)) return false;
        if (version_ >= 103) p += 24;  // list box
        l.stride = vert_size(vt) + col_size(ct);
        l.vert_off = p;
        l.col_off = p + vert_size(vt);
        if (p + (size_t)l.stride * l.count > frame_.size()) {
            err_ = "list data runs past the frame end";
            return false;
        }
        p += (size_t)l.stride * l.count;
        if (version_ == 101) {
            uint32_t ncl = 0;
            uint64_t sz = 0;
            if (!rd(frame_, p, ncl) || !rd(frame_, p, sz)) return false;
            p += sz;
        }
        lists_.push_back(l);
    }
    loaded_ = idx;
    return true;
}

bool MMPLDSource::Serve(geocalls::MultiParticleDataCall& call, int fn) {
    call.SetFrameCount(frames_);
    auto& bb = call.AccessBoundingBoxes();
    std::memcpy(bb.box, bbox_, sizeof bbox_);
    bb.valid = true;
    call.SetDataHash(hash_);
    if (fn == 1) return true;
    if (!load_frame(call.FrameID() < frames_ ? call.FrameID() : frames_ - 1)) return false;
    call.SetParticleListCount((unsigned int)lists_.size());
    for (size_t i = 0; i < lists_.size(); ++i) {
        const List& l = lists_[i];
        auto& pp = call.AccessParticles((unsigned int)i);
        pp.SetCount(l.count);
        pp.SetGlobalRadius(l.radius);
        pp.SetGlobalColour(l.rgba[0], l.rgba[1], l.rgba[2], l.rgba[3]);
        pp.SetColourMapIndexValues(l.cmin, l.cmax);
        pp.SetVertexData(l.vt, l.vt != SSP::VERTDATA_NONE ? frame_.data() + l.vert_off : nullptr, l.stride);
        pp.SetColourData(l.ct, l.ct != SSP::COLDATA_NONE ? frame_.data() + l.col_off : nullptr, l.stride);
    }
    return true;
}

}  // namespace megamol::mmhrt_plugin
