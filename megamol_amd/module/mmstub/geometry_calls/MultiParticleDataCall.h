#pragma once
#include <cstdint>
#include <functional>
#include <vector>
#include "mmcore/CallerSlot.h"
#include "mmstd/renderer/CallRender3D.h"
namespace megamol::geocalls {
// plugins/geometry_calls/include/geometry_calls/SimpleSphericalParticles.h (accessors used by renderers)
class SimpleSphericalParticles {
public:
    enum VertexDataType { VERTDATA_NONE = 0, VERTDATA_FLOAT_XYZ = 1, VERTDATA_FLOAT_XYZR = 2, VERTDATA_SHORT_XYZ = 3, VERTDATA_DOUBLE_XYZ = 4 };
    enum ColourDataType { COLDATA_NONE = 0, COLDATA_UINT8_RGB = 1, COLDATA_UINT8_RGBA = 2, COLDATA_FLOAT_RGB = 3, COLDATA_FLOAT_RGBA = 4, COLDATA_FLOAT_I = 5, COLDATA_USHORT_RGBA = 6, COLDATA_DOUBLE_I = 7 };
    void SetVertexData(VertexDataType t, const void* p, unsigned int stride = 0) { vt_ = t; vp_ = p; vs_ = stride; }
    void SetColourData(ColourDataType t, const void* p, unsigned int stride = 0) { ct_ = t; cp_ = p; cs_ = stride; }
    void SetCount(uint64_t n) { n_ = n; }
    void SetGlobalRadius(float r) { r_ = r; }
    void SetGlobalColour(unsigned r, unsigned g, unsigned b, unsigned a = 255) { c_[0] = r; c_[1] = g; c_[2] = b; c_[3] = a; }
    void SetColourMapIndexValues(float mn, float mx) { mn_ = mn; mx_ = mx; }
    VertexDataType GetVertexDataType() const { return vt_; }
    const void* GetVertexData() const { return vp_; }
    unsigned int GetVertexDataStride() const { return vs_; }
    ColourDataType GetColourDataType() const { return ct_; }
    const void* GetColourData() const { return cp_; }
    unsigned int GetColourDataStride() const { return cs_; }
    uint64_t GetCount() const { return n_; }
    float GetGlobalRadius() const { return r_; }
    const unsigned char* GetGlobalColour() const { return c_; }
    float GetMinColourIndexValue() const { return mn_; }
    float GetMaxColourIndexValue() const { return mx_; }
private:
    VertexDataType vt_ = VERTDATA_NONE; ColourDataType ct_ = COLDATA_NONE;
    const void* vp_ = nullptr; const void* cp_ = nullptr; unsigned int vs_ = 0, cs_ = 0;
    uint64_t n_ = 0; float r_ = 0.5f; unsigned char c_[4] = {192, 192, 192, 255}; float mn_ = 0, mx_ = 1;
};
// MultiParticleDataCall: functions 0 = GetData, 1 = GetExtent (AbstractGetData3DCall.h:159-164)
class MultiParticleDataCall : public core::Call {
public:
    using Particles = SimpleSphericalParticles;
    std::function<bool(MultiParticleDataCall&, int)> source;
    bool operator()(int fn) { return source ? source(*this, fn) : false; }
    void SetFrameID(unsigned int f, bool force = false) { frame_ = f; (void)force; }
    unsigned int FrameID() const { return frame_; }
    void SetDataHash(size_t h) { hash_ = h; }
    size_t DataHash() const { return hash_; }
    void SetFrameCount(unsigned int n) { frames_ = n; }
    unsigned int FrameCount() const { return frames_; }
    void SetParticleListCount(unsigned int n) { lists_.resize(n); }
    unsigned int GetParticleListCount() const { return (unsigned int)lists_.size(); }
    Particles& AccessParticles(unsigned int i) { return lists_[i]; }
    core::BoundingBoxes_2& AccessBoundingBoxes() { return bbox_; }
private:
    unsigned int frame_ = 0, frames_ = 1; size_t hash_ = 0;
    std::vector<Particles> lists_; core::BoundingBoxes_2 bbox_;
};
struct MultiParticleDataCallDescription {};
}  // namespace megamol::geocalls
