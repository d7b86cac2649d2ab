#pragma once
#include "mmstd/renderer/CallRender3D.h"
namespace megamol::core::view {
// plugins/mmstd/include/mmstd/renderer/Renderer3DModule.h: create/release + the two callbacks
class Renderer3DModule {
public:
    virtual ~Renderer3DModule() = default;
    bool Create() { return created_ = create(); }
    void Release() { if (created_) release(); created_ = false; }
    bool CallGetExtents(CallRender3D& c) { return GetExtents(c); }
    bool CallRender(CallRender3D& c) { return Render(c); }
protected:
    virtual bool create() = 0;
    virtual void release() = 0;
    virtual bool GetExtents(CallRender3D& call) = 0;
    virtual bool Render(CallRender3D& call) = 0;
    template <class S> void MakeSlotAvailable(S*) {}
private:
    bool created_ = false;
};
}  // namespace megamol::core::view
