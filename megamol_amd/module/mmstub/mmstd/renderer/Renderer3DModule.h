#pragma once
#include <map>
#include <string>
#include "mmcore/CallerSlot.h"
#include "mmcore/param/ParamSlot.h"
#include "mmstd/renderer/CallRender3D.h"
namespace megamol::core::view {
// plugins/mmstd/include/mmstd/renderer/Renderer3DModule.h: create/release + the two callbacks;
// slots made available are found by name, as core::Module::FindSlot does for the module graph
class Renderer3DModule {
public:
    virtual ~Renderer3DModule() = default;
    bool Create() { return created_ = create(); }
    void Release() { if (created_) release(); created_ = false; }
    bool CallGetExtents(CallRender3D& c) { return GetExtents(c); }
    bool CallRender(CallRender3D& c) { return Render(c); }
    param::ParamSlot* FindParamSlot(const std::string& name) { auto i = params_.find(name); return i == params_.end() ? nullptr : i->second; }
    CallerSlot* FindCallerSlot(const std::string& name) { auto i = callers_.find(name); return i == callers_.end() ? nullptr : i->second; }
protected:
    virtual bool create() = 0;
    virtual void release() = 0;
    virtual bool GetExtents(CallRender3D& call) = 0;
    virtual bool Render(CallRender3D& call) = 0;
    void MakeSlotAvailable(param::ParamSlot* s) { params_[s->Name()] = s; }
    void MakeSlotAvailable(CallerSlot* s) { callers_[s->Name()] = s; }
private:
    bool created_ = false;
    std::map<std::string, param::ParamSlot*> params_;
    std::map<std::string, CallerSlot*> callers_;
};
}  // namespace megamol::core::view
