#pragma once
#include <memory>
#include "mmcore/CallerSlot.h"
#include "mmcore/view/CPUFramebuffer.h"
#include "mmcore/view/Camera.h"
#include "mmstub_vec.h"
namespace megamol::core {
struct BoundingBoxes_2 { float box[6] = {0, 0, 0, 0, 0, 0}; bool valid = false; void Clear() { valid = false; } };
namespace view {
// plugins/mmstd/include/mmstd/renderer/CallRender3D.h (BaseCallRender<CPUFramebuffer>)
class CallRender3D : public Call {
public:
    float Time() const { return time_; }
    void SetTime(float t) { time_ = t; }
    void SetTimeFramesCount(unsigned int n) { frames_ = n; }
    unsigned int TimeFramesCount() const { return frames_; }
    BoundingBoxes_2& AccessBoundingBoxes() { return bbox_; }
    Camera GetCamera() const { return cam_; }
    void SetCamera(const Camera& c) { cam_ = c; }
    stub::vec4 BackgroundColor() const { return bg_; }
    void SetBackgroundColor(stub::vec4 c) { bg_ = c; }
    std::shared_ptr<CPUFramebuffer> GetFramebuffer() { return fbo_; }
    void SetFramebuffer(std::shared_ptr<CPUFramebuffer> f) { fbo_ = f; }
private:
    float time_ = 0;
    unsigned int frames_ = 1;
    BoundingBoxes_2 bbox_;
    Camera cam_;
    stub::vec4 bg_{0, 0, 0, 1};
    std::shared_ptr<CPUFramebuffer> fbo_;
};
}  // namespace view
}  // namespace megamol::core
