#pragma once
#include "mmstub_vec.h"
namespace megamol::core::view {
// core/include/mmcore/view/Camera.h: Pose + PerspectiveParameters (fovy in radians)
class Camera {
public:
    struct Pose { stub::vec3 position, direction, up, right; };
    struct ImagePlaneTile { stub::vec2 tile_start{0, 0}, tile_end{1, 1}; };
    struct PerspectiveParameters { float fovy = 0.7f, aspect = 1.0f, near_plane = 0.1f, far_plane = 100.0f; ImagePlaneTile image_plane_tile; };
    template <class T> T get() const;
    void setPose(const Pose& p) { pose_ = p; }
    void setPerspectiveProjection(const PerspectiveParameters& p) { persp_ = p; }
private:
    Pose pose_;
    PerspectiveParameters persp_;
};
template <> inline Camera::Pose Camera::get<Camera::Pose>() const { return pose_; }
template <> inline Camera::PerspectiveParameters Camera::get<Camera::PerspectiveParameters>() const { return persp_; }
}  // namespace megamol::core::view
