/*
 * HIPSphereRenderer — MegaMol renderer module for particle spheres on MI355X.
 *
 * Drop-in for the optix_hpg sphere path (plugins/optix_hpg/src/optix/SphereGeometry.h:22-82 +
 * AbstractRenderer.h/Renderer.h): consumes geocalls::MultiParticleDataCall through the caller
 * slot "inData" and fills the CallRender3D CPU framebuffer (RGBA8 colour + depth) like the
 * mmospray renderer does (plugins/mmospray/src/OSPRayRenderer.cpp:340-345). All work runs in
 * libmmhrt (include/mmhrt.h).
 *
 * Parameters mirror optix_hpg Renderer (Renderer.cpp:22-39: spp, max bounces, accumulate,
 * intensity) and the OSPRay scivis AO knobs (AbstractOSPRayRenderer.cpp:370-377).
 */
#pragma once

#include <memory>
#include <string>

#include "geometry_calls/MultiParticleDataCall.h"
#include "mmcore/CallerSlot.h"
#include "mmcore/param/ParamSlot.h"
#include "mmstd/renderer/CallRender3D.h"
#include "mmstd/renderer/Renderer3DModule.h"
#include "mmhrt.h"

namespace megamol::mmhrt_plugin {

class HIPSphereRenderer : public core::view::Renderer3DModule {
public:
    static const char* ClassName() { return "HIPSphereRenderer"; }
    static const char* Description() { return "Particle-sphere ray tracer on MI355X (HIP): Lambert, AO, bounces"; }
    static bool IsAvailable() { return true; }

    HIPSphereRenderer();
    ~HIPSphereRenderer() override;

    // Index the next Render blends at (0: a fresh frame); AbstractRenderer.cpp:108 advances it.
    uint32_t ProgressiveFrameIndex() const { return frame_idx_; }

protected:
    bool create() override;
    void release() override;
    bool GetExtents(core::view::CallRender3D& call) override;
    bool Render(core::view::CallRender3D& call) override;

    core::CallerSlot in_data_slot_;  // connected by the module graph

private:
    bool upload(geocalls::MultiParticleDataCall& in);

    core::param::ParamSlot spp_slot_, max_bounces_slot_, accumulate_slot_, mode_slot_, ao_distance_slot_,
        morton_bits_slot_, denoise_slot_, weights_slot_;

    mmhrt_ctx ctx_ = nullptr;
    unsigned int frame_id_ = ~0u;
    size_t data_hash_ = ~size_t(0);
    uint32_t frame_idx_ = 0;  // progressive accumulation index (AbstractRenderer.cpp:108)
    mmhrt_camera last_cam_{};
    mmhrt_render_params last_params_{};
    bool last_accumulate_param_ = true;
    std::string loaded_weights_;
    bool has_data_ = false;
};

}  // namespace megamol::mmhrt_plugin
