/*
 * HIPSphereRenderer — see HIPSphereRenderer.h.
 *
 * Render follows AbstractRenderer::Render (plugins/optix_hpg/src/optix/AbstractRenderer.cpp:27-111):
 * fetch the data (SphereGeometry::get_data_cb, SphereGeometry.cpp:230-268), re-upload and rebuild
 * when the frame or data hash changed, reset the progressive frame index on any change of data,
 * viewport, camera, parameters or background, launch, advance the frame index.
 */
#include "HIPSphereRenderer.h"

#include <cmath>
#include <cstring>
#include <fstream>
#include <vector>

#include "mmcore/param/BoolParam.h"
#include "mmcore/param/EnumParam.h"
#include "mmcore/param/FilePathParam.h"
#include "mmcore/param/FloatParam.h"
#include "mmcore/param/IntParam.h"
#include "mmcore/utility/log/Log.h"

using namespace megamol;
using namespace megamol::mmhrt_plugin;
using core::utility::log::Log;

HIPSphereRenderer::HIPSphereRenderer()
        : in_data_slot_("inData", "particle data (MultiParticleDataCall)")
        , spp_slot_("spp", "samples per pixel")
        , max_bounces_slot_("max bounces", "diffuse bounces (mode Bounce)")
        , accumulate_slot_("accumulate", "progressive accumulation over frames")
        , mode_slot_("mode", "shading: Lambert, AO or diffuse bounces")
        , ao_distance_slot_("AOdistance", "ambient occlusion ray length")
        , morton_bits_slot_("mortonBits", "Morton code width of the LBVH (30 or 63)")
        , denoise_slot_("denoise", "run the learned denoiser on AO frames")
        , weights_slot_("denoiserWeights", "denoiser weights file") {
    in_data_slot_.SetCompatibleCall<geocalls::MultiParticleDataCallDescription>();
    MakeSlotAvailable(&in_data_slot_);
    spp_slot_ << new core::param::IntParam(1, 1);
    MakeSlotAvailable(&spp_slot_);
    max_bounces_slot_ << new core::param::IntParam(0, 0);
    MakeSlotAvailable(&max_bounces_slot_);
    accumulate_slot_ << new core::param::BoolParam(true);
    MakeSlotAvailable(&accumulate_slot_);
    auto* mode = new core::param::EnumParam(MMHRT_MODE_LAMBERT);
    mode->SetTypePair(MMHRT_MODE_LAMBERT, "Lambert");
    mode->SetTypePair(MMHRT_MODE_AO, "AO");
    mode->SetTypePair(MMHRT_MODE_BOUNCE, "Bounce");
    mode_slot_ << mode;
    MakeSlotAvailable(&mode_slot_);
    ao_distance_slot_ << new core::param::FloatParam(4.0f, 0.0f);
    MakeSlotAvailable(&ao_distance_slot_);
    morton_bits_slot_ << new core::param::IntParam(30, 30, 63);
    MakeSlotAvailable(&morton_bits_slot_);
    denoise_slot_ << new core::param::BoolParam(false);
    MakeSlotAvailable(&denoise_slot_);
    weights_slot_ << new core::param::FilePathParam("");
    MakeSlotAvailable(&weights_slot_);
}

HIPSphereRenderer::~HIPSphereRenderer() {
    this->Release();
}

bool HIPSphereRenderer::create() {
    if (mmhrt_create(0, &ctx_) != MMHRT_OK) {
        Log::DefaultLog.WriteError("[HIPSphereRenderer] %s", mmhrt_last_error(nullptr));
        return false;
    }
    return true;
}

void HIPSphereRenderer::release() {
    if (ctx_) mmhrt_destroy(ctx_);
    ctx_ = nullptr;
}

bool HIPSphereRenderer::GetExtents(core::view::CallRender3D& call) {
    // AbstractRenderer::GetExtents (AbstractRenderer.cpp:114-129)
    auto* in = in_data_slot_.CallAs<geocalls::MultiParticleDataCall>();
    if (in == nullptr) {
        call.SetTimeFramesCount(1);
        call.AccessBoundingBoxes().Clear();
        return true;
    }
    in->SetFrameID(static_cast<unsigned int>(call.Time()));
    if (!(*in)(1)) return false;
    call.SetTimeFramesCount(in->FrameCount());
    call.AccessBoundingBoxes() = in->AccessBoundingBoxes();
    return true;
}

// One list of SimpleSphericalParticles -> mmhrt_sphere_list (zero-copy strided views).
bool HIPSphereRenderer::upload(geocalls::MultiParticleDataCall& in) {
    const unsigned int n = in.GetParticleListCount();
    std::vector<mmhrt_sphere_list> lists(n);
    for (unsigned int i = 0; i < n; ++i) {
        auto& p = in.AccessParticles(i);
        mmhrt_sphere_list& l = lists[i];
        std::memset(&l, 0, sizeof l);
        l.vert_type = static_cast<int32_t>(p.GetVertexDataType());
        l.col_type = static_cast<int32_t>(p.GetColourDataType());
        l.vert = p.GetVertexData();
        l.col = p.GetColourData();
        l.vert_stride = p.GetVertexDataStride();
        l.col_stride = p.GetColourDataStride();
        l.count = p.GetCount();
        l.global_radius = p.GetGlobalRadius();
        std::memcpy(l.global_rgba, p.GetGlobalColour(), 4);
        l.col_min = p.GetMinColourIndexValue();
        l.col_max = p.GetMaxColourIndexValue();
    }
    if (mmhrt_set_spheres(ctx_, lists.data(), n, in.DataHash(), in.FrameID()) != MMHRT_OK) {
        Log::DefaultLog.WriteError("[HIPSphereRenderer] %s", mmhrt_last_error(ctx_));
        return false;
    }
    int bits = morton_bits_slot_.Param<core::param::IntParam>()->Value() >= 63 ? 63 : 30;
    if (mmhrt_build(ctx_, MMHRT_BUILD_REBUILD, bits) != MMHRT_OK) {
        Log::DefaultLog.WriteError("[HIPSphereRenderer] %s", mmhrt_last_error(ctx_));
        return false;
    }
    return true;
}

bool HIPSphereRenderer::Render(core::view::CallRender3D& call) {
    if (ctx_ == nullptr) return false;
    auto* in = in_data_slot_.CallAs<geocalls::MultiParticleDataCall>();
    if (in == nullptr) return false;
    // SphereGeometry::get_data_cb: frame id, extents, data
    in->SetFrameID(static_cast<unsigned int>(call.Time()));
    if (!(*in)(1)) return false;
    if (!(*in)(0)) return false;
    bool changed = false;
    if (in->FrameID() != frame_id_ || in->DataHash() != data_hash_ || !has_data_) {
        if (!upload(*in)) return false;
        frame_id_ = in->FrameID();
        data_hash_ = in->DataHash();
        has_data_ = true;
        changed = true;
    }
    auto fbo = call.GetFramebuffer();
    if (!fbo || fbo->width == 0 || fbo->height == 0) return false;  // OSPRayRenderer.cpp:166-167
    core::view::Camera cam = call.GetCamera();
    auto pose = cam.get<core::view::Camera::Pose>();
    auto intr = cam.get<core::view::Camera::PerspectiveParameters>();
    mmhrt_camera c{};
    for (int k = 0; k < 3; ++k) {
        c.position[k] = pose.position[k];
        c.direction[k] = pose.direction[k];
        c.up[k] = pose.up[k];
    }
    c.fovy = intr.fovy;
    c.aspect = intr.aspect;
    c.near_plane = intr.near_plane;
    c.far_plane = intr.far_plane;
    c.tile_start[0] = intr.image_plane_tile.tile_start.x;
    c.tile_start[1] = intr.image_plane_tile.tile_start.y;
    c.tile_end[0] = intr.image_plane_tile.tile_end.x;
    c.tile_end[1] = intr.image_plane_tile.tile_end.y;
    mmhrt_render_params p{};
    p.width = fbo->width;
    p.height = fbo->height;
    p.spp = static_cast<uint32_t>(spp_slot_.Param<core::param::IntParam>()->Value());
    p.mode = mode_slot_.Param<core::param::EnumParam>()->Value();
    p.max_bounces = max_bounces_slot_.Param<core::param::IntParam>()->Value();
    p.ao_distance = ao_distance_slot_.Param<core::param::FloatParam>()->Value();
    // The reference blends every frame with frameIdx > 0 into the previous one whatever its
    // accumulate parameter says (raygen.cu:254-260 never reads FrameState::accumulate, which
    // Renderer.cpp:162 stores): the module does the same. The parameter stays for project files
    // and, as in the reference, a change of it restarts the accumulation.
    p.accumulate = 1;
    const bool accumulate_param = accumulate_slot_.Param<core::param::BoolParam>()->Value();
    auto bg = call.BackgroundColor();
    p.background[0] = bg.r;
    p.background[1] = bg.g;
    p.background[2] = bg.b;
    p.background[3] = bg.a;
    const bool denoise = denoise_slot_.Param<core::param::BoolParam>()->Value() && p.mode == MMHRT_MODE_AO;
    if (denoise) p.flags |= MMHRT_FLAG_AUX;
    // change detection resets the progressive accumulation (Renderer.cpp:93-176)
    p.frame_idx = 0;
    if (changed || std::memcmp(&c, &last_cam_, sizeof c) != 0 || std::memcmp(&p, &last_params_, sizeof p) != 0 ||
        accumulate_param != last_accumulate_param_)
        frame_idx_ = 0;
    last_cam_ = c;
    last_params_ = p;
    last_accumulate_param_ = accumulate_param;
    p.frame_idx = frame_idx_;
    if (mmhrt_render(ctx_, &c, &p, nullptr) != MMHRT_OK) {  // out == NULL: the frame stays on the device
        Log::DefaultLog.WriteError("[HIPSphereRenderer] %s", mmhrt_last_error(ctx_));
        return false;
    }
    if (denoise) {
        std::string w = weights_slot_.Param<core::param::FilePathParam>()->Value().string();
        if (w != loaded_weights_) {
            // weights file: see megamol_amd/denoiser.py (magic, version, layer table, meta, fp32)
            std::ifstream f(w, std::ios::binary);
            std::vector<char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
            uint64_t np = 0;
            mmhrt_denoiser_param_count(ctx_, &np);
            if (buf.size() < 16 || std::memcmp(buf.data(), "MMHRTDN1", 8) != 0 || buf.size() < np * 4) {
                Log::DefaultLog.WriteError("[HIPSphereRenderer] cannot read denoiser weights %s", w.c_str());
                return false;
            }
            const float* params = reinterpret_cast<const float*>(buf.data() + buf.size() - np * 4);
            if (mmhrt_denoiser_set_params(ctx_, params, np) != MMHRT_OK) return false;
            loaded_weights_ = w;
        }
        float dist = std::sqrt(c.position[0] * c.position[0] + c.position[1] * c.position[1] +
                               c.position[2] * c.position[2]);
        float bounds[6];
        if (mmhrt_get_info(ctx_, nullptr, bounds, nullptr) == MMHRT_OK) {
            float mid[3] = {0.5f * (bounds[0] + bounds[3]), 0.5f * (bounds[1] + bounds[4]),
                            0.5f * (bounds[2] + bounds[5])};
            float d2 = 0.0f;
            for (int k = 0; k < 3; ++k) d2 += (c.position[k] - mid[k]) * (c.position[k] - mid[k]);
            dist = std::sqrt(d2);
        }
        if (mmhrt_denoise(ctx_, dist) != MMHRT_OK) {
            Log::DefaultLog.WriteError("[HIPSphereRenderer] %s", mmhrt_last_error(ctx_));
            return false;
        }
    }
    // CPUFramebuffer: RGBA8 colour + depth, as OSPRayRenderer fills it (OSPRayRenderer.cpp:340-345)
    fbo->colorBuffer.resize(static_cast<size_t>(p.width) * p.height);
    fbo->depthBuffer.resize(static_cast<size_t>(p.width) * p.height);
    mmhrt_framebuffer out{};
    out.rgba8 = fbo->colorBuffer.data();
    out.depth = fbo->depthBuffer.data();
    if (mmhrt_download(ctx_, &out) != MMHRT_OK) {
        Log::DefaultLog.WriteError("[HIPSphereRenderer] %s", mmhrt_last_error(ctx_));
        return false;
    }
    fbo->depthBufferActive = true;
    ++frame_idx_;  // AbstractRenderer.cpp:108
    return true;
}
