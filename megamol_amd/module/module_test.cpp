// Drives HIPSphereRenderer the way a MegaMol view does (GetExtents, then Render into a CPU
// framebuffer) with an interleaved MMPLD-style particle list, and checks the framebuffer against
// a direct call of the C ABI with the same inputs. Exit status 0 = identical.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "HIPSphereRenderer.h"

using namespace megamol;

namespace {
class TestRenderer : public mmhrt_plugin::HIPSphereRenderer {
public:
    core::CallerSlot& in_slot() { return in_data_slot_; }
};
}  // namespace

int main() {
    const int n = 20000, W = 160, H = 120;
    // interleaved records: xyzr (16 B) + rgba float (16 B), like an MMPLD list (stride 32)
    std::vector<float> rec(static_cast<size_t>(n) * 8);
    std::mt19937 g(5);
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    for (int i = 0; i < n; ++i) {
        float* r = &rec[static_cast<size_t>(i) * 8];
        r[0] = u(g); r[1] = u(g); r[2] = u(g); r[3] = 0.005f + 0.02f * u(g);
        r[4] = u(g); r[5] = u(g); r[6] = u(g); r[7] = 1.0f;
    }
    geocalls::MultiParticleDataCall data;
    int data_calls = 0;
    data.source = [&](geocalls::MultiParticleDataCall& c, int fn) {
        if (fn == 1) {
            c.SetFrameCount(1);
            return true;
        }
        ++data_calls;
        c.SetParticleListCount(1);
        auto& p = c.AccessParticles(0);
        p.SetCount(n);
        p.SetVertexData(geocalls::SimpleSphericalParticles::VERTDATA_FLOAT_XYZR, rec.data(), 32);
        p.SetColourData(geocalls::SimpleSphericalParticles::COLDATA_FLOAT_RGBA, rec.data() + 4, 32);
        c.SetDataHash(42);
        return true;
    };
    TestRenderer mod;
    if (!mod.Create()) {
        std::fprintf(stderr, "create failed\n");
        return 2;
    }
    mod.in_slot().ConnectCall(&data);  // inside MegaMol the module graph connects the slot

    core::view::CallRender3D call;
    core::view::Camera cam;
    core::view::Camera::Pose pose;
    pose.position = {0.5f, 0.5f, 2.2f};
    pose.direction = {0.0f, 0.0f, -1.0f};
    pose.up = {0.0f, 1.0f, 0.0f};
    pose.right = {1.0f, 0.0f, 0.0f};
    core::view::Camera::PerspectiveParameters pp;
    pp.fovy = 0.6981317f;
    pp.aspect = float(W) / float(H);
    pp.near_plane = 0.1f;
    pp.far_plane = 10.0f;
    cam.setPose(pose);
    cam.setPerspectiveProjection(pp);
    call.SetCamera(cam);
    call.SetBackgroundColor({0.1f, 0.2f, 0.3f, 1.0f});
    auto fbo = std::make_shared<core::view::CPUFramebuffer>();
    fbo->width = W;
    fbo->height = H;
    call.SetFramebuffer(fbo);
    if (!mod.CallGetExtents(call) || call.TimeFramesCount() != 1) {
        std::fprintf(stderr, "GetExtents failed\n");
        return 3;
    }
    if (!mod.CallRender(call)) {
        std::fprintf(stderr, "Render failed\n");
        return 4;
    }
    // direct C ABI render of the same inputs
    mmhrt_ctx ctx;
    if (mmhrt_create(0, &ctx) != MMHRT_OK) return 5;
    mmhrt_sphere_list l{};
    l.vert_type = MMHRT_VERTDATA_FLOAT_XYZR;
    l.col_type = MMHRT_COLDATA_FLOAT_RGBA;
    l.vert = rec.data();
    l.col = rec.data() + 4;
    l.vert_stride = l.col_stride = 32;
    l.count = n;
    l.global_radius = 0.5f;
    l.global_rgba[0] = l.global_rgba[1] = l.global_rgba[2] = 192;
    l.global_rgba[3] = 255;
    l.col_max = 1.0f;
    if (mmhrt_set_spheres(ctx, &l, 1, 42, 0) != MMHRT_OK || mmhrt_build(ctx, MMHRT_BUILD_REBUILD, 30) != MMHRT_OK)
        return 6;
    mmhrt_camera c{};
    for (int k = 0; k < 3; ++k) {
        c.position[k] = pose.position[k];
        c.direction[k] = pose.direction[k];
        c.up[k] = pose.up[k];
    }
    c.fovy = pp.fovy;
    c.aspect = pp.aspect;
    c.near_plane = pp.near_plane;
    c.far_plane = pp.far_plane;
    c.tile_end[0] = c.tile_end[1] = 1.0f;
    mmhrt_render_params p{};
    p.width = W;
    p.height = H;
    p.spp = 1;
    p.mode = MMHRT_MODE_LAMBERT;
    p.ao_distance = 4.0f;
    p.accumulate = 1;
    p.background[0] = 0.1f;
    p.background[1] = 0.2f;
    p.background[2] = 0.3f;
    p.background[3] = 1.0f;
    std::vector<uint32_t> rgba(W * H);
    std::vector<float> depth(W * H);
    mmhrt_framebuffer out{};
    out.rgba8 = rgba.data();
    out.depth = depth.data();
    if (mmhrt_render(ctx, &c, &p, &out) != MMHRT_OK) return 7;
    size_t diff = 0, hits = 0;
    for (size_t i = 0; i < rgba.size(); ++i) {
        diff += fbo->colorBuffer[i] != rgba[i] || std::memcmp(&fbo->depthBuffer[i], &depth[i], 4) != 0;
        hits += depth[i] < 1.0f;
    }
    // a second Render with nothing changed accumulates (frame index 1) and reads the data again
    if (!mod.CallRender(call)) return 8;
    mmhrt_destroy(ctx);
    mod.Release();
    std::printf("module: %zu pixels differ, %zu hit pixels, %d data calls\n", diff, hits, data_calls);
    return (diff == 0 && hits > 1000 && data_calls == 2) ? 0 : 1;
}
