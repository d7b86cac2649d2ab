/*
 * headless — runs MegaMol-style screenshot projects without a GUI: an MMPLDDataSource feeding
 * HIPSphereRenderer through a MultiParticleDataCall, a view camera restored from a camstore
 * settings string, mmRenderNextFrame and mmScreenshot (the projects written by
 * utils/MMPLD/GenTestFiles.pl:119-206: View3DGL, MMPLDDataSource, a sphere renderer, the camera
 * at :199, 1280x720).
 *
 *   headless --camstore JSON [--mmpld FILE]... [--param NAME=VALUE]... [--bg R,G,B,A]
 *            [--script CMDS] --out PREFIX
 *
 * For every MMPLD file k in order, the script runs; its commands are separated by ';':
 *   render            GetExtents + Render into a CPUFramebuffer (mmRenderNextFrame), then write
 *                     PREFIX_k_f.rgba (width*height RGBA8) and PREFIX_k_f.depth (float) and, as a
 *                     screenshot (mmScreenshot), PREFIX_k_f.ppm
 *   param NAME=VALUE  mmSetParamValue on the renderer
 *   bg R,G,B,A        background colour of the view
 *   move DX,DY,DZ     moves the camera position
 *   rehash            the data source reports new data (a new data hash)
 * The default script is "render". For every render a line "frame k f index I camera ..." goes to
 * stdout: I is the progressive index the frame was blended at, and the camera is printed exactly
 * (hex floats) as the renderer received it. Exit status 0 when every step succeeded.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "HIPSphereRenderer.h"
#include "MMPLDSource.h"

using namespace megamol;

namespace {

// The numbers of a JSON member "key":[a,b,...] or "key":a in a flat settings object.
std::vector<double> json_numbers(const std::string& js, const std::string& key) {
    std::vector<double> out;
    size_t p = js.find("\"" + key + "\"");
    if (p == std::string::npos) return out;
    p = js.find(':', p);
    if (p == std::string::npos) return out;
    ++p;
    while (p < js.size() && (js[p] == ' ')) ++p;
    const bool arr = p < js.size() && js[p] == '[';
    if (arr) ++p;
    while (p < js.size()) {
        char* e = nullptr;
        double v = std::strtod(js.c_str() + p, &e);
        if (e == js.c_str() + p) break;
        out.push_back(v);
        p = (size_t)(e - js.c_str());
        while (p < js.size() && (js[p] == ' ' || js[p] == ',')) ++p;
        if (!arr || (p < js.size() && js[p] == ']')) break;
    }
    return out;
}

// v rotated by the unit quaternion (x, y, z, w): v + 2w (q x v) + 2 q x (q x v)
stub::vec3 rotate(const double q[4], double vx, double vy, double vz) {
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    const double tx = 2.0 * (y * vz - z * vy), ty = 2.0 * (z * vx - x * vz), tz = 2.0 * (x * vy - y * vx);
    stub::vec3 r;
    r.x = (float)(vx + w * tx + (y * tz - z * ty));
    r.y = (float)(vy + w * ty + (z * tx - x * tz));
    r.z = (float)(vz + w * tz + (x * ty - y * tx));
    return r;
}

std::vector<std::string> split(const std::string& s, char sep) {
    std::vector<std::string> out;
    size_t a = 0;
    while (a <= s.size()) {
        size_t b = s.find(sep, a);
        if (b == std::string::npos) b = s.size();
        std::string t = s.substr(a, b - a);
        while (!t.empty() && t.front() == ' ') t.erase(t.begin());
        while (!t.empty() && t.back() == ' ') t.pop_back();
        if (!t.empty()) out.push_back(t);
        a = b + 1;
    }
    return out;
}

bool write_file(const std::string& path, const void* p, size_t n) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fwrite(p, 1, n, f) == n;
    return std::fclose(f) == 0 && ok;
}

// screenshot: binary PPM, top row first (pixel row 0 is the bottom of the image)
bool write_ppm(const std::string& path, const std::vector<uint32_t>& rgba, unsigned w, unsigned h) {
    std::vector<uint8_t> b;
    std::string hdr = "P6\n" + std::to_string(w) + " " + std::to_string(h) + "\n255\n";
    b.assign(hdr.begin(), hdr.end());
    for (unsigned y = h; y-- > 0;)
        for (unsigned x = 0; x < w; ++x) {
            uint32_t c = rgba[(size_t)y * w + x];
            b.push_back((uint8_t)(c & 0xff));
            b.push_back((uint8_t)((c >> 8) & 0xff));
            b.push_back((uint8_t)((c >> 16) & 0xff));
        }
    return write_file(path, b.data(), b.size());
}

struct Renderer : public mmhrt_plugin::HIPSphereRenderer {};

}  // namespace

int main(int argc, char** argv) {
    std::string camstore, out = "headless", script = "render";
    std::vector<std::string> files, params;
    float bg[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    for (int i = 1; i + 1 < argc; i += 2) {
        std::string k = argv[i], v = argv[i + 1];
        if (k == "--camstore") camstore = v;
        else if (k == "--mmpld") files.push_back(v);
        else if (k == "--param") params.push_back(v);
        else if (k == "--out") out = v;
        else if (k == "--script") script = v;
        else if (k == "--bg") {
            auto c = split(v, ',');
            for (int q = 0; q < 4 && q < (int)c.size(); ++q) bg[q] = std::strtof(c[q].c_str(), nullptr);
        } else {
            std::fprintf(stderr, "unknown argument %s\n", k.c_str());
            return 2;
        }
    }
    if (files.empty() || camstore.empty()) {
        std::fprintf(stderr, "usage: headless --camstore JSON --mmpld FILE [...] --out PREFIX\n");
        return 2;
    }
    // the view camera from the camstore settings (position, orientation quaternion x y z w,
    // half aperture angle, clipping planes, resolution gate)
    auto pos = json_numbers(camstore, "position");
    auto ori = json_numbers(camstore, "orientation");
    auto ha = json_numbers(camstore, "half_aperture_angle");
    auto nearp = json_numbers(camstore, "near_clipping_plane");
    auto farp = json_numbers(camstore, "far_clipping_plane");
    auto gate = json_numbers(camstore, "resolution_gate");
    if (pos.size() != 3 || ori.size() != 4 || ha.size() != 1 || gate.size() != 2) {
        std::fprintf(stderr, "camstore settings lack position, orientation, half_aperture_angle or resolution_gate\n");
        return 2;
    }
    const unsigned W = (unsigned)gate[0], H = (unsigned)gate[1];
    const double q[4] = {ori[0], ori[1], ori[2], ori[3]};
    core::view::Camera::Pose pose;
    pose.position.x = (float)pos[0];
    pose.position.y = (float)pos[1];
    pose.position.z = (float)pos[2];
    pose.direction = rotate(q, 0.0, 0.0, -1.0);  // Camera.h:49-52
    pose.up = rotate(q, 0.0, 1.0, 0.0);
    pose.right = rotate(q, 1.0, 0.0, 0.0);
    core::view::Camera::PerspectiveParameters pp;
    pp.fovy = (float)(2.0 * ha[0]);
    pp.aspect = (float)W / (float)H;
    pp.near_plane = nearp.empty() ? 0.1f : (float)nearp[0];
    pp.far_plane = farp.empty() ? 100.0f : (float)farp[0];

    Renderer mod;
    if (!mod.Create()) {
        std::fprintf(stderr, "renderer create failed\n");
        return 3;
    }
    for (const auto& kv : params) {
        size_t e = kv.find('=');
        auto* slot = e == std::string::npos ? nullptr : mod.FindParamSlot(kv.substr(0, e));
        if (!slot || !slot->Parameter()->ParseValue(kv.substr(e + 1))) {
            std::fprintf(stderr, "cannot set parameter %s\n", kv.c_str());
            return 2;
        }
    }
    auto cmds = split(script, ';');
    int rc = 0;
    for (size_t k = 0; k < files.size() && rc == 0; ++k) {
        mmhrt_plugin::MMPLDSource src;
        if (!src.Open(files[k])) {
            std::fprintf(stderr, "%s: %s\n", files[k].c_str(), src.Error().c_str());
            rc = 4;
            break;
        }
        src.SetDataHash(1000 + k);
        geocalls::MultiParticleDataCall data;
        data.source = [&src](geocalls::MultiParticleDataCall& c, int fn) { return src.Serve(c, fn); };
        mod.FindCallerSlot("inData")->ConnectCall(&data);  // mmCreateCall(MultiParticleDataCall, ...)
        core::view::CallRender3D call;
        auto fbo = std::make_shared<core::view::CPUFramebuffer>();
        fbo->width = W;
        fbo->height = H;
        call.SetFramebuffer(fbo);
        core::view::Camera::Pose cur = pose;
        float cbg[4] = {bg[0], bg[1], bg[2], bg[3]};
        size_t hash = 1000 + k;
        int f = 0;
        for (const auto& cmd : cmds) {
            auto w = split(cmd, ' ');
            if (w.empty()) continue;
            if (w[0] == "render") {
                core::view::Camera cam;
                cam.setPose(cur);
                cam.setPerspectiveProjection(pp);
                call.SetCamera(cam);
                call.SetBackgroundColor({cbg[0], cbg[1], cbg[2], cbg[3]});
                if (!mod.CallGetExtents(call) || !mod.CallRender(call)) {
                    std::fprintf(stderr, "%s: render %d failed\n", files[k].c_str(), f);
                    rc = 5;
                    break;
                }
                // a Render that rebuilt or changed anything started over at index 0
                const uint32_t used = mod.ProgressiveFrameIndex() - 1;
                std::string base = out + "_" + std::to_string(k) + "_" + std::to_string(f);
                if (!write_file(base + ".rgba", fbo->colorBuffer.data(), fbo->colorBuffer.size() * 4) ||
                    !write_file(base + ".depth", fbo->depthBuffer.data(), fbo->depthBuffer.size() * 4) ||
                    !write_ppm(base + ".ppm", fbo->colorBuffer, W, H)) {
                    std::fprintf(stderr, "cannot write %s\n", base.c_str());
                    rc = 6;
                    break;
                }
                std::printf("frame %zu %d index %u width %u height %u camera %a %a %a %a %a %a %a %a %a %a %a %a %a "
                            "bg %a %a %a %a\n",
                            k, f, used, W, H, cur.position.x, cur.position.y, cur.position.z, cur.direction.x,
                            cur.direction.y, cur.direction.z, cur.up.x, cur.up.y, cur.up.z, pp.fovy, pp.aspect,
                            pp.near_plane, pp.far_plane, cbg[0], cbg[1], cbg[2], cbg[3]);
                ++f;
            } else if (w[0] == "param" && w.size() == 2) {
                size_t e = w[1].find('=');
                auto* slot = e == std::string::npos ? nullptr : mod.FindParamSlot(w[1].substr(0, e));
                if (!slot || !slot->Parameter()->ParseValue(w[1].substr(e + 1))) {
                    std::fprintf(stderr, "cannot set parameter %s\n", w[1].c_str());
                    rc = 2;
                    break;
                }
            } else if (w[0] == "bg" && w.size() == 2) {
                auto c = split(w[1], ',');
                for (int q2 = 0; q2 < 4 && q2 < (int)c.size(); ++q2) cbg[q2] = std::strtof(c[q2].c_str(), nullptr);
            } else if (w[0] == "move" && w.size() == 2) {
                auto c = split(w[1], ',');
                if (c.size() == 3) {
                    cur.position.x += std::strtof(c[0].c_str(), nullptr);
                    cur.position.y += std::strtof(c[1].c_str(), nullptr);
                    cur.position.z += std::strtof(c[2].c_str(), nullptr);
                }
            } else if (w[0] == "rehash") {
                src.SetDataHash(++hash + 1000000);
            } else {
                std::fprintf(stderr, "unknown script command %s\n", cmd.c_str());
                rc = 2;
                break;
            }
        }
        std::fflush(stdout);
    }
    mod.Release();
    return rc;
}
