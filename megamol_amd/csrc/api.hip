// api.hip — the extern "C" boundary declared in include/mmhrt.h.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "../../include/mmhrt.h"
#include "ctx.hpp"
#include "denoise.hpp"

namespace mmhrt {

int fail(mmhrt_ctx_s* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}
int hip_fail(mmhrt_ctx_s* c, hipError_t e, const char* what) {
    return fail(c, e == hipErrorOutOfMemory ? MMHRT_E_OOM : MMHRT_E_HIP,
                std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
}

namespace {
// SimpleSphericalParticles.cpp:7-9
const uint32_t kVertSize[5] = {0, 12, 16, 6, 24};
const uint32_t kColSize[8] = {0, 3, 4, 12, 16, 4, 8, 8};

thread_local std::string g_create_error;

float elapsed(mmhrt_ctx_s* c, int a, int b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c->tm.ev[a], c->tm.ev[b]) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0f;
    }
    return ms;
}

}  // namespace

// Camera -> per-frame constants: Renderer.cpp:136-172 (near plane of the ray model fixed at 100)
// and the glm::frustum entries (2,2) and (3,2) that Renderer.cpp:141-146 extracts.
void frame_consts(const mmhrt_camera* cam, const mmhrt_render_params* p, FrameConsts& f) {
    memset(&f, 0, sizeof f);
    for (int i = 0; i < 3; ++i) {
        f.center[i] = cam->position[i];
        f.front[i] = cam->direction[i];
        f.up[i] = cam->up[i];
    }
    const float* d = cam->direction;
    const float* u = cam->up;
    f.right[0] = d[1] * u[2] - u[1] * d[2];
    f.right[1] = d[2] * u[0] - u[2] * d[0];
    f.right[2] = d[0] * u[1] - u[0] * d[1];
    float hfov = 0.5f * cam->fovy;
    f.th = tanf(hfov) * 100.0f;
    f.rw = f.th * cam->aspect;
    f.nearp = 100.0f;
    float n = cam->near_plane, fa = cam->far_plane;
    f.depth_a = -(fa + n) / (fa - n);
    f.depth_c = -(2.0f * fa * n) / (fa - n);
    f.pad = std::max(std::max(fabsf(cam->position[0]), fabsf(cam->position[1])), fabsf(cam->position[2]));
    f.ao_distance = p->ao_distance;
    for (int i = 0; i < 4; ++i) f.bg[i] = p->background[i];
    f.width = p->width;
    f.height = p->height;
    f.spp = p->spp;
    f.frame_idx = p->frame_idx;
    f.mode = p->mode;
    f.accumulate = p->accumulate;
    f.want_aux = (p->flags & MMHRT_FLAG_AUX) ? 1 : 0;
    f.max_bounces = p->max_bounces;
    if (p->window[2] > 0 && p->window[3] > 0) {
        f.win_x0 = p->window[0];
        f.win_y0 = p->window[1];
        f.win_w = p->window[2];
        f.win_h = p->window[3];
    } else {
        f.win_x0 = f.win_y0 = 0;
        f.win_w = p->width;
        f.win_h = p->height;
    }
}

// Argument and state checks shared by the render entry points.
int check_render_args(mmhrt_ctx_s* c, const mmhrt_camera* cam, const mmhrt_render_params* p, const char* who) {
    const std::string w(who);
    if (!cam || !p) return fail(c, MMHRT_E_INVALID, w + ": camera or params is NULL");
    if (p->width == 0 || p->height == 0) return fail(c, MMHRT_E_INVALID, w + ": viewport is 0x0");
    if ((uint64_t)p->width * p->height > (1ull << 31)) return fail(c, MMHRT_E_UNSUPPORTED, w + ": viewport too large");
    if (p->spp == 0) return fail(c, MMHRT_E_INVALID, w + ": spp must be >= 1");
    if (p->mode < 0 || p->mode > 2) return fail(c, MMHRT_E_INVALID, w + ": unknown mode");
    if (!c->built || c->built_n != c->n) return fail(c, MMHRT_E_STATE, w + ": build the LBVH first");
    if (!(cam->fovy > 0.0f) || !(cam->aspect > 0.0f)) return fail(c, MMHRT_E_INVALID, w + ": bad fovy or aspect");
    if (p->window[2] > 0 && ((uint64_t)p->window[0] + p->window[2] > p->width ||
                             (uint64_t)p->window[1] + p->window[3] > p->height || p->window[3] == 0))
        return fail(c, MMHRT_E_INVALID, w + ": window outside the frame");
    return MMHRT_OK;
}

int check_ctx(mmhrt_ctx c) {
    if (!c) return MMHRT_E_INVALID;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    return MMHRT_OK;
}
}  // namespace mmhrt

using namespace mmhrt;

extern "C" {

int32_t mmhrt_create(int32_t device, mmhrt_ctx* out) {
    g_create_error.clear();
    if (!out) {
        g_create_error = "mmhrt_create: out is NULL";
        return MMHRT_E_INVALID;
    }
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        g_create_error = "mmhrt_create: no HIP device available";
        return MMHRT_E_HIP;
    }
    if (device < 0 || device >= ndev) {
        g_create_error = "mmhrt_create: device index out of range";
        return MMHRT_E_INVALID;
    }
    mmhrt_ctx_s* c = new (std::nothrow) mmhrt_ctx_s();
    if (!c) {
        g_create_error = "mmhrt_create: host allocation failed";
        return MMHRT_E_OOM;
    }
    c->device = device;
    if ((e = hipSetDevice(device)) != hipSuccess || (e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
        g_create_error = std::string("mmhrt_create: ") + hipGetErrorString(e);
        delete c;
        return MMHRT_E_HIP;
    }
    hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, device);
    if (c->cu_count <= 0) c->cu_count = 1;
    for (int i = 0; i < 12; ++i) hipEventCreate(&c->tm.ev[i]);
    c->tm.init = true;
    // the device address of this device's scratch array for mmhrt_render_noalloc; on failure that
    // entry point reports the reason and everything else works as before
    if (noalloc_init(c) != 0) c->noalloc_err = c->err;
    c->err.clear();
    // pinned words through which a build tells the next one its straddler rounds (all ones: none yet)
    if (hipHostMalloc((void**)&c->rounds_host, BOX_SLOTS * sizeof(uint32_t), hipHostMallocDefault) == hipSuccess)
        memset(c->rounds_host, 0xff, BOX_SLOTS * sizeof(uint32_t));
    else {
        (void)hipGetLastError();
        c->rounds_host = nullptr;
    }
    *out = c;
    return MMHRT_OK;
}

int32_t mmhrt_destroy(mmhrt_ctx c) {
    if (!c) return MMHRT_E_INVALID;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    release(c->sph); release(c->col); release(c->staging);
    release(c->meet); release(c->meet_far); release(c->ka); release(c->kb); release(c->kla); release(c->klb);
    release(c->va); release(c->vb); release(c->hist); release(c->scan_tmp); release(c->sort_fault);
    release(c->pair_count); release(c->xlist);
    release(c->leaf_sph); release(c->leaf_id); release(c->nodes); release(c->flags); release(c->flags2); release(c->box_cnt);
    release(c->partial); release(c->info); release(c->tf_buf);
    release(c->fb_col); release(c->fb_depth); release(c->fb_t); release(c->fb_id); release(c->fb_rgba8);
    release(c->fb_aux); release(c->spill); release(c->counters);
    release(c->ts_sph); release(c->ts_col); release(c->bt_col); release(c->bt_depth); release(c->bt_t);
    release(c->bt_id); release(c->bt_rgba8); release(c->batch_ctr); release(c->batch_prefix);
    release(c->batch_jobs);
    for (auto& sl : c->ts_slots) {
        release(sl.nodes); release(sl.leaf); release(sl.id); release(sl.info); release(sl.dim); release(sl.box);
        release(sl.rmax);
    }
    c->ts_slots.clear();
    if (c->rounds_host) (void)hipHostFree(c->rounds_host);
    c->rounds_host = nullptr;
    if (c->dn) {
        c->dn->release_all();
        delete c->dn;
        c->dn = nullptr;
    }
    if (c->tm.init)
        for (int i = 0; i < 12; ++i) hipEventDestroy(c->tm.ev[i]);
    hipStreamDestroy(c->stream);
    delete c;
    return MMHRT_OK;
}

const char* mmhrt_last_error(mmhrt_ctx c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int32_t mmhrt_set_transfer_function(mmhrt_ctx c, const float* rgba, uint32_t n, float vmin, float vmax) {
    if (int e = check_ctx(c)) return e;
    if (n > 0 && !rgba) return fail(c, MMHRT_E_INVALID, "set_transfer_function: rgba is NULL");
    c->tf.assign(rgba, rgba + (size_t)n * 4);
    c->tf_min = vmin;
    c->tf_max = vmax;
    return MMHRT_OK;
}

int32_t mmhrt_set_spheres(mmhrt_ctx c, const mmhrt_sphere_list* lists, uint32_t n_lists, uint64_t data_hash,
                          uint32_t frame_id) {
    if (int e = check_ctx(c)) return e;
    if (n_lists > 0 && !lists) return fail(c, MMHRT_E_INVALID, "set_spheres: lists is NULL");
    // validate and size
    uint64_t total = 0, staging = 0;
    struct Span {
        const uint8_t* hv;
        uint64_t vbytes;
        const uint8_t* hc;
        uint64_t cbytes;
        bool col_in_vert;
        uint64_t vs, cs;
    };
    std::vector<Span> spans(n_lists);
    for (uint32_t l = 0; l < n_lists; ++l) {
        const mmhrt_sphere_list& L = lists[l];
        Span& s = spans[l];
        memset(&s, 0, sizeof s);
        if (L.count == 0 || L.vert_type == MMHRT_VERTDATA_NONE) continue;  // SphereGeometry.cpp:98-100
        if (L.vert_type < 0 || L.vert_type > 4)
            return fail(c, MMHRT_E_INVALID, "set_spheres: list " + std::to_string(l) + " has an unknown vertex type");
        if (L.col_type < 0 || L.col_type > 7)
            return fail(c, MMHRT_E_INVALID, "set_spheres: list " + std::to_string(l) + " has an unknown colour type");
        if (!L.vert) return fail(c, MMHRT_E_INVALID, "set_spheres: list " + std::to_string(l) + " vertex pointer is NULL");
        if (L.col_type != MMHRT_COLDATA_NONE && !L.col)
            return fail(c, MMHRT_E_INVALID, "set_spheres: list " + std::to_string(l) + " colour pointer is NULL");
        s.vs = L.vert_stride ? L.vert_stride : kVertSize[L.vert_type];
        s.cs = L.col_stride ? L.col_stride : kColSize[L.col_type];
        if (s.vs < kVertSize[L.vert_type] || (L.col_type != 0 && s.cs < kColSize[L.col_type]))
            return fail(c, MMHRT_E_INVALID, "set_spheres: list " + std::to_string(l) + " stride smaller than element");
        s.hv = (const uint8_t*)L.vert;
        s.vbytes = (L.count - 1) * s.vs + kVertSize[L.vert_type];
        if (L.col_type != MMHRT_COLDATA_NONE) {
            s.hc = (const uint8_t*)L.col;
            s.cbytes = (L.count - 1) * s.cs + kColSize[L.col_type];
            // interleaved record (MMPLD layout): colour bytes lie inside the vertex span
            s.col_in_vert = s.hc >= s.hv && s.hc + s.cbytes <= s.hv + s.vbytes + (s.vs - kVertSize[L.vert_type]) &&
                            s.cs == s.vs;
            if (s.col_in_vert) s.vbytes = std::max(s.vbytes, (uint64_t)(s.hc - s.hv) + s.cbytes);
        }
        total += L.count;
        staging += ((s.vbytes + 255) & ~255ull) + (s.col_in_vert ? 0 : ((s.cbytes + 255) & ~255ull));
    }
    if (total > 0x7fffffffull) return fail(c, MMHRT_E_UNSUPPORTED, "set_spheres: more than 2^31-1 spheres");
    int e = 0;
    if ((e = ensure(c, c->sph, total, "spheres"))) return e;
    if ((e = ensure(c, c->col, total, "colours"))) return e;
    if ((e = ensure(c, c->staging, staging + 256, "staging"))) return e;
    // transfer function (AddParticleColors.cpp:59-74)
    float tf_vmin = c->tf_min, tf_vmax = c->tf_max;
    if (tf_vmin > tf_vmax) {
        tf_vmin = INFINITY;
        tf_vmax = -INFINITY;
        for (uint32_t l = 0; l < n_lists; ++l) {
            tf_vmin = std::min(tf_vmin, lists[l].col_min);
            tf_vmax = std::max(tf_vmax, lists[l].col_max);
        }
    }
    float tf_fac = 1.0f / (tf_vmax - tf_vmin + 1e-8f);
    uint32_t tf_n = (uint32_t)(c->tf.size() / 4);
    float* tf_dev = nullptr;
    if (tf_n) {
        if ((e = ensure(c, c->tf_buf, (size_t)tf_n * 4, "transfer function"))) return e;
        tf_dev = c->tf_buf.p;
        hipError_t he = hipMemcpyAsync(tf_dev, c->tf.data(), (size_t)tf_n * 16, hipMemcpyHostToDevice, c->stream);
        if (he != hipSuccess) return hip_fail(c, he, "transfer function upload");
    }
    uint64_t off = 0, sp = 0;
    for (uint32_t l = 0; l < n_lists; ++l) {
        const mmhrt_sphere_list& L = lists[l];
        const Span& s = spans[l];
        if (L.count == 0 || L.vert_type == MMHRT_VERTDATA_NONE) continue;
        uint8_t* dv = c->staging.p + sp;
        hipError_t he = hipMemcpyAsync(dv, s.hv, s.vbytes, hipMemcpyHostToDevice, c->stream);
        if (he != hipSuccess) return hip_fail(c, he, "particle upload");
        sp += (s.vbytes + 255) & ~255ull;
        uint8_t* dc = dv;
        if (L.col_type != MMHRT_COLDATA_NONE) {
            if (s.col_in_vert) {
                dc = dv + (s.hc - s.hv);
            } else {
                dc = c->staging.p + sp;
                he = hipMemcpyAsync(dc, s.hc, s.cbytes, hipMemcpyHostToDevice, c->stream);
                if (he != hipSuccess) return hip_fail(c, he, "colour upload");
                sp += (s.cbytes + 255) & ~255ull;
            }
        }
        ListDesc d;
        d.vert = dv;
        d.col = dc;
        d.vstride = s.vs;
        d.cstride = s.cs;
        d.count = L.count;
        d.out_offset = off;
        d.vtype = L.vert_type;
        d.ctype = L.col_type;
        d.gradius = L.global_radius;
        for (int q = 0; q < 4; ++q) d.gcol[q] = (float)L.global_rgba[q] / 255.0f;  // SphereGeometry.cpp:179-181
        d.use_tf = tf_n > 0 && (L.col_type == MMHRT_COLDATA_FLOAT_I || L.col_type == MMHRT_COLDATA_DOUBLE_I);
        if ((e = launch_convert_list(c, d, tf_dev, tf_n, tf_vmin, tf_fac))) return e;
        off += L.count;
    }
    hipError_t he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) return hip_fail(c, he, "set_spheres");
    c->n = total;
    c->data_hash = data_hash;
    c->frame_id = frame_id;
    c->info_valid = false;
    return MMHRT_OK;
}

int32_t mmhrt_set_spheres_device(mmhrt_ctx c, const void* xyzr_dev, const void* rgba_dev, uint64_t count,
                                 const uint8_t global_rgba[4]) {
    if (int e = check_ctx(c)) return e;
    if (count > 0 && !xyzr_dev) return fail(c, MMHRT_E_INVALID, "set_spheres_device: xyzr is NULL");
    if (count > 0x7fffffffull) return fail(c, MMHRT_E_UNSUPPORTED, "set_spheres_device: more than 2^31-1 spheres");
    int e = 0;
    if ((e = ensure(c, c->sph, count, "spheres"))) return e;
    if ((e = ensure(c, c->col, count, "colours"))) return e;
    float4 g = make_float4(1.f, 1.f, 1.f, 1.f);
    if (global_rgba)
        g = make_float4(global_rgba[0] / 255.0f, global_rgba[1] / 255.0f, global_rgba[2] / 255.0f,
                        global_rgba[3] / 255.0f);
    if ((e = launch_copy_packed(c, (const float4*)xyzr_dev, (const float4*)rgba_dev, count, g))) return e;
    c->n = count;
    c->info_valid = false;
    return MMHRT_OK;
}

int32_t mmhrt_build(mmhrt_ctx c, int32_t policy, int32_t morton_bits) {
    if (int e = check_ctx(c)) return e;
    if (morton_bits != 30 && morton_bits != 63) return fail(c, MMHRT_E_INVALID, "build: morton_bits must be 30 or 63");
    if (policy < 0 || policy > 2) return fail(c, MMHRT_E_INVALID, "build: unknown policy");
    if (c->n == 0) return fail(c, MMHRT_E_STATE, "build: no spheres set");
    if (c->opt_accel == 1) {
        // treelet P-k-d trees keep no topology that moved spheres could refit: every policy builds
        // (the reference's treelets renderer rebuilds on every data change, TreeletsRenderer.cpp:95-110)
        c->refits_since_build = 0;
        return build_pkd(c, morton_bits);
    }
    if (c->opt_accel != 0) return fail(c, MMHRT_E_INVALID, "build: option accel must be 0 (LBVH) or 1 (treelets)");
    bool can_refit = c->built && c->accel == 0 && c->built_n == c->n && c->bits == morton_bits;
    if (policy == MMHRT_BUILD_REFIT) {
        if (!can_refit) return fail(c, MMHRT_E_STATE, "build: refit needs an LBVH over the same sphere count");
        return refit_lbvh(c);
    }
    if (policy == MMHRT_BUILD_AUTO && can_refit && c->refits_since_build < c->opt_refit_period) {
        c->refits_since_build++;
        return refit_lbvh(c);
    }
    c->refits_since_build = 0;
    return build_lbvh(c, morton_bits);
}

int32_t mmhrt_render(mmhrt_ctx c, const mmhrt_camera* cam, const mmhrt_render_params* p, mmhrt_framebuffer* out) {
    if (int e = check_ctx(c)) return e;
    if (int e = check_render_args(c, cam, p, "render")) return e;
    const bool win = p->window[2] > 0;
    uint32_t ow = win ? p->window[2] : p->width, oh = win ? p->window[3] : p->height;
    uint32_t ox = win ? p->window[0] : 0, oy = win ? p->window[1] : 0;
    // blending reads the stored frame: it must be the same window of a frame of the same size
    if (p->accumulate && p->frame_idx > 0 &&
        (c->fb_w != ow || c->fb_h != oh || c->fb_x0 != ox || c->fb_y0 != oy || c->fb_full_w != p->width ||
         c->fb_full_h != p->height))
        return fail(c, MMHRT_E_STATE, "render: accumulation needs the previous frame of the same window and size");
    FrameConsts f;
    frame_consts(cam, p, f);
    if (int e = render_frame(c, f)) return e;
    c->last_aux = (p->flags & MMHRT_FLAG_AUX) != 0;
    for (int i = 0; i < 4; ++i) c->last_bg[i] = p->background[i];
    if (out && !(p->flags & MMHRT_FLAG_NO_DOWNLOAD)) return mmhrt_download(c, out);
    return MMHRT_OK;
}

int32_t mmhrt_render_noalloc(mmhrt_ctx c, const mmhrt_camera* cam, const mmhrt_render_params* p,
                             const mmhrt_device_framebuffer* out) {
    if (int e = check_ctx(c)) return e;
    if (int e = check_render_args(c, cam, p, "render_noalloc")) return e;
    if (!c->noalloc_err.empty()) return fail(c, MMHRT_E_STATE, "render_noalloc: " + c->noalloc_err);
    if (!out || !out->color || !out->depth || !out->t || !out->hit_id || !out->rgba8)
        return fail(c, MMHRT_E_INVALID, "render_noalloc: out needs device colour, depth, t, hit id and rgba8 buffers");
    if ((p->flags & MMHRT_FLAG_AUX) && !out->aux)
        return fail(c, MMHRT_E_INVALID, "render_noalloc: MMHRT_FLAG_AUX needs an aux buffer");
    const bool win = p->window[2] > 0;
    const uint32_t ow = win ? p->window[2] : p->width, oh = win ? p->window[3] : p->height;
    if (out->width != ow || out->height != oh)
        return fail(c, MMHRT_E_INVALID, "render_noalloc: out->width x out->height must be the render window (" +
                                            std::to_string(ow) + "x" + std::to_string(oh) + ")");
    FrameConsts f;
    frame_consts(cam, p, f);
    FrameOutputs o{static_cast<float4*>(out->color), static_cast<float*>(out->depth), static_cast<float*>(out->t),
                   static_cast<int32_t*>(out->hit_id), static_cast<uint32_t*>(out->rgba8),
                   static_cast<__half*>(out->aux)};
    return render_noalloc(c, f, o);
}

int32_t mmhrt_download(mmhrt_ctx c, mmhrt_framebuffer* out) {
    if (int e = check_ctx(c)) return e;
    if (!out) return fail(c, MMHRT_E_INVALID, "download: out is NULL");
    if (c->fb_w == 0) return fail(c, MMHRT_E_STATE, "download: nothing rendered yet");
    size_t P = (size_t)c->fb_w * c->fb_h;
    hipError_t e = hipSuccess;
    if (out->rgba8 && e == hipSuccess) e = hipMemcpyAsync(out->rgba8, c->fb_rgba8.p, P * 4, hipMemcpyDeviceToHost, c->stream);
    if (out->depth && e == hipSuccess) e = hipMemcpyAsync(out->depth, c->fb_depth.p, P * 4, hipMemcpyDeviceToHost, c->stream);
    if (out->color && e == hipSuccess) e = hipMemcpyAsync(out->color, c->fb_col.p, P * 16, hipMemcpyDeviceToHost, c->stream);
    if (out->t && e == hipSuccess) e = hipMemcpyAsync(out->t, c->fb_t.p, P * 4, hipMemcpyDeviceToHost, c->stream);
    if (out->hit_id && e == hipSuccess) e = hipMemcpyAsync(out->hit_id, c->fb_id.p, P * 4, hipMemcpyDeviceToHost, c->stream);
    if (out->aux && e == hipSuccess) {
        if (!c->fb_aux.p || c->fb_aux.cap < P * 8) return fail(c, MMHRT_E_STATE, "download: no aux buffer rendered");
        e = hipMemcpyAsync(out->aux, c->fb_aux.p, P * 16, hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "download");
    return MMHRT_OK;
}

int32_t mmhrt_get_device_framebuffer(mmhrt_ctx c, mmhrt_device_framebuffer* out) {
    if (int e = check_ctx(c)) return e;
    if (!out) return fail(c, MMHRT_E_INVALID, "get_device_framebuffer: out is NULL");
    if (c->fb_w == 0) return fail(c, MMHRT_E_STATE, "get_device_framebuffer: nothing rendered yet");
    size_t P = (size_t)c->fb_w * c->fb_h;
    out->color = c->fb_col.p;
    out->depth = c->fb_depth.p;
    out->t = c->fb_t.p;
    out->hit_id = c->fb_id.p;
    out->rgba8 = c->fb_rgba8.p;
    out->aux = (c->last_aux && c->fb_aux.p && c->fb_aux.cap >= P * 8) ? c->fb_aux.p : nullptr;
    out->stream = c->stream;
    out->width = c->fb_w;
    out->height = c->fb_h;
    return MMHRT_OK;
}

int32_t mmhrt_sync(mmhrt_ctx c) {
    if (int e = check_ctx(c)) return e;
    hipError_t e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? MMHRT_OK : hip_fail(c, e, "sync");
}

int32_t mmhrt_get_info(mmhrt_ctx c, uint64_t* n_spheres, float* bounds6, uint64_t* n_nodes) {
    if (int e = check_ctx(c)) return e;
    if (n_spheres) *n_spheres = c->n;
    if (n_nodes)
        *n_nodes = !c->built ? 0 : c->accel == 1 ? (uint64_t)c->pkd_L - 1 : (c->built_n > 1 ? c->built_n - 1 : 1);
    if (bounds6) {
        if (c->n == 0) return fail(c, MMHRT_E_STATE, "get_info: no spheres set");
        if (!c->info_valid) {
            if (int e = reduce_scene(c)) return e;
            hipError_t he = hipMemcpyAsync(&c->info_host, c->info.p, sizeof(SceneInfo), hipMemcpyDeviceToHost, c->stream);
            if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
            if (he != hipSuccess) return hip_fail(c, he, "get_info");
            c->info_valid = true;
        }
        for (int a = 0; a < 3; ++a) {
            bounds6[a] = c->info_host.blo[a];
            bounds6[3 + a] = c->info_host.bhi[a];
        }
    }
    return MMHRT_OK;
}

int32_t mmhrt_get_bvh(mmhrt_ctx c, float* nodes, uint32_t* leaf_ids, uint64_t* keys) {
    if (int e = check_ctx(c)) return e;
    if (!c->built) return fail(c, MMHRT_E_STATE, "get_bvh: nothing built");
    if (c->accel != 0) return fail(c, MMHRT_E_STATE, "get_bvh: the current structure is not an LBVH");
    uint64_t n = c->built_n, m = n > 1 ? n - 1 : 1;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && nodes) e = hipMemcpy(nodes, c->nodes.p, m * sizeof(Node), hipMemcpyDeviceToHost);
    if (e == hipSuccess && leaf_ids) e = hipMemcpy(leaf_ids, c->leaf_id.p, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && keys) {
        // the sorted keys stay in the first ping-pong buffer until the next build
        if (c->bits == 63) {
            e = hipMemcpy(keys, c->kla.p, n * 8, hipMemcpyDeviceToHost);
        } else {
            std::vector<uint32_t> k32(n);
            e = hipMemcpy(k32.data(), c->ka.p, n * 4, hipMemcpyDeviceToHost);
            for (uint64_t i = 0; i < n; ++i) keys[i] = k32[i];
        }
    }
    if (e != hipSuccess) return hip_fail(c, e, "get_bvh");
    return MMHRT_OK;
}

int32_t mmhrt_get_treelets(mmhrt_ctx c, uint32_t* sizes, float* spheres, uint32_t* ids, uint8_t* dims,
                           float* treelet_boxes, float* treelet_rmax, float* top_nodes) {
    if (int e = check_ctx(c)) return e;
    if (!c->built || c->accel != 1) return fail(c, MMHRT_E_STATE, "get_treelets: no treelet P-k-d trees built");
    const uint64_t n = c->built_n;
    const uint32_t T = c->pkd_T, L = c->pkd_L;
    if (sizes) {
        sizes[0] = c->pkd_M;
        sizes[1] = T;
        sizes[2] = L;
    }
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && spheres) e = hipMemcpy(spheres, c->leaf_sph.p, n * 16, hipMemcpyDeviceToHost);
    if (e == hipSuccess && ids) e = hipMemcpy(ids, c->leaf_id.p, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && dims) e = hipMemcpy(dims, c->pkd_dim.p, n, hipMemcpyDeviceToHost);
    if (e == hipSuccess && treelet_boxes) e = hipMemcpy(treelet_boxes, c->tl_box.p, (size_t)T * 24, hipMemcpyDeviceToHost);
    if (e == hipSuccess && treelet_rmax) e = hipMemcpy(treelet_rmax, c->tl_rmax.p, (size_t)T * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && top_nodes) e = hipMemcpy(top_nodes, c->nodes.p, (size_t)(L - 1) * sizeof(Node), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "get_treelets");
    return MMHRT_OK;
}

int32_t mmhrt_get_timings(mmhrt_ctx c, float* t, int32_t n) {
    if (int e = check_ctx(c)) return e;
    if (!t || n <= 0) return fail(c, MMHRT_E_INVALID, "get_timings: bad output");
    hipStreamSynchronize(c->stream);
    float v[9] = {elapsed(c, 0, 2), elapsed(c, 2, 3), elapsed(c, 3, 4), elapsed(c, 4, 5),
                  elapsed(c, 6, 7), 0.0f,           elapsed(c, 0, 5), elapsed(c, 6, 7),
                  elapsed(c, 8, 9)};
    for (int i = 0; i < n && i < 9; ++i) t[i] = v[i];
    return MMHRT_OK;
}

int32_t mmhrt_get_counters(mmhrt_ctx c, uint64_t* visits, uint64_t* tests, uint64_t* rays) {
    if (int e = check_ctx(c)) return e;
    if (!c->counters.p) return fail(c, MMHRT_E_STATE, "get_counters: nothing rendered");
    uint32_t h[16];
    hipError_t e = hipMemcpyAsync(h, c->counters.p, sizeof h, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "get_counters");
    uint64_t v, s, r;
    memcpy(&v, h + 2, 8);
    memcpy(&s, h + 4, 8);
    memcpy(&r, h + 6, 8);
    if (visits) *visits = v;
    if (tests) *tests = s;
    if (rays) *rays = r;
    c->stack_overflow = h[1];
    memcpy(&c->pair_visits, h + 8, 8);
    memcpy(&c->leaf_fetches, h + 10, 8);
    return MMHRT_OK;
}

static int64_t* option_slot(mmhrt_ctx c, const char* key) {
    if (!key) return nullptr;
    std::string k(key);
    if (k == "stack_lds") return &c->opt_stack_lds;
    if (k == "persistent") return &c->opt_persistent;
    if (k == "refill") return &c->opt_refill;
    if (k == "count_visits") return &c->opt_count;
    if (k == "blocks_per_cu") return &c->opt_blocks_per_cu;
    if (k == "tile_w") return &c->opt_tile_w;
    if (k == "tile_h") return &c->opt_tile_h;
    if (k == "refit_period") return &c->opt_refit_period;
    if (k == "dn_k32") return &c->opt_dn_k32;
    if (k == "dn_conv_shape") return &c->opt_dn_conv_shape;
    if (k == "build_chunk") return &c->opt_build_chunk;
    if (k == "sample_lanes") return &c->opt_sample_lanes;
    if (k == "build_ls") return &c->opt_build_ls;
    if (k == "sort_ri") return &c->opt_sort_ri;
    if (k == "build_grid") return &c->opt_build_grid;
    if (k == "sort_onesweep") return &c->opt_sort_onesweep;
    if (k == "sort_digit") return &c->opt_sort_digit;
    if (k == "leaf_pairs") return &c->opt_leaf_pairs;
    if (k == "build_tail") return &c->opt_build_tail;
    if (k == "build_wave") return &c->opt_build_wave;
    if (k == "sort_xcd") return &c->opt_sort_xcd;
    if (k == "build_span") return &c->opt_build_span;
    if (k == "accel") return &c->opt_accel;
    if (k == "fail_alloc") return &c->opt_fail_alloc;
    if (k == "treelet") return &c->opt_treelet;
    if (k == "batch_launch") return &c->opt_batch_launch;
    if (k == "stack_overflow") return &c->stack_overflow;
    if (k == "pair_visits") return &c->pair_visits;
    if (k == "leaf_fetches") return &c->leaf_fetches;
    return nullptr;
}

int32_t mmhrt_set_option(mmhrt_ctx c, const char* key, int64_t value) {
    if (!c) return MMHRT_E_INVALID;
    int64_t* s = option_slot(c, key);
    if (!s) return fail(c, MMHRT_E_INVALID, std::string("set_option: unknown key ") + (key ? key : "(null)"));
    *s = value;
    return MMHRT_OK;
}

int32_t mmhrt_get_option(mmhrt_ctx c, const char* key, int64_t* value) {
    if (!c || !value) return MMHRT_E_INVALID;
    if (key && std::string(key) == "sort_fault") {  // read from the device after the queued work
        uint32_t f = 0;
        if (c->sort_fault.p) {
            hipError_t e = hipMemcpyAsync(&f, c->sort_fault.p, 4, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) return hip_fail(c, e, "get_option sort_fault");
        }
        *value = f;
        return MMHRT_OK;
    }
    if (key && std::string(key) == "leaf_pair_count") {  // counted on the device over the current tree
        if (int e = count_leaf_pairs(c, value)) return e;
        return MMHRT_OK;
    }
    int64_t* s = option_slot(c, key);
    if (!s) return fail(c, MMHRT_E_INVALID, std::string("get_option: unknown key ") + (key ? key : "(null)"));
    *value = *s;
    return MMHRT_OK;
}

}  // extern "C"
