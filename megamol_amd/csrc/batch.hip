// batch.hip — batched time series: many timesteps and cameras in one hipGraph launch.
//
// No single counterpart in the reference: MegaMol renders one frame per Render call and rebuilds
// its OptiX BVH whenever the frame changes (SphereGeometry.cpp:252-257, always
// OPTIX_BUILD_OPERATION_BUILD). Here all timesteps stay resident in HBM; for each timestep the
// LBVH is rebuilt or refitted (refit keeps the topology and recomputes the boxes) into a structure
// of its own, and then one persistent launch traces every (timestep, camera) job of the series,
// stealing tiles across all of them (option batch_launch 0: one launch per timestep, right after
// its build, over the structure buffers the timesteps share). The whole sequence is captured
// once into a hipGraph and launched.
#include <string.h>

#include <algorithm>
#include <vector>

#include "ctx.hpp"

using namespace mmhrt;

namespace {
void frame_consts_batch(const mmhrt_camera* cam, const mmhrt_render_params* p, FrameConsts& f) {
    // the same derivation as mmhrt_render (Renderer.cpp:136-172)
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
    f.max_bounces = p->max_bounces;
    if (p->window[2] > 0) {
        f.win_x0 = p->window[0];
        f.win_y0 = p->window[1];
        f.win_w = p->window[2];
        f.win_h = p->window[3];
    } else {
        f.win_w = p->width;
        f.win_h = p->height;
    }
}
}  // namespace

extern "C" {

int32_t mmhrt_set_timesteps(mmhrt_ctx c, const float* xyzr, uint32_t timesteps, uint64_t count, const float* rgba,
                            const uint8_t global_rgba[4]) {
    if (!c) return MMHRT_E_INVALID;
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MMHRT_E_HIP, "hipSetDevice");
    if (!xyzr || timesteps == 0 || count == 0) return fail(c, MMHRT_E_INVALID, "set_timesteps: empty input");
    if (count > 0x7fffffffull) return fail(c, MMHRT_E_UNSUPPORTED, "set_timesteps: more than 2^31-1 spheres");
    size_t total = (size_t)timesteps * count;
    int e = 0;
    if ((e = ensure(c, c->ts_sph, total, "timestep spheres"))) return e;
    if ((e = ensure(c, c->ts_col, count, "timestep colours"))) return e;
    hipError_t he = hipMemcpyAsync(c->ts_sph.p, xyzr, total * 16, hipMemcpyHostToDevice, c->stream);
    if (he != hipSuccess) return hip_fail(c, he, "set_timesteps upload");
    // the timesteps keep their own colour buffer, so that mmhrt_set_spheres (the context's
    // spheres and colours) and the series never overwrite each other
    float4 g = make_float4(1.f, 1.f, 1.f, 1.f);
    if (global_rgba)
        g = make_float4(global_rgba[0] / 255.0f, global_rgba[1] / 255.0f, global_rgba[2] / 255.0f, global_rgba[3] / 255.0f);
    if (rgba) {
        he = hipMemcpyAsync(c->ts_col.p, rgba, count * 16, hipMemcpyHostToDevice, c->stream);
        if (he != hipSuccess) return hip_fail(c, he, "set_timesteps colours");
    } else {
        launch_fill_f4(c->stream, c->ts_col.p, g, count);
    }
    he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) return hip_fail(c, he, "set_timesteps");
    c->ts_count = timesteps;
    c->ts_n = count;
    return MMHRT_OK;
}

int32_t mmhrt_render_batch(mmhrt_ctx c, const mmhrt_job* jobs, uint32_t n_jobs, int32_t refit_period,
                           int32_t morton_bits, int32_t use_graph, uint32_t* rgba8_out, float* depth_out,
                           float* ms_out) {
    if (!c) return MMHRT_E_INVALID;
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MMHRT_E_HIP, "hipSetDevice");
    if (!jobs || n_jobs == 0) return fail(c, MMHRT_E_INVALID, "render_batch: no jobs");
    if (c->ts_count == 0) return fail(c, MMHRT_E_STATE, "render_batch: call mmhrt_set_timesteps first");
    if (morton_bits != 30 && morton_bits != 63) return fail(c, MMHRT_E_INVALID, "render_batch: morton_bits");
    if (refit_period < 1) refit_period = 1;
    // validate jobs, size the outputs, group by timestep (kept in submission order inside a timestep)
    uint64_t total_px = 0;
    std::vector<std::vector<uint32_t>> by_t(c->ts_count);
    for (uint32_t j = 0; j < n_jobs; ++j) {
        const mmhrt_job& J = jobs[j];
        const mmhrt_render_params& p = J.params;
        if (J.timestep >= c->ts_count) return fail(c, MMHRT_E_INVALID, "render_batch: timestep out of range");
        if (p.width == 0 || p.height == 0 || p.spp == 0 || (p.mode != 0 && p.mode != 1))
            return fail(c, MMHRT_E_INVALID, "render_batch: job needs a viewport, spp >= 1 and mode Lambert or AO");
        if (!(J.camera.fovy > 0.0f) || !(J.camera.aspect > 0.0f))
            return fail(c, MMHRT_E_INVALID, "render_batch: bad camera");
        if (p.window[2] > 0 && ((uint64_t)p.window[0] + p.window[2] > p.width ||
                                (uint64_t)p.window[1] + p.window[3] > p.height || p.window[3] == 0))
            return fail(c, MMHRT_E_INVALID, "render_batch: window outside the frame");
        uint64_t w = p.window[2] > 0 ? p.window[2] : p.width, h = p.window[2] > 0 ? p.window[3] : p.height;
        total_px = std::max<uint64_t>(total_px, J.out_offset + w * h);
        by_t[J.timestep].push_back(j);
    }
    if (total_px > (1ull << 32)) return fail(c, MMHRT_E_UNSUPPORTED, "render_batch: outputs too large");
    const uint32_t T = c->ts_count;
    // one launch: all jobs of one mode (Lambert, AO) in one persistent launch after every
    // timestep's structure is built; otherwise one launch per timestep after its build, whose jobs
    // then share the mode
    const bool one = c->opt_batch_launch != 0;
    if (!one) {
        for (uint32_t t = 0; t < T; ++t) {
            int mode = -1;
            for (uint32_t j : by_t[t]) {
                if (mode >= 0 && jobs[j].params.mode != mode)
                    return fail(c, MMHRT_E_INVALID, "render_batch: jobs of one timestep must share the mode");
                mode = jobs[j].params.mode;
            }
        }
    }
    // launch groups: consecutive jobs in launch order (by timestep, or by mode then timestep)
    std::vector<uint32_t> order;
    struct Launch {
        uint32_t first, count, prefix, timestep;
        int mode;
    };
    std::vector<Launch> launches;
    for (int m = 0; m < (one ? 2 : 1); ++m) {
        for (uint32_t t = 0; t < T; ++t) {
            for (uint32_t j : by_t[t]) {
                if (one && jobs[j].params.mode != m) continue;
                if (launches.empty() || (one ? launches.back().mode != m : launches.back().timestep != t))
                    launches.push_back(Launch{(uint32_t)order.size(), 0, 0, t, jobs[j].params.mode});
                order.push_back(j);
                ++launches.back().count;
            }
        }
    }
    int e = 0;
    if ((e = ensure(c, c->bt_col, total_px, "batch colour"))) return e;
    if ((e = ensure(c, c->bt_depth, total_px, "batch depth"))) return e;
    if ((e = ensure(c, c->bt_t, total_px, "batch t"))) return e;
    if ((e = ensure(c, c->bt_id, total_px, "batch id"))) return e;
    if ((e = ensure(c, c->bt_rgba8, total_px, "batch rgba8"))) return e;
    if ((e = ensure(c, c->batch_ctr, launches.size() * 16, "batch counters"))) return e;
    const size_t js = batch_job_size();
    if ((e = ensure(c, c->batch_jobs, js * n_jobs, "batch jobs"))) return e;
    // The builds read the timestep slices and their count; the context's own spheres, count and
    // LBVH state come back when this call returns (the tree in the buffers is then the last
    // timestep's, so the context needs a build before its next render).
    // The per-timestep structures are released too: they are rebuilt by every call, and holding
    // them (one LBVH per timestep) would keep memory a later, larger scene needs.
    struct Restore {
        mmhrt_ctx_s* c;
        uint64_t n;
        ~Restore() {
            c->sph_cur = nullptr;
            c->n = n;
            c->built = false;
            c->info_valid = false;
            for (auto& sl : c->ts_slots) {
                release(sl.nodes);
                release(sl.leaf);
                release(sl.id);
                release(sl.info);
                release(sl.dim);
                release(sl.box);
                release(sl.rmax);
            }
            c->ts_slots.clear();
        }
    } restore{c, c->n};
    // every buffer the per-timestep build and trace touch is allocated before any capture
    c->sph_cur = c->ts_sph.p;
    c->n = c->ts_n;
    // treelet P-k-d trees (option accel 1) are rebuilt for every timestep: they have no refit
    const bool tl = c->opt_accel == 1;
    if ((e = tl ? build_pkd(c, morton_bits) : build_lbvh(c, morton_bits))) return e;
    {
        unsigned grid = (unsigned)std::max<int64_t>(1, c->cu_count * c->opt_blocks_per_cu);
        if ((e = ensure(c, c->spill, (size_t)grid * 256 * 120, "traversal spill"))) return e;
    }
    // One launch reads every timestep's structure, so each timestep but the last builds into a
    // slot of its own (the last one into the context's buffers), and each rebuild but the last
    // keeps its leaf order (which its refits share) in a slot too. bind(t) swaps timestep t's
    // buffers into the context, so that the build and the job table see them; a second call
    // swaps them back.
    const uint32_t period = tl ? 1u : (uint32_t)refit_period;
    const uint32_t G = (T - 1) / period + 1;  // rebuild groups
    if (one && T > 1) {
        // sized for this series (the context's own buffers may be larger, from an earlier scene)
        const uint64_t n = c->ts_n;
        const size_t n_nodes = tl ? (size_t)c->pkd_L - 1 : (size_t)(n > 1 ? n - 1 : 1);
        if (c->ts_slots.size() < T - 1) c->ts_slots.resize(T - 1);
        for (uint32_t t = 0; t + 1 < T; ++t) {
            SeriesSlot& sl = c->ts_slots[t];
            if ((e = ensure(c, sl.nodes, n_nodes, "series nodes"))) return e;
            if ((e = ensure(c, sl.leaf, n, "series leaf spheres"))) return e;
            if ((e = ensure(c, sl.info, 1, "series scene info"))) return e;
            if (t + 1 < G && (e = ensure(c, sl.id, n, "series leaf ids"))) return e;
            if (tl) {
                if ((e = ensure(c, sl.dim, n, "series split axes"))) return e;
                if ((e = ensure(c, sl.box, (size_t)c->pkd_T * 6, "series treelet boxes"))) return e;
                if ((e = ensure(c, sl.rmax, c->pkd_T, "series treelet radii"))) return e;
            }
        }
    }
    auto bind = [&](uint32_t t) {
        if (!one) return;
        if (t + 1 < T) {
            SeriesSlot& sl = c->ts_slots[t];
            std::swap(c->nodes, sl.nodes);
            std::swap(c->leaf_sph, sl.leaf);
            std::swap(c->info, sl.info);
            if (tl) {
                std::swap(c->pkd_dim, sl.dim);
                std::swap(c->tl_box, sl.box);
                std::swap(c->tl_rmax, sl.rmax);
            }
        }
        const uint32_t g = t / period;
        if (g + 1 < G) std::swap(c->leaf_id, c->ts_slots[g].id);
    };
    // job table and tile prefixes (the jobs of a launch are contiguous, each launch's prefix
    // array starts at 0)
    std::vector<uint8_t> hjobs(js * n_jobs);
    std::vector<uint32_t> hprefix;
    for (Launch& L : launches) {
        L.prefix = (uint32_t)hprefix.size();
        uint32_t acc = 0;
        hprefix.push_back(0);
        for (uint32_t k = L.first; k < L.first + L.count; ++k) {
            const mmhrt_job& J = jobs[order[k]];
            FrameConsts f;
            frame_consts_batch(&J.camera, &J.params, f);
            bind(J.timestep);
            const int tiles = batch_fill_job(c, f, J.out_offset, hjobs.data() + js * k);
            bind(J.timestep);
            acc += (uint32_t)tiles;
            hprefix.push_back(acc);
        }
    }
    if ((e = ensure(c, c->batch_prefix, hprefix.size(), "batch tile prefix"))) return e;
    hipError_t he = hipMemcpyAsync(c->batch_jobs.p, hjobs.data(), hjobs.size(), hipMemcpyHostToDevice, c->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(c->batch_prefix.p, hprefix.data(), hprefix.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) return hip_fail(c, he, "render_batch upload");
    auto launch = [&](size_t li) {
        const Launch& L = launches[li];
        launch_trace_batch(c, c->batch_jobs.p + js * L.first, c->batch_prefix.p + L.prefix, L.count,
                           c->batch_ctr.p + 16 * li, L.mode, tl);
    };
    auto issue = [&]() -> int {
        size_t li = 0;
        for (uint32_t t = 0; t < T; ++t) {
            c->sph_cur = c->ts_sph.p + (size_t)t * c->ts_n;
            bind(t);
            int r = tl ? build_pkd(c, morton_bits)
                       : (t % (uint32_t)refit_period == 0) ? build_lbvh(c, morton_bits) : refit_lbvh(c);
            bind(t);
            if (r) return r;
            if (!one && li < launches.size() && launches[li].timestep == t) launch(li++);
        }
        for (; one && li < launches.size(); ++li) launch(li);
        return 0;
    };
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    int rc = 0;
    if (use_graph) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        he = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (he == hipSuccess) {
            c->capturing = true;
            rc = issue();
            c->capturing = false;
            hipError_t he2 = hipStreamEndCapture(c->stream, &graph);
            if (rc == 0 && he2 != hipSuccess) rc = hip_fail(c, he2, "render_batch capture");
        } else {
            rc = hip_fail(c, he, "render_batch begin capture");
        }
        if (rc == 0) {
            he = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (he != hipSuccess) rc = hip_fail(c, he, "render_batch instantiate");
        }
        if (rc == 0) {
            hipEventRecord(e0, c->stream);
            he = hipGraphLaunch(exec, c->stream);
            hipEventRecord(e1, c->stream);
            if (he != hipSuccess) rc = hip_fail(c, he, "render_batch graph launch");
        }
        if (rc == 0) {
            he = hipStreamSynchronize(c->stream);
            if (he != hipSuccess) rc = hip_fail(c, he, "render_batch");
        }
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
    } else {
        hipEventRecord(e0, c->stream);
        rc = issue();
        hipEventRecord(e1, c->stream);
        if (rc == 0) {
            he = hipStreamSynchronize(c->stream);
            if (he != hipSuccess) rc = hip_fail(c, he, "render_batch");
        }
    }
    if (rc == 0 && ms_out) hipEventElapsedTime(ms_out, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    if (rgba8_out) he = hipMemcpy(rgba8_out, c->bt_rgba8.p, total_px * 4, hipMemcpyDeviceToHost);
    if (he == hipSuccess && depth_out) he = hipMemcpy(depth_out, c->bt_depth.p, total_px * 4, hipMemcpyDeviceToHost);
    if (he != hipSuccess) return hip_fail(c, he, "render_batch download");
    return MMHRT_OK;
}

}  // extern "C"
