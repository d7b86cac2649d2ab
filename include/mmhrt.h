/*
 * mmhrt — MI355X (gfx950) particle-sphere ray tracer, C ABI.
 *
 * This is the drop-in boundary for MegaMol's sphere ray-tracing path. A MegaMol
 * renderer module (or any other host: ctypes, a benchmark, a test harness) calls
 * these functions; everything behind them is hand-written HIP for gfx950.
 *
 * Reference interfaces replaced (paths relative to the MegaMol source tree):
 *   - plugins/optix_hpg/src/optix/SphereGeometry.cpp:78-227   (assertData: per-list
 *     conversion, upload, bounds kernel, optixAccelBuild)        -> mmhrt_set_spheres + mmhrt_build
 *   - plugins/datatools/src/AddParticleColors.cpp:40-117        -> mmhrt_set_transfer_function
 *   - plugins/optix_hpg/src/optix/Renderer.cpp:136-172          (camera -> FrameState)
 *     plugins/optix_hpg/src/optix/AbstractRenderer.cpp:27-111   (Render: change detection, launch)
 *     plugins/optix_hpg/cuda_resources/raygen.cu:79-278          (raygen + traceRay) -> mmhrt_render
 *   - frontend/services/cuda/CUDA_Service.cpp:30-56             (device context)    -> mmhrt_create
 *
 * Conventions
 *   - Every function returns an int32 status (MMHRT_OK == 0). On failure the
 *     context keeps a message, readable through mmhrt_last_error().
 *   - A context is single-threaded; calls are synchronous unless stated.
 *   - Inputs passed as host pointers are copied before the call returns, matching
 *     the lock/unlock lifetime of MultiParticleDataCall frames
 *     (plugins/moldyn/src/io/MMPLDDataSource.cpp:443-461).
 *   - Pixel (x, y) has y = 0 at the bottom row, like the reference's CUDA surface
 *     (raygen.cu:241-242 maps y = 0 to v = -th).
 */
#ifndef MMHRT_H
#define MMHRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMHRT_API __attribute__((visibility("default")))

/* ---- status codes ------------------------------------------------------- */
enum {
    MMHRT_OK = 0,
    MMHRT_E_INVALID = 1,     /* bad argument (null pointer, bad enum, size 0 where >0 needed) */
    MMHRT_E_HIP = 2,         /* a HIP runtime call failed */
    MMHRT_E_OOM = 3,         /* device allocation failed */
    MMHRT_E_UNSUPPORTED = 4, /* valid request that this build does not implement */
    MMHRT_E_STATE = 5        /* call order violated (e.g. render before build) */
};

/* ---- particle list: mirrors geocalls::SimpleSphericalParticles ----------- */
/* Vertex types: plugins/geometry_calls/include/geometry_calls/SimpleSphericalParticles.h:27-33 */
enum {
    MMHRT_VERTDATA_NONE = 0,
    MMHRT_VERTDATA_FLOAT_XYZ = 1,
    MMHRT_VERTDATA_FLOAT_XYZR = 2,
    MMHRT_VERTDATA_SHORT_XYZ = 3,
    MMHRT_VERTDATA_DOUBLE_XYZ = 4
};
/* Colour types: SimpleSphericalParticles.h:41-50 (note: these are the in-memory enum
 * values, not the MMPLD file codes of MMPLDDataSource.cpp:117-156). */
enum {
    MMHRT_COLDATA_NONE = 0,
    MMHRT_COLDATA_UINT8_RGB = 1,
    MMHRT_COLDATA_UINT8_RGBA = 2,
    MMHRT_COLDATA_FLOAT_RGB = 3,
    MMHRT_COLDATA_FLOAT_RGBA = 4,
    MMHRT_COLDATA_FLOAT_I = 5,
    MMHRT_COLDATA_USHORT_RGBA = 6,
    MMHRT_COLDATA_DOUBLE_I = 7
};

/* One particle list, as AccessParticles(i) exposes it
 * (plugins/geometry_calls/include/geometry_calls/AbstractParticleDataCall.h:61-81).
 * Strides of 0 mean "tightly packed" (the element size of SimpleSphericalParticles.cpp:7-9). */
typedef struct mmhrt_sphere_list {
    int32_t vert_type;
    int32_t col_type;
    const void* vert;       /* host pointer to the first vertex */
    const void* col;        /* host pointer to the first colour (ignored for COLDATA_NONE) */
    uint64_t vert_stride;   /* bytes between vertices */
    uint64_t col_stride;    /* bytes between colours */
    uint64_t count;
    float global_radius;    /* used unless vert_type == FLOAT_XYZR */
    uint8_t global_rgba[4]; /* used when the list has no per-particle colour */
    float col_min;          /* colour-map index range of FLOAT_I / DOUBLE_I lists */
    float col_max;
} mmhrt_sphere_list;

/* ---- camera: mirrors core::view::Camera Pose + PerspectiveParameters ----- */
/* core/include/mmcore/view/Camera.h:41-115; fovy is in radians (Camera.h:367). */
typedef struct mmhrt_camera {
    float position[3];
    float direction[3];
    float up[3];
    float fovy;       /* full vertical field of view, radians */
    float aspect;     /* width / height */
    float near_plane; /* projection near plane (depth conversion only) */
    float far_plane;  /* projection far plane (depth conversion only) */
    float tile_start[2]; /* image-plane tile (Camera.h:94-99), default {0,0}; carried for the
                            interface but not read: the reference's frame state has no tile and
                            its raygen spans the full image plane (Renderer.cpp:70-93, raygen.cu:241-242).
                            Use the render window to render part of a frame. */
    float tile_end[2];   /* default {1,1} */
} mmhrt_camera;

/* ---- render parameters: mirrors the optix_hpg Renderer params ------------ */
/* spp / max bounces / accumulate: plugins/optix_hpg/src/optix/Renderer.cpp:22-39;
 * AO samples / distance: plugins/mmospray/src/AbstractOSPRayRenderer.cpp:370-377. */
enum {
    MMHRT_MODE_LAMBERT = 0, /* primary rays + albedo*(0.2+0.6|N.d|) (raygen.cu:140-142) */
    MMHRT_MODE_AO = 1,      /* primary rays + 1 cosine-weighted AO ray per sample */
    MMHRT_MODE_BOUNCE = 2   /* diffuse multi-bounce path of raygen.cu:144-154 */
};
enum {
    MMHRT_FLAG_AUX = 1u << 0,        /* also write fp16 denoiser features (albedo, normal, depth) */
    MMHRT_FLAG_NO_DOWNLOAD = 1u << 1 /* keep the frame on the device; read it later with mmhrt_download */
};
typedef struct mmhrt_render_params {
    uint32_t width;
    uint32_t height;
    uint32_t spp;          /* samples per pixel, >= 1 */
    int32_t mode;          /* MMHRT_MODE_* */
    float ao_distance;     /* AO ray length (tmax) */
    int32_t max_bounces;   /* MMHRT_MODE_BOUNCE only */
    uint32_t frame_idx;    /* RNG frame index (the TEA seed of raygen.cu:182) */
    int32_t accumulate;    /* 1: with frame_idx > 0 the frame is blended into the previous one,
                              old + (new - old) / (frame_idx + 1), as raygen.cu:254-260 always does
                              (the reference's raygen never reads its own accumulate flag,
                              Renderer.cpp:162); the previous frame must be the same window of a
                              frame of the same size (else MMHRT_E_STATE). 0: frame_idx only seeds
                              the random streams and nothing is blended. */
    float background[4];   /* RGBA background (CallRender3D::BackgroundColor) */
    uint32_t flags;        /* MMHRT_FLAG_* */
    uint32_t window[4];    /* x0, y0, w, h: render only this pixel window of the width x height frame
                              (w == 0: the whole frame); outputs are w x h and equal that crop of the
                              full frame bit for bit */
} mmhrt_render_params;

/* ---- framebuffer: host destinations (any pointer may be NULL = skip) ------ */
typedef struct mmhrt_framebuffer {
    uint32_t* rgba8;   /* width*height RGBA8, as CPUFramebuffer::colorBuffer (CPUFramebuffer.h:21) */
    float* depth;      /* pseudo depth 0.5*(C/t - A)+0.5, 1.0 on miss (raygen.cu:262-268) */
    float* color;      /* width*height*4 float RGBA before quantisation */
    float* t;          /* linear ray distance of the nearest primary hit, +inf on miss */
    int32_t* hit_id;   /* global particle index of the nearest primary hit, -1 on miss */
    uint16_t* aux;     /* width*height*8 fp16: albedo rgb, normal xyz, linear depth, visibility */
} mmhrt_framebuffer;

/* ---- framebuffer: device views of the last rendered frame ----------------- */
typedef struct mmhrt_device_framebuffer {
    void* color;     /* width*height float4 RGBA (RGBA32F, the reference's colour surface) */
    void* depth;     /* width*height float pseudo depth (R32F, the reference's depth surface) */
    void* t;         /* width*height float linear distance */
    void* hit_id;    /* width*height int32 global particle index, -1 on miss */
    void* rgba8;     /* width*height RGBA8 */
    void* aux;       /* width*height*8 fp16 features, NULL unless the frame had MMHRT_FLAG_AUX */
    void* stream;    /* hipStream_t the frame was rendered on: consumers order their work after it */
    uint32_t width, height;
} mmhrt_device_framebuffer;

/* ---- build policy --------------------------------------------------------- */
enum {
    MMHRT_BUILD_REBUILD = 0, /* full LBVH build (the reference always builds: SphereGeometry.cpp:197-214) */
    MMHRT_BUILD_REFIT = 1,   /* keep topology, recompute boxes from moved spheres */
    MMHRT_BUILD_AUTO = 2     /* refit while the measured cost ratio allows, else rebuild */
};

typedef struct mmhrt_ctx_s* mmhrt_ctx;

/* Device context. Replaces CUDA_Service::init (frontend/services/cuda/CUDA_Service.cpp:30-56). */
MMHRT_API int32_t mmhrt_create(int32_t device, mmhrt_ctx* out);
MMHRT_API int32_t mmhrt_destroy(mmhrt_ctx ctx);
/* Message of the last failed call on this context ("" if none). ctx may be NULL for create errors. */
MMHRT_API const char* mmhrt_last_error(mmhrt_ctx ctx);

/* Upload and convert particle lists: SphereGeometry::assertData (SphereGeometry.cpp:94-146).
 * Lists with count 0 or VERTDATA_NONE are skipped; particle IDs are the running index over
 * all non-skipped lists in order. data_hash / frame_id mirror MultiParticleDataCall
 * (AbstractGetDataCall.h:54, AbstractGetData3DCall.h:49) and are reported back unchanged. */
MMHRT_API int32_t mmhrt_set_spheres(mmhrt_ctx ctx, const mmhrt_sphere_list* lists, uint32_t n_lists,
                                    uint64_t data_hash, uint32_t frame_id);
/* Same, with data already in device memory as packed float4 xyzr and optional float4 rgba
 * (NULL: every sphere takes global_rgba/255). Device pointers are copied on-device. */
MMHRT_API int32_t mmhrt_set_spheres_device(mmhrt_ctx ctx, const void* xyzr_dev, const void* rgba_dev,
                                           uint64_t count, const uint8_t global_rgba[4]);
/* Transfer function for FLOAT_I / DOUBLE_I lists: AddParticleColors::manipulateData
 * (plugins/datatools/src/AddParticleColors.cpp:40-117). rgba has n*4 floats. If vmin > vmax the
 * range is taken from the lists' [col_min, col_max] (AddParticleColors.cpp:59-71). Applies to the
 * next mmhrt_set_spheres; n == 0 removes it (intensity lists then use their global colour,
 * SphereGeometry.cpp:102-105). */
MMHRT_API int32_t mmhrt_set_transfer_function(mmhrt_ctx ctx, const float* rgba, uint32_t n, float vmin, float vmax);

/* Build (or refit) the LBVH over the current spheres: replaces optixAccelBuild
 * (SphereGeometry.cpp:197-214). morton_bits is 30 or 63. With option "accel" 1 it builds the
 * treelet P-k-d trees instead (plugins/optix_owl/src/TreeletsRenderer.cpp:95-110: the Morton
 * order cut into treelets of option "treelet" spheres, each a left-balanced k-d tree over its
 * spheres, a tree of boxes on top), which have no refit: every policy rebuilds them. */
MMHRT_API int32_t mmhrt_build(mmhrt_ctx ctx, int32_t policy, int32_t morton_bits);

/* Render one frame: AbstractRenderer::Render + raygen_program (AbstractRenderer.cpp:27-111,
 * raygen.cu:159-278). out may be NULL together with MMHRT_FLAG_NO_DOWNLOAD. */
MMHRT_API int32_t mmhrt_render(mmhrt_ctx ctx, const mmhrt_camera* cam, const mmhrt_render_params* params,
                               mmhrt_framebuffer* out);
/* Copy the last rendered frame to host buffers (any pointer NULL = skip). */
MMHRT_API int32_t mmhrt_download(mmhrt_ctx ctx, mmhrt_framebuffer* out);

/* Device pointers of the last rendered frame for zero-copy chaining on the GPU: the analogue of
 * the CUDA surfaces and exec stream of CallRender3DCUDA (plugins/optix_hpg/src/CallRender3DCUDA.h:10-26).
 * The buffers belong to the context and stay valid until the next render, a resize or
 * mmhrt_destroy; the frame may still be in flight on `stream`. */
MMHRT_API int32_t mmhrt_get_device_framebuffer(mmhrt_ctx ctx, mmhrt_device_framebuffer* out);

/* Render one frame into device buffers the caller owns, allocating nothing. No counterpart in the
 * reference: for callers that cannot allocate device memory (work under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The frame is the one mmhrt_render produces
 * (same kernels, same bits); out holds device pointers to colour (float4), depth, t (float),
 * hit_id (int32), rgba8 (uint32) and, with MMHRT_FLAG_AUX, aux (8 fp16 per pixel) of
 * out->width x out->height pixels, which must be the render window (stream is not read). With
 * accumulate and frame_idx > 0 the colour buffer must hold the previous frame. The call is
 * asynchronous on the context's stream (mmhrt_sync waits for it); it needs a build, and the
 * scene, the structure and out must stay alive until it completes. The trace's scratch is a
 * fixed device array of 62.9 MB per GPU (tile counter, flags and the traversal stacks of at most
 * 131,072 persistent lanes: the persistent grid is capped to two workgroups per CU, so this entry
 * point traces slower than mmhrt_render), held until the process ends and shared by every call on
 * that GPU: calls that overlap on one GPU race. Frames larger than 2^31 pixels are rejected. */
MMHRT_API int32_t mmhrt_render_noalloc(mmhrt_ctx ctx, const mmhrt_camera* cam, const mmhrt_render_params* params,
                                       const mmhrt_device_framebuffer* out);

/* Wait for all queued work on the context's stream. */
MMHRT_API int32_t mmhrt_sync(mmhrt_ctx ctx);

/* Scene facts: sphere count, scene bounds (6 floats: min xyz, max xyz incl. radii), LBVH node count. */
MMHRT_API int32_t mmhrt_get_info(mmhrt_ctx ctx, uint64_t* n_spheres, float* bounds6, uint64_t* n_nodes);

/* Copy the LBVH to host for structural checks: nodes (n_nodes*16 floats: child0 lo.xyz hi.xyz,
 * child1 lo.xyz hi.xyz, then as int32 bit patterns the two child references (>= 0: node index,
 * < 0: ~k for leaf k), the pair flags and the subtree height), leaf_ids (n_spheres uint32,
 * particle index of leaf k), keys (n_spheres uint64 Morton keys in leaf order). A node whose
 * two children are leaves k, k+1 (a leaf pair) is not stored, except the root: its parent holds
 * ~k with pair flag bit 0 (child0) or bit 1 (child1) set, and its slot is left unwritten. Any
 * pointer may be NULL. */
MMHRT_API int32_t mmhrt_get_bvh(mmhrt_ctx ctx, float* nodes, uint32_t* leaf_ids, uint64_t* keys);

/* Copy the treelet P-k-d trees (built with option "accel" 1) to host for structural checks:
 * sizes[3] = treelet size M, treelet count T, top-tree leaves L (a power of two >= max(T, 2));
 * spheres (n*4 floats) and ids (n, particle index) in treelet heap order: node i of treelet t at
 * position t*M + i, children 2i+1 and 2i+2 (plugins/optix_owl/src/PKDCreate.h:86-160 recBuild,
 * TreeletsCreate.h:117-166 prePartition_inPlace); dims (n bytes, split axis of each node);
 * treelet_boxes (T*6: sphere bounds), treelet_rmax (T: largest radius); top_nodes ((L-1)*16
 * floats in the layout of mmhrt_get_bvh: child ref >= 0 a top node, ~t treelet t, INT32_MIN an
 * absent child whose box holds 3e38 everywhere). Any pointer may be NULL. */
MMHRT_API int32_t mmhrt_get_treelets(mmhrt_ctx ctx, uint32_t* sizes, float* spheres, uint32_t* ids, uint8_t* dims,
                                     float* treelet_boxes, float* treelet_rmax, float* top_nodes);

/* Per-kernel timing of the last mmhrt_build / mmhrt_render on the context's stream, measured
 * with HIP events: times_ms[0..8] = morton, sort, hierarchy, bounds, trace, resolve, total build,
 * total render, and the hierarchy's chunk pass alone (n up to 9). */
MMHRT_API int32_t mmhrt_get_timings(mmhrt_ctx ctx, float* times_ms, int32_t n);

/* Instrumented trace counters of the last render (when enabled with mmhrt_set_option
 * "count_visits"): node visits, sphere tests, rays traced. Also refreshes the options
 * "stack_overflow" (1 if a traversal stack ever exceeded its capacity; 0 expected),
 * "pair_visits" (the node visits that were leaf pairs: two 16-byte spheres read instead of a
 * 64-byte node) and "leaf_fetches" (spheres read for the leaf children of stored nodes). */
MMHRT_API int32_t mmhrt_get_counters(mmhrt_ctx ctx, uint64_t* node_visits, uint64_t* sphere_tests, uint64_t* rays);

/* ---- batched time series ---------------------------------------------------
 * No single counterpart in the reference, which renders one frame per Render call and rebuilds
 * its BVH on every frame change (SphereGeometry.cpp:252-257). All timesteps stay resident; per
 * timestep the LBVH is rebuilt (t % refit_period == 0) or refitted into a structure of its own,
 * then one persistent launch per mode (Lambert, AO) traces every job of the series; with option
 * "batch_launch" 0 each timestep's jobs are traced right after its build instead (one launch per
 * timestep, whose jobs must then share the mode; no per-timestep structures are kept). use_graph
 * captures the whole sequence into one hipGraph and launches it once. */
typedef struct mmhrt_job {
    uint32_t timestep;
    uint32_t reserved;
    mmhrt_camera camera;
    mmhrt_render_params params; /* mode Lambert or AO; window honoured; no accumulation */
    uint64_t out_offset;        /* first pixel of this job in the batch outputs */
} mmhrt_job;
/* Upload `timesteps` x `count` packed float4 xyzr spheres (timestep-major) and one colour per
 * sphere shared by all timesteps (rgba NULL: global_rgba / 255). */
MMHRT_API int32_t mmhrt_set_timesteps(mmhrt_ctx ctx, const float* xyzr, uint32_t timesteps, uint64_t count,
                                      const float* rgba, const uint8_t global_rgba[4]);
/* Render all jobs; rgba8_out / depth_out (NULL to skip) receive the concatenated job outputs;
 * ms_out the device time of the whole sequence (graph launch or eager issue). */
MMHRT_API int32_t mmhrt_render_batch(mmhrt_ctx ctx, const mmhrt_job* jobs, uint32_t n_jobs, int32_t refit_period,
                                     int32_t morton_bits, int32_t use_graph, uint32_t* rgba8_out, float* depth_out,
                                     float* ms_out);

/* ---- learned denoiser ------------------------------------------------------
 * No counterpart in the reference (optix_hpg has no denoiser). A 4-level U-Net (3x3 convs,
 * 32/48/64/96 channels, ReLU, skip concatenation) run and trained with BF16 MFMA kernels that
 * accumulate in FP32. It predicts per-pixel AO visibility v (output channel 0) and a colour
 * residual r (channels 1-3); the output colour is albedo * v + r on hit pixels and the
 * background elsewhere (the residual corrects the one-sample albedo towards the pixel's mean).
 * Parameters (fp32, canonical order): per layer W[tap][c_in][c_out] then bias[c_out]; layers
 * enc0a(16->32) enc0b(32->32) enc1a(32->48) enc1b(48->48) enc2a(48->64) enc2b(64->64)
 * bota(64->96) botb(96->96) dec2a(64+96->64) dec2b(64->64) dec1a(48+64->48) dec1b(48->48)
 * dec0a(32+48->32) dec0b(32->32) out(32->16, channels 0-3 used, no ReLU).
 * Feature record per pixel (12 fp16): colour rgb, albedo rgb, normal xyz, linear depth t
 * (0 on miss), visibility, 0 — the colour channels of a render plus its MMHRT_FLAG_AUX buffer.
 * depth_scale divides t before it enters the network. */
MMHRT_API int32_t mmhrt_denoiser_init(mmhrt_ctx ctx, uint64_t seed);
MMHRT_API int32_t mmhrt_denoiser_param_count(mmhrt_ctx ctx, uint64_t* n);
MMHRT_API int32_t mmhrt_denoiser_get_params(mmhrt_ctx ctx, float* out, uint64_t n);
MMHRT_API int32_t mmhrt_denoiser_set_params(mmhrt_ctx ctx, const float* in, uint64_t n);
/* Denoise the last rendered frame in place on the device (it must carry MMHRT_FLAG_AUX). */
MMHRT_API int32_t mmhrt_denoise(mmhrt_ctx ctx, float depth_scale);
/* Denoise host features [b][h][w][12] fp16 into out_color [b][h][w][3] fp32. */
MMHRT_API int32_t mmhrt_denoise_features(mmhrt_ctx ctx, const uint16_t* feats, uint32_t b, uint32_t h, uint32_t w,
                                         float depth_scale, const float bg[4], float* out_color);
/* One training step on a batch: loss = mean squared error of the composed colour against
 * target [b][h][w][3]. grads_out (NULL to skip) receives d(loss)/d(params) before the update;
 * apply = 0 computes the gradients only; otherwise Adam (0.9, 0.999, 1e-8) with step lr. */
MMHRT_API int32_t mmhrt_denoiser_train_step(mmhrt_ctx ctx, const uint16_t* feats, const float* target, uint32_t b,
                                            uint32_t h, uint32_t w, float depth_scale, const float bg[4], float lr,
                                            int32_t apply, float* loss_out, float* grads_out);

/* Tuning knobs (autotuner output): "stack_lds" (LDS stack entries per lane: 8, 16, 24 or 32, default 16),
 * "tile_w", "tile_h", "persistent" (0/1), "refill" (ballot refill threshold in lanes; -1, the
 * default: 16 for 1-spp frames, 0 otherwise), "blocks_per_cu" (persistent
 * workgroups per CU), "refit_period" (MMHRT_BUILD_AUTO: refits between rebuilds), "build_chunk"
 * (leaves per LBVH chunk workgroup: 1024, 512 or 256), "build_ls" (level-synchronous chunk
 * pass, default 1; 0: one climbing thread per leaf), "build_grid" (test hook: at most this many
 * persistent chunk workgroups, 0 = as many as fit), "sort_ri" (radix sort keys per thread per
 * tile: 8, 16 or 32, default 16), "sort_onesweep" (single-pass radix passes with look-back, default 0),
 * "sort_digit" (radix digit width: 8, the default, or 0 for wide digits: three 10-bit passes over
 * 30-bit keys, six 11-bit passes over 63-bit keys),
 * "leaf_pairs" (default 1: nodes over two leaves are not stored, their parents reference the
 * pair; 0: every node is stored), "build_tail" (straddler worklist size from which one
 * workgroup finishes the tree, default 512; used when "build_span" is 0), "build_span" (straddlers
 * merged by one workgroup per span of this many chunks before the tail, default 256; 0: kernel
 * rounds), "build_wave" (chunk pass: once a round would merge at most this many splits, the
 * rest of the chunk climbs without rounds, one lane per split; default 64, one wavefront; up to
 * the chunk size; 0: rounds to the top),
 * "accel" (structure the next build makes: 0 LBVH, the default; 1 treelet P-k-d trees, about one
 * byte per sphere beyond the spheres instead of up to 64), "treelet" (spheres per treelet of that
 * structure: 256, 512 or 1024, default 1024), "batch_launch" (mmhrt_render_batch: 1, the
 * default, one trace launch per mode over all timesteps; 0 one launch per timestep),
 * "sample_lanes" (samples of a pixel on the lanes for power-of-two spp,
 * default 1), "dn_k32" (denoiser forward on the 16x16x32 MFMA kernels, default 1),
 * "dn_conv_shape" (test hook: forced launch shape of those kernels, 0 = automatic),
 * "fail_alloc" (test hook: the k-th device allocation from now fails with MMHRT_E_OOM, as a real
 * failure would, and the option returns to 0; 0, the default, injects nothing).
 * Diagnostics: "count_visits" (0/1) and, read-only in effect, "stack_overflow", "pair_visits",
 * "leaf_fetches" (refreshed by mmhrt_get_counters), "sort_fault", and "leaf_pair_count" (get
 * only: the leaf pairs of the current tree, counted on the device from its sorted keys). */
MMHRT_API int32_t mmhrt_set_option(mmhrt_ctx ctx, const char* key, int64_t value);
MMHRT_API int32_t mmhrt_get_option(mmhrt_ctx ctx, const char* key, int64_t* value);

#ifdef __cplusplus
}
#endif
#endif /* MMHRT_H */
