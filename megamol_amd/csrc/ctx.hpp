// ctx.hpp — context state of the mmhrt library and the launchers each .hip unit exports.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "mmhrt.h"
#include "rt_math.hpp"

namespace mmhrt {
namespace dn {
struct Denoiser;
}

// Counters of the straddler rounds of an LBVH build (lbvh.hip): completion state, the worklist
// size of each of the 32 kernel rounds and their successor, the rounds run, the generation.
constexpr int BOX_ROUNDS = 32;
constexpr int BOX_SLOTS = BOX_ROUNDS + 4;

// BVH2 node, 64 bytes: both child boxes in one node so that one visit tests two children
// with four 16-byte loads. child ref >= 0: internal node index; < 0: leaf ~k (k-th sphere in
// Morton order), or with the side's pair flag in d.z (bit 0: child 0, bit 1: child 1) the leaf
// pair k, k + 1, whose node is not stored. w of d holds the subtree height (leaves have height 0).
struct __align__(16) Node {
    float4 a;  // c0.lo.x c0.lo.y c0.lo.z c0.hi.x
    float4 b;  // c0.hi.y c0.hi.z c1.lo.x c1.lo.y
    float4 c;  // c1.lo.z c1.hi.x c1.hi.y c1.hi.z
    int4 d;    // child0, child1, pair flags, height
};

// Scene facts computed on the device by the bounds reduction.
struct SceneInfo {
    float clo[3], chi[3];  // centre bounds
    float blo[3], bhi[3];  // sphere bounds (centre -+ radius)
    float absmax;          // max over spheres of max(|x|,|y|,|z|) + r
    float scale;           // Morton cells per world unit
    uint32_t height;       // root height after a build
    uint32_t pair_h;       // set before a hierarchy pass: 1 = leaf pairs are not stored, 0 = every node is
};

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;  // elements
};

// The structure of one timestep of a batched series, kept until the single trace launch over all
// timesteps has read it: the LBVH (or treelets' top tree) nodes, leaf spheres, scene facts and, for
// treelets, the split axes, treelet boxes and radii; `id` holds the leaf order of one rebuild (its
// refits share it).
struct SeriesSlot {
    DevBuf<Node> nodes;
    DevBuf<float4> leaf;
    DevBuf<uint32_t> id;
    DevBuf<SceneInfo> info;
    DevBuf<uint8_t> dim;
    DevBuf<float> box, rmax;
};

struct Timers {
    hipEvent_t ev[12];
    bool init = false;
};

}  // namespace mmhrt

struct mmhrt_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::string noalloc_err;  // why mmhrt_render_noalloc cannot run on this device ("" = it can)
    int cu_count = 0;

    // scene (original particle order)
    uint64_t n = 0;
    mmhrt::DevBuf<float4> sph, col;
    mmhrt::DevBuf<uint8_t> staging;
    uint64_t data_hash = 0;
    uint32_t frame_id = 0;

    // transfer function (AddParticleColors semantics)
    std::vector<float> tf;
    float tf_min = 1.0f, tf_max = 0.0f;
    mmhrt::DevBuf<float> tf_buf;

    // LBVH
    int bits = 30;
    bool built = false;
    uint64_t built_n = 0;
    mmhrt::DevBuf<uint32_t> ka, kb;  // 32-bit key ping-pong
    mmhrt::DevBuf<uint64_t> kla, klb;  // 64-bit key ping-pong
    mmhrt::DevBuf<uint32_t> va, vb;  // values ping-pong
    mmhrt::DevBuf<uint32_t> hist, scan_tmp;
    mmhrt::DevBuf<uint32_t> sort_fault;  // set by a single-pass radix pass whose look-back wait ran out
    mmhrt::DevBuf<uint64_t> pair_count;  // instrumentation: leaf pairs of the current tree
    mmhrt::DevBuf<float4> leaf_sph;
    mmhrt::DevBuf<uint32_t> leaf_id;
    mmhrt::DevBuf<mmhrt::Node> nodes;
    mmhrt::DevBuf<uint32_t> flags;   // arrivals at the global meeting places (generation numbers)
    mmhrt::DevBuf<uint32_t> flags2;  // the chunk pass's arrivals at chunk boundaries, per side (generation numbers)
    uint32_t build_gen = 0;          // generation number of the last hierarchy pass
    bool bvh_pairs = true;           // the current tree may reference leaf pairs (built with leaf_pairs on)
    uint32_t* rounds_host = nullptr; // pinned: the last build's straddler counters (BOX_SLOTS words)
    mmhrt::DevBuf<mmhrt::Node> meet;     // per split position: the records of its two children (build scratch)
    mmhrt::DevBuf<int4> meet_far;  // per split position: the children's far range ends and parent deltas
    mmhrt::DevBuf<uint32_t> xlist;  // span pass: splits exported to the tail (two halves)
    mmhrt::DevBuf<uint32_t> box_cnt;  // worklist sizes of the bottom-up box rounds
    mmhrt::DevBuf<float> partial;
    mmhrt::DevBuf<mmhrt::SceneInfo> info;

    // treelet P-k-d trees (accel 1): leaf_sph / leaf_id hold the spheres in treelet heap order,
    // nodes the top tree of boxes over the treelets
    int accel = 0;                    // structure of the current build: 0 LBVH, 1 treelet P-k-d trees
    uint32_t pkd_M = 0, pkd_T = 0;    // treelet size (power of two) and count of the current build
    uint32_t pkd_L = 0;               // leaves of the top tree (a power of two >= max(T, 2))
    mmhrt::DevBuf<uint8_t> pkd_dim;   // split axis of every node
    mmhrt::DevBuf<float> tl_box;      // T x 6: sphere bounds of each treelet
    mmhrt::DevBuf<float> tl_rmax;     // T: largest radius of each treelet
    mmhrt::SceneInfo info_host{};
    bool info_valid = false;

    // frame
    uint32_t fb_w = 0, fb_h = 0;            // size of the last rendered window (the stored frame)
    uint32_t fb_x0 = 0, fb_y0 = 0;          // its origin in the full frame
    uint32_t fb_full_w = 0, fb_full_h = 0;  // the full frame it was cut from
    mmhrt::DevBuf<float4> fb_col;
    mmhrt::DevBuf<float> fb_depth, fb_t;
    mmhrt::DevBuf<int32_t> fb_id;
    mmhrt::DevBuf<uint32_t> fb_rgba8;
    mmhrt::DevBuf<__half> fb_aux;
    mmhrt::DevBuf<int32_t> spill;
    mmhrt::DevBuf<uint32_t> counters;  // [0] tile counter, [2..7] visit counters (u64 pairs)

    // options
    int64_t opt_stack_lds = 16;  // 16 leaves the LDS for 7 wavefronts per SIMD
    int64_t opt_persistent = 1;
    int64_t opt_refill = -1;  // ballot-refill threshold; -1: 16 for 1-spp frames (measured 13% faster), else off
    int64_t opt_count = 0;
    int64_t opt_blocks_per_cu = 8;
    int64_t opt_tile_w = 8, opt_tile_h = 8;
    int64_t opt_refit_period = 4;  // MMHRT_BUILD_AUTO: refits between rebuilds
    int64_t opt_dn_k32 = 1;        // denoiser forward with the 16x16x32 MFMA kernels (0: 16x16x16)
    int64_t opt_dn_conv_shape = 0; // forced k32 launch shape (test hook), 0 = automatic
    int64_t opt_build_chunk = 1024; // leaves per LBVH chunk workgroup: 1024, 512 or 256
    int64_t opt_sample_lanes = 1;   // sample-parallel traversal for power-of-two spp (Lambert/AO)
    int64_t opt_build_ls = 1;       // level-synchronous LBVH chunk pass (0: one climbing thread per leaf)
    int64_t opt_sort_ri = 16;       // radix sort keys per thread per tile: 8, 16 or 32 (32 for 32-bit keys only)
    int64_t opt_sort_onesweep = 0;  // single-pass radix passes with decoupled look-back (0: histogram + scans + scatter)
    int64_t opt_sort_digit = 8;     // radix digit width: 8, or 0 for wide digits (10 bits for 30-bit keys, 11 for 63-bit)
    int64_t opt_build_grid = 0;     // test hook: cap on the persistent chunk workgroups (0 = as many as fit)
    int64_t opt_leaf_pairs = 1;     // nodes over two leaves are not stored (their parents reference the pair)
    int64_t opt_build_tail = 512;   // straddler worklist size from which one workgroup finishes the tree
    int64_t opt_build_wave = 64;    // chunk pass: from this many merges on, lanes climb instead of rounds (0: never)
    int64_t opt_sort_xcd = 1;       // radix scatter: each XCD takes one contiguous range of tiles (0: round robin)
    int64_t opt_build_span = 256;   // straddlers: one workgroup per span of this many chunks, then the tail (0: rounds)
    int64_t opt_fail_alloc = 0;     // test hook: the k-th device allocation from now fails with E_OOM (0: none)
    int64_t opt_accel = 0;          // structure the next build makes: 0 LBVH, 1 treelet P-k-d trees
    int64_t opt_treelet = 1024;     // spheres per treelet of the P-k-d structure: 256, 512 or 1024
    int64_t opt_batch_launch = 1;   // batched series: 1 one trace launch over all timesteps (per mode), 0 one per timestep
    int64_t refits_since_build = 0;
    int64_t stack_overflow = 0;  // set by mmhrt_get_counters from the trace kernel's flag
    int64_t pair_visits = 0;     // set by mmhrt_get_counters: visits of leaf pairs (of the node visits)
    int64_t leaf_fetches = 0;    // set by mmhrt_get_counters: sphere loads of leaf children of stored nodes

    // batched time series: all timesteps resident, one hipGraph for the whole sequence
    mmhrt::DevBuf<float4> ts_sph;      // timesteps x n spheres
    mmhrt::DevBuf<float4> ts_col;      // n colours shared by all timesteps
    uint32_t ts_count = 0;
    uint64_t ts_n = 0;
    const float4* sph_cur = nullptr;   // spheres the next build reads (a timestep slice), null = sph
    bool capturing = false;            // stream capture in progress: no event records
    mmhrt::DevBuf<float4> bt_col;
    mmhrt::DevBuf<float> bt_depth, bt_t;
    mmhrt::DevBuf<int32_t> bt_id;
    mmhrt::DevBuf<uint32_t> bt_rgba8, batch_ctr, batch_prefix;
    mmhrt::DevBuf<uint8_t> batch_jobs;
    std::vector<mmhrt::SeriesSlot> ts_slots;  // per timestep but the last (one-launch series)

    // learned denoiser (created on first use)
    mmhrt::dn::Denoiser* dn = nullptr;
    bool last_aux = false;
    float last_bg[4] = {0, 0, 0, 1};

    // timing
    mmhrt::Timers tm;
    float t_build[5] = {0, 0, 0, 0, 0};  // morton(+bounds), sort, hierarchy, boxes, total
    float t_render[3] = {0, 0, 0};       // trace, resolve, total
    uint64_t last_visits = 0, last_tests = 0, last_rays = 0;
};

namespace mmhrt {
// error helper: records message, returns code
int fail(mmhrt_ctx_s* c, int code, const std::string& msg);
int hip_fail(mmhrt_ctx_s* c, hipError_t e, const char* what);

template <typename T>
int ensure(mmhrt_ctx_s* c, DevBuf<T>& b, size_t n, const char* what) {
    if (b.cap >= n && b.p) return 0;
    if (c->opt_fail_alloc > 0 && --c->opt_fail_alloc == 0)  // injected failure (test hook "fail_alloc")
        return fail(c, 3, std::string("device allocation failed: ") + what + " (injected)");
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        if (e != hipSuccess) return hip_fail(c, e, what);
    }
    size_t bytes = (n ? n : 1) * sizeof(T);
    hipError_t e = hipMalloc((void**)&b.p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return fail(c, 3, std::string("device allocation failed: ") + what + " (" + std::to_string(bytes) + " bytes)");
    }
    b.cap = n ? n : 1;
    return 0;
}
template <typename T>
void release(DevBuf<T>& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// api.hip
int check_ctx(mmhrt_ctx_s* c);
int check_render_args(mmhrt_ctx_s* c, const mmhrt_camera* cam, const mmhrt_render_params* p, const char* who);
void frame_consts(const mmhrt_camera* cam, const mmhrt_render_params* p, FrameConsts& f);

// ingest.hip
struct ListDesc {
    const uint8_t* vert;
    const uint8_t* col;
    uint64_t vstride, cstride, count, out_offset;
    int32_t vtype, ctype;
    float gradius;
    float gcol[4];
    int32_t use_tf;
};
int launch_convert_list(mmhrt_ctx_s* c, const ListDesc& d, const float* tf_dev, uint32_t tf_n, float tf_vmin,
                        float tf_fac);
int launch_copy_packed(mmhrt_ctx_s* c, const float4* xyzr, const float4* rgba, uint64_t n, float4 gcol);
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, uint64_t n);
void launch_fill_f4(hipStream_t s, float4* p, float4 v, uint64_t n);

// lbvh.hip
int build_lbvh(mmhrt_ctx_s* c, int bits);
int refit_lbvh(mmhrt_ctx_s* c);
int reduce_scene(mmhrt_ctx_s* c);
int count_leaf_pairs(mmhrt_ctx_s* c, int64_t* out);

int sort_spheres(mmhrt_ctx_s* c, int bits);

// pkd.hip
int build_pkd(mmhrt_ctx_s* c, int bits);

// trace.hip
constexpr int TRACE_STACK_CAP = 128;                       // traversal stack entries per lane
constexpr int TRACE_SPILL_PER_LANE = TRACE_STACK_CAP - 8;  // the most any LDS stack depth (>= 8) spills
constexpr uint32_t TRACE_CTR_WORDS = 16;                   // tile counter, overflow flag, 5 u64 counters
struct FrameOutputs {
    float4* color;
    float* depth;
    float* t;
    int32_t* hit_id;
    uint32_t* rgba8;
    __half* aux;
};
int render_frame(mmhrt_ctx_s* c, const FrameConsts& fc);
int render_frame_into(mmhrt_ctx_s* c, const FrameConsts& fc, const FrameOutputs& out, uint32_t* ctr, int32_t* spill,
                      uint32_t lanes, bool count);
uint32_t tile_log2(int64_t tw);

// noalloc.hip
int noalloc_init(mmhrt_ctx_s* c);
int render_noalloc(mmhrt_ctx_s* c, const FrameConsts& fc, const FrameOutputs& out);
size_t noalloc_scratch_bytes();
size_t batch_job_size();
int batch_fill_job(mmhrt_ctx_s* c, const FrameConsts& fc, size_t out_offset, void* dst);
void launch_trace_batch(mmhrt_ctx_s* c, const void* jobs, const uint32_t* prefix, uint32_t n_jobs, uint32_t* ctr,
                        int mode, bool treelets);

// event record on the context stream, skipped while the stream is being captured into a graph
inline void rec(mmhrt_ctx_s* c, int i) {
    if (!c->capturing) (void)hipEventRecord(c->tm.ev[i], c->stream);
}
// spheres the build reads: the current timestep slice, or the context's spheres
inline const float4* cur_spheres(const mmhrt_ctx_s* c) { return c->sph_cur ? c->sph_cur : c->sph.p; }
}  // namespace mmhrt
