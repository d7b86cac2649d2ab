// trace.hip — ray traversal and shading, replacing optixLaunch of raygen_program
// (plugins/optix_hpg/src/optix/AbstractRenderer.cpp:106, cuda_resources/raygen.cu:79-278) and the
// intersection / closest-hit programs (cuda_resources/sphere.cuh:43-57, 94-131).
//
// Launch shape: a persistent grid (blocks_per_cu x CUs blocks of 256 threads). Each wavefront
// takes 8x8-pixel tiles from a global atomic counter (work stealing) until none are left; a
// lane owns one pixel and loops over its samples in order, so the per-pixel random stream is the
// reference's one (TEA seed, two LCG draws per sample for the jitter).
//
// Traversal: BVH2 with both child boxes in the 64-byte node (four 16-byte loads per visit),
// ordered front-to-back, leaves tested inline with one 16-byte sphere load. The traversal stack
// keeps its top S entries in LDS (lane-strided, conflict free) and spills deeper entries to a
// per-lane global region. Closest hit keeps the lowest particle ID among equal t; boxes whose
// entry distance equals the current best are still visited so that ties resolve the same way
// whatever the visiting order. AO rays stop at the first occluder (any hit).
#include <float.h>

#include "ctx.hpp"

namespace mmhrt {
namespace {

constexpr int TB = 256;         // threads per block
constexpr int STACK_CAP = 128;  // LDS + spill entries per lane (LBVH height <= 63 + 31)

struct TraceArgs {
    const Node* nodes;
    const float4* leaf_sph;
    const uint32_t* leaf_id;
    const float4* col;
    const SceneInfo* info;
    int32_t* spill;
    uint32_t spill_stride;  // lanes in the grid
    uint32_t* ctr;          // [0] tile counter, [1] overflow flag, [2..] u64 counters
    uint32_t n_tiles, tiles_x;
    float4* out_col;
    float* out_depth;
    float* out_t;
    int32_t* out_id;
    uint32_t* out_rgba8;
    __half* out_aux;
    FrameConsts fc;
};

template <int S>
struct LaneStack {
    int32_t* lds;  // element k at lds[k * TB]
    int32_t* spill;
    uint32_t stride;
    int sp;
    uint32_t* overflow;
    __device__ __forceinline__ void push(int32_t v) {
        if (sp < S) {
            lds[sp * TB] = v;
        } else if (sp < STACK_CAP) {
            spill[(size_t)(sp - S) * stride] = v;
        } else {
            *overflow = 1u;  // cannot happen for a valid LBVH; recorded, entry dropped
            return;
        }
        ++sp;
    }
    __device__ __forceinline__ int32_t pop() {
        --sp;
        return sp < S ? lds[sp * TB] : spill[(size_t)(sp - S) * stride];
    }
};

struct RayPre {
    V3 o, d, inv, oil, oih;
};

__device__ __forceinline__ RayPre prep(V3 o, V3 d, float pad) {
    RayPre r;
    r.o = o;
    r.d = d;
    float dx = fabsf(d.x) < 1e-30f ? __builtin_copysignf(1e-30f, d.x) : d.x;
    float dy = fabsf(d.y) < 1e-30f ? __builtin_copysignf(1e-30f, d.y) : d.y;
    float dz = fabsf(d.z) < 1e-30f ? __builtin_copysignf(1e-30f, d.z) : d.z;
    r.inv = V3{1.0f / dx, 1.0f / dy, 1.0f / dz};
    // box widened by pad: (lo - pad - o) * inv = lo * inv - (o + pad) * inv
    r.oil = V3{(o.x + pad) * r.inv.x, (o.y + pad) * r.inv.y, (o.z + pad) * r.inv.z};
    r.oih = V3{(o.x - pad) * r.inv.x, (o.y - pad) * r.inv.y, (o.z - pad) * r.inv.z};
    return r;
}

// Slab test of one child box (lo, hi); returns the entry distance, or +inf when missed.
__device__ __forceinline__ float slab(const RayPre& r, float lx, float ly, float lz, float hx, float hy, float hz,
                                      float tmin, float tmax) {
    float x0 = __builtin_fmaf(lx, r.inv.x, -r.oil.x), x1 = __builtin_fmaf(hx, r.inv.x, -r.oih.x);
    float y0 = __builtin_fmaf(ly, r.inv.y, -r.oil.y), y1 = __builtin_fmaf(hy, r.inv.y, -r.oih.y);
    float z0 = __builtin_fmaf(lz, r.inv.z, -r.oil.z), z1 = __builtin_fmaf(hz, r.inv.z, -r.oih.z);
    float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
    float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
    return tn <= tf ? tn : INFINITY;
}

struct Count {
    uint64_t visits = 0, tests = 0, rays = 0;
};

// Closest hit: returns the leaf index of the nearest reported intersection (or -1) and its t.
template <int S, bool COUNT>
__device__ int trace_closest(const TraceArgs& A, const RayPre& r, float tmin, float tmax, float& t_best,
                             LaneStack<S>& st, Count& cnt) {
    int best = -1;
    t_best = tmax;
    int32_t node = 0;
    st.sp = 0;
    for (;;) {
        const Node* nd = A.nodes + node;
        float4 a = nd->a, b = nd->b, c = nd->c;
        int4 dd = nd->d;
        if (COUNT) cnt.visits++;
        float tn0 = slab(r, a.x, a.y, a.z, a.w, b.x, b.y, tmin, t_best);
        float tn1 = slab(r, b.z, b.w, c.x, c.y, c.z, c.w, tmin, t_best);
        int32_t ch0 = dd.x, ch1 = dd.y;
        if (tn0 != INFINITY && ch0 < 0) {
            int leaf = ~ch0;
            float t = sphere_t(r.o, r.d, A.leaf_sph[leaf], tmin, tmax);
            if (COUNT) cnt.tests++;
            if (t >= 0.0f) {
                if (t < t_best) {
                    t_best = t;
                    best = leaf;
                } else if (t == t_best && best >= 0 && A.leaf_id[leaf] < A.leaf_id[best]) {
                    best = leaf;
                }
            }
            tn0 = INFINITY;
        }
        if (tn1 != INFINITY && ch1 < 0) {
            int leaf = ~ch1;
            float t = sphere_t(r.o, r.d, A.leaf_sph[leaf], tmin, tmax);
            if (COUNT) cnt.tests++;
            if (t >= 0.0f) {
                if (t < t_best) {
                    t_best = t;
                    best = leaf;
                } else if (t == t_best && best >= 0 && A.leaf_id[leaf] < A.leaf_id[best]) {
                    best = leaf;
                }
            }
            tn1 = INFINITY;
        }
        bool h0 = tn0 <= t_best, h1 = tn1 <= t_best;
        if (h0 && h1) {
            bool first0 = tn0 <= tn1;
            st.push(first0 ? ch1 : ch0);
            node = first0 ? ch0 : ch1;
        } else if (h0) {
            node = ch0;
        } else if (h1) {
            node = ch1;
        } else {
            if (st.sp == 0) break;
            node = st.pop();
        }
    }
    return best;
}

// Any hit in (tmin, tmax) ignoring leaf `skip`.
template <int S, bool COUNT>
__device__ bool trace_any(const TraceArgs& A, const RayPre& r, float tmin, float tmax, int skip, LaneStack<S>& st,
                          Count& cnt) {
    int32_t node = 0;
    st.sp = 0;
    for (;;) {
        const Node* nd = A.nodes + node;
        float4 a = nd->a, b = nd->b, c = nd->c;
        int4 dd = nd->d;
        if (COUNT) cnt.visits++;
        float tn0 = slab(r, a.x, a.y, a.z, a.w, b.x, b.y, tmin, tmax);
        float tn1 = slab(r, b.z, b.w, c.x, c.y, c.z, c.w, tmin, tmax);
        int32_t ch0 = dd.x, ch1 = dd.y;
        bool h0 = tn0 != INFINITY, h1 = tn1 != INFINITY;
        if (h0 && ch0 < 0) {
            int leaf = ~ch0;
            if (leaf != skip) {
                if (COUNT) cnt.tests++;
                if (sphere_t(r.o, r.d, A.leaf_sph[leaf], tmin, tmax) >= 0.0f) return true;
            }
            h0 = false;
        }
        if (h1 && ch1 < 0) {
            int leaf = ~ch1;
            if (leaf != skip) {
                if (COUNT) cnt.tests++;
                if (sphere_t(r.o, r.d, A.leaf_sph[leaf], tmin, tmax) >= 0.0f) return true;
            }
            h1 = false;
        }
        if (h0 && h1) {
            bool first0 = tn0 <= tn1;
            st.push(first0 ? ch1 : ch0);
            node = first0 ? ch0 : ch1;
        } else if (h0) {
            node = ch0;
        } else if (h1) {
            node = ch1;
        } else {
            if (st.sp == 0) return false;
            node = st.pop();
        }
    }
}

// Diffuse bounce direction of raygen.cu:71-77,149-152: rejection sample in the unit cube,
// the three draws taken x, y, z in order.
__device__ __forceinline__ V3 random_in_unit_sphere(uint32_t& seed) {
    V3 p;
    do {
        float rx = rnd(seed);
        float ry = rnd(seed);
        float rz = rnd(seed);
        p = V3{2.0f * rx - 1.0f, 2.0f * ry - 1.0f, 2.0f * rz - 1.0f};
    } while (dot(p, p) >= 1.0f);
    return p;
}

template <int S, int MODE, bool COUNT>
__device__ void shade_pixel(const TraceArgs& A, uint32_t px, uint32_t py, float pad, LaneStack<S>& st, Count& cnt) {
    const FrameConsts& f = A.fc;
    uint32_t pix = py * f.width + px;  // full-frame index: the random stream of the pixel
    uint32_t opix = (py - f.win_y0) * f.win_w + (px - f.win_x0);  // index in the output window
    uint32_t seed = tea16(pix, f.frame_idx);
    float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
    float depth = FLT_MAX;
    float best_t = INFINITY;
    int best_leaf = -1;
    float aux[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float vis_sum = 0.0f;
    for (uint32_t s = 0; s < f.spp; ++s) {
        float jx = rnd(seed);
        float jy = rnd(seed);
        V3 o, d;
        primary_ray(f, px, py, jx, jy, o, d);
        RayPre r = prep(o, d, pad);
        float t;
        if (COUNT) cnt.rays++;
        int leaf = trace_closest<S, COUNT>(A, r, 1e-6f, 1e20f, t, st, cnt);
        float4 sc;
        float ray_depth = FLT_MAX;
        if (leaf < 0) {
            if (MODE == 1)
                sc = make_float4(f.bg[0], f.bg[1], f.bg[2], f.bg[3]);
            else
                sc = make_float4(1.0f * 0.8f, 1.0f * 0.8f, 1.0f * 0.8f, 1.0f);
        } else {
            ray_depth = t;
            float4 sp = A.leaf_sph[leaf];
            uint32_t id = A.leaf_id[leaf];
            V3 P = V3{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z};
            V3 N = V3{P.x - sp.x, P.y - sp.y, P.z - sp.z};
            if (dot(N, d) > 0.0f) N = neg(N);
            N = normalize(N);
            float4 alb = A.col[id];
            if (MODE == 1) {
                uint32_t aseed = tea16(pix, 0xA0000000u + f.frame_idx * 4096u + s);
                V3 ad = ao_direction(N, aseed);
                RayPre ra = prep(P, ad, pad);
                if (COUNT) cnt.rays++;
                bool occ = trace_any<S, COUNT>(A, ra, 0.0f, f.ao_distance, leaf, st, cnt);
                float V = occ ? 0.0f : 1.0f;
                sc = make_float4(alb.x * V, alb.y * V, alb.z * V, 1.0f);
                vis_sum += V;
            } else if (MODE == 2 && f.max_bounces > 0) {
                // raygen.cu:122-155 with maxBounces > 0
                V3 cc = V3{alb.x, alb.y, alb.z};
                V3 ro = o, rd = d;
                float rt = t;
                V3 n = N;
                float4 res;
                for (int dep = 0;; ++dep) {
                    if (dep >= f.max_bounces) {
                        res = make_float4(0.1f, 0.1f, 0.1f, 1.0f);
                        break;
                    }
                    V3 so = V3{ro.x + rt * rd.x, ro.y + rt * rd.y, ro.z + rt * rd.z};
                    V3 ru = random_in_unit_sphere(seed);
                    V3 sd = normalize(V3{n.x + ru.x, n.y + ru.y, n.z + ru.z});
                    RayPre rb = prep(so, sd, pad);
                    float tb;
                    if (COUNT) cnt.rays++;
                    int lb = trace_closest<S, COUNT>(A, rb, 1e-3f, 1e8f, tb, st, cnt);
                    if (lb < 0) {
                        res = make_float4(cc.x * 0.8f, cc.y * 0.8f, cc.z * 0.8f, 1.0f);
                        break;
                    }
                    float4 sb = A.leaf_sph[lb];
                    float4 ab = A.col[A.leaf_id[lb]];
                    V3 Pb = V3{so.x + tb * sd.x, so.y + tb * sd.y, so.z + tb * sd.z};
                    V3 Nb = V3{Pb.x - sb.x, Pb.y - sb.y, Pb.z - sb.z};
                    if (dot(Nb, sd) > 0.0f) Nb = neg(Nb);
                    n = normalize(Nb);
                    cc = V3{cc.x * ab.x, cc.y * ab.y, cc.z * ab.z};
                    ro = so;
                    rd = sd;
                    rt = tb;
                }
                sc = res;
            } else {
                float sh = 0.2f + 0.6f * fabsf(dot(N, d));
                sc = make_float4(alb.x * sh, alb.y * sh, alb.z * sh, 1.0f);
            }
            if (t < best_t) {
                best_t = t;
                best_leaf = leaf;
            }
            if (s == 0) {
                aux[0] = alb.x; aux[1] = alb.y; aux[2] = alb.z;
                aux[3] = N.x; aux[4] = N.y; aux[5] = N.z;
                aux[6] = t;
            }
        }
        col.x += sc.x;
        col.y += sc.y;
        col.z += sc.z;
        col.w += sc.w;
        depth = fminf(depth, ray_depth);
    }
    float fs = (float)f.spp;
    col = make_float4(col.x / fs, col.y / fs, col.z / fs, col.w / fs);
    if (f.accumulate && f.frame_idx > 0) {
        // raygen.cu:254-260, lerp(a, b, t) = a + t * (b - a) (utils_device.h:152-154)
        float4 old = A.out_col[opix];
        float a = 1.0f / (float)(f.frame_idx + 1);
        col = make_float4(old.x + a * (col.x - old.x), old.y + a * (col.y - old.y), old.z + a * (col.z - old.z),
                          old.w + a * (col.w - old.w));
    }
    if (depth < FLT_MAX) {
        depth = (f.depth_c / depth) - f.depth_a;
        depth = 0.5f * (depth + 1.0f);
    } else {
        depth = 1.0f;
        col = make_float4(f.bg[0], f.bg[1], f.bg[2], f.bg[3]);
    }
    A.out_col[opix] = col;
    A.out_depth[opix] = depth;
    A.out_t[opix] = best_t;
    A.out_id[opix] = best_leaf >= 0 ? (int32_t)A.leaf_id[best_leaf] : -1;
    A.out_rgba8[opix] = pack_rgba8(col);
    if (f.want_aux) {
        aux[7] = vis_sum / fs;
        uint32_t w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            w[k] = (uint32_t)__half_as_ushort(__float2half(aux[2 * k])) |
                   ((uint32_t)__half_as_ushort(__float2half(aux[2 * k + 1])) << 16);
        *reinterpret_cast<uint4*>(A.out_aux + (size_t)opix * 8) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

template <int S, int MODE, bool COUNT, bool STEAL>
__global__ __launch_bounds__(TB) void k_trace(TraceArgs A) {
    __shared__ int32_t lds_stack[S * TB];
    const uint32_t gtid = blockIdx.x * TB + threadIdx.x;
    LaneStack<S> st{lds_stack + threadIdx.x, A.spill + gtid, A.spill_stride, 0, A.ctr + 1};
    Count cnt;
    // slab padding: a bound on the rounding of the FP32 sphere and slab tests, which scale with
    // the magnitude of ray origins and sphere centres (fc.pad carries max |camera coordinate|)
    const float amax = A.info->absmax;
    const float pad = (fmaxf(A.fc.pad, amax) + amax) * (1.0f / 131072.0f);
    const uint32_t lane = __lane_id();
    const uint32_t gwave = gtid >> 6;
    const uint32_t nwaves = (gridDim.x * TB) >> 6;
    for (uint32_t it = 0;; ++it) {
        uint32_t tile;
        if (STEAL) {
            uint32_t t0 = 0;
            if (lane == 0) t0 = atomicAdd(A.ctr, 1u);
            tile = __shfl(t0, 0, 64);
        } else {
            tile = gwave + it * nwaves;
        }
        if (tile >= A.n_tiles) break;
        uint32_t tx = tile % A.tiles_x, ty = tile / A.tiles_x;
        uint32_t px = A.fc.win_x0 + tx * 8 + (lane & 7), py = A.fc.win_y0 + ty * 8 + (lane >> 3);
        if (px < A.fc.win_x0 + A.fc.win_w && py < A.fc.win_y0 + A.fc.win_h)
            shade_pixel<S, MODE, COUNT>(A, px, py, pad, st, cnt);
    }
    if (COUNT) {
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 2), (unsigned long long)cnt.visits);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 4), (unsigned long long)cnt.tests);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 6), (unsigned long long)cnt.rays);
    }
}

template <int S, int MODE, bool COUNT>
void launch3(const TraceArgs& A, unsigned grid, bool steal, hipStream_t s) {
    if (steal)
        hipLaunchKernelGGL((k_trace<S, MODE, COUNT, true>), dim3(grid), dim3(TB), 0, s, A);
    else
        hipLaunchKernelGGL((k_trace<S, MODE, COUNT, false>), dim3(grid), dim3(TB), 0, s, A);
}
template <int S, int MODE>
void launch2(const TraceArgs& A, unsigned grid, bool steal, bool count, hipStream_t s) {
    if (count)
        launch3<S, MODE, true>(A, grid, steal, s);
    else
        launch3<S, MODE, false>(A, grid, steal, s);
}
template <int S>
void launch1(const TraceArgs& A, unsigned grid, bool steal, bool count, int mode, hipStream_t s) {
    if (mode == 1)
        launch2<S, 1>(A, grid, steal, count, s);
    else if (mode == 2)
        launch2<S, 2>(A, grid, steal, count, s);
    else
        launch2<S, 0>(A, grid, steal, count, s);
}

}  // namespace

int render_frame(mmhrt_ctx_s* c, const FrameConsts& fc) {
    uint32_t W = fc.win_w, H = fc.win_h;  // output size = the rendered window
    size_t P = (size_t)W * H;
    int e = 0;
    if ((e = ensure(c, c->fb_col, P, "framebuffer colour"))) return e;
    if ((e = ensure(c, c->fb_depth, P, "framebuffer depth"))) return e;
    if ((e = ensure(c, c->fb_t, P, "framebuffer t"))) return e;
    if ((e = ensure(c, c->fb_id, P, "framebuffer id"))) return e;
    if ((e = ensure(c, c->fb_rgba8, P, "framebuffer rgba8"))) return e;
    if (fc.want_aux && (e = ensure(c, c->fb_aux, P * 8, "framebuffer aux"))) return e;
    if ((e = ensure(c, c->counters, 16, "counters"))) return e;
    unsigned grid = (unsigned)(c->cu_count * c->opt_blocks_per_cu);
    if (grid == 0) grid = 1;
    uint32_t lanes = grid * TB;
    int S = (int)c->opt_stack_lds;
    if (S != 8 && S != 16 && S != 24 && S != 32) S = 16;
    if ((e = ensure(c, c->spill, (size_t)lanes * (STACK_CAP - 8), "traversal spill"))) return e;
    c->fb_w = W;
    c->fb_h = H;
    TraceArgs A;
    A.nodes = c->nodes.p;
    A.leaf_sph = c->leaf_sph.p;
    A.leaf_id = c->leaf_id.p;
    A.col = c->col.p;
    A.info = c->info.p;
    A.spill = c->spill.p;
    A.spill_stride = lanes;
    A.ctr = c->counters.p;
    A.tiles_x = (W + 7) / 8;
    A.n_tiles = A.tiles_x * ((H + 7) / 8);
    A.out_col = c->fb_col.p;
    A.out_depth = c->fb_depth.p;
    A.out_t = c->fb_t.p;
    A.out_id = c->fb_id.p;
    A.out_rgba8 = c->fb_rgba8.p;
    A.out_aux = fc.want_aux ? c->fb_aux.p : nullptr;
    A.fc = fc;
    launch_fill_u32(c->stream, c->counters.p, 0u, 16);
    hipEventRecord(c->tm.ev[6], c->stream);
    bool steal = c->opt_persistent != 0;
    bool count = c->opt_count != 0;
    switch (S) {
    case 8: launch1<8>(A, grid, steal, count, fc.mode, c->stream); break;
    case 24: launch1<24>(A, grid, steal, count, fc.mode, c->stream); break;
    case 32: launch1<32>(A, grid, steal, count, fc.mode, c->stream); break;
    default: launch1<16>(A, grid, steal, count, fc.mode, c->stream); break;
    }
    hipEventRecord(c->tm.ev[7], c->stream);
    hipError_t he = hipGetLastError();
    if (he != hipSuccess) return hip_fail(c, he, "trace launch");
    return 0;
}

}  // namespace mmhrt
