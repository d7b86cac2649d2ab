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
// ordered front-to-back, leaves tested inline with one 16-byte sphere load; a node over two
// leaves is not stored and its visit reads the two spheres instead (leaf pairs). The traversal stack
// keeps its top S entries in LDS (lane-strided, conflict free) and spills deeper entries to a
// per-lane global region. Closest hit keeps the lowest particle ID among equal t; boxes whose
// entry distance equals the current best are still visited so that ties resolve the same way
// whatever the visiting order. AO rays stop at the first occluder (any hit).
#include <float.h>

#include <algorithm>
#include <cstring>

#include "ctx.hpp"

namespace mmhrt {
namespace {

constexpr int TB = 256;                  // threads per block
constexpr int STACK_CAP = TRACE_STACK_CAP;  // LDS + spill entries per lane (LBVH height <= 63 + 31)

struct TraceArgs {
    const Node* nodes;
    const float4* leaf_sph;
    const uint32_t* leaf_id;
    const float4* col;
    const SceneInfo* info;
    int32_t* spill;
    uint32_t spill_stride;  // lanes in the grid
    uint32_t* ctr;          // [0] tile counter, [1] overflow flag, [2..11] u64 counters
    uint32_t n_tiles, tiles_x;
    uint32_t tw_log2;  // pixel tile = (1 << tw_log2) x (64 >> tw_log2), one wavefront
    float4* out_col;
    float* out_depth;
    float* out_t;
    int32_t* out_id;
    uint32_t* out_rgba8;
    __half* out_aux;
    FrameConsts fc;
    // treelet P-k-d trees (nodes then hold the top tree over the treelets)
    const uint8_t* pkd_dim;  // split axis of the node at each position
    const float* tl_rmax;    // largest radius of each treelet
    uint32_t n_sph;          // spheres
    uint32_t tl_log2;        // log2 of the treelet size
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
    uint64_t visits = 0, tests = 0, rays = 0, pair_visits = 0, leaf_fetches = 0;
};

// The fetch of one visit. node >= 0: an internal node (four 16-byte loads). node < 0: the leaf
// pair ~node, ~node + 1, which the build does not store: its two spheres are read in the same
// four loads (each twice; the second reads hit the same cache lines) and turned in place into
// the node they stand for: child boxes the leaves' boxes (centre -+ radius, as the build
// computes them), children the two leaves. Waves without a pair skip that branch. PAIRS = false:
// the tree has none (built with leaf pairs off) and a visit is the node fetch alone.
struct Visit {
    float4 a, b, c;
    int4 d;
};
template <bool COUNT, bool PAIRS>
__device__ __forceinline__ Visit fetch(const TraceArgs& A, int32_t node, Count& cnt) {
    const bool pair = PAIRS && node < 0;
    const float4* q = pair ? A.leaf_sph + ~node : reinterpret_cast<const float4*>(A.nodes + node);
    const float4* q23 = pair ? q : q + 2;
    Visit v;
    v.a = q[0];
    v.b = q[1];
    v.c = q23[0];
    v.d = reinterpret_cast<const int4*>(q23)[1];
    if (COUNT) {
        cnt.visits++;
        if (pair) cnt.pair_visits++;
    }
    if (pair) {
        const float4 s0 = v.c, s1 = make_float4(__int_as_float(v.d.x), __int_as_float(v.d.y),
                                                __int_as_float(v.d.z), __int_as_float(v.d.w));
        v.a = make_float4(s0.x - s0.w, s0.y - s0.w, s0.z - s0.w, s0.x + s0.w);
        v.b = make_float4(s0.y + s0.w, s0.z + s0.w, s1.x - s1.w, s1.y - s1.w);
        v.c = make_float4(s1.z - s1.w, s1.x + s1.w, s1.y + s1.w, s1.z + s1.w);
        v.d = make_int4(node, node - 1, 0, 1);  // ~l, ~(l + 1)
    }
    return v;
}

// One node visit of a closest-hit ray (state: node, stack, t_best, best leaf). Returns true when
// the ray is finished. A closest-hit trace is this step repeated from the root.
template <int S, bool COUNT, bool PAIRS = true>
__device__ __forceinline__ bool step_closest(const TraceArgs& A, const RayPre& r, float tmin, float tmax, float& t_best,
                                             int& best, int32_t& node, LaneStack<S>& st, Count& cnt) {
    const bool pair_visit = PAIRS && node < 0;
    const Visit v = fetch<COUNT, PAIRS>(A, node, cnt);
    float tn0 = slab(r, v.a.x, v.a.y, v.a.z, v.a.w, v.b.x, v.b.y, tmin, t_best);
    float tn1 = slab(r, v.b.z, v.b.w, v.c.x, v.c.y, v.c.z, v.c.w, tmin, t_best);
    const int32_t ch0 = v.d.x, ch1 = v.d.y;
    // a leaf child (ref < 0 without its pair flag: bit 0 / bit 1 of d.z) is tested here; a leaf
    // pair child is descended into like an internal node. Both leaf spheres are fetched before
    // either is tested: one memory round trip (for a pair's own two, cache hits).
    const bool lf0 = tn0 != INFINITY && ch0 < 0 && !(PAIRS && (v.d.z & 1));
    const bool lf1 = tn1 != INFINITY && ch1 < 0 && !(PAIRS && (v.d.z & 2));
    float4 sp0 = make_float4(0.f, 0.f, 0.f, 0.f), sp1 = sp0;
    if (lf0) sp0 = A.leaf_sph[~ch0];
    if (lf1) sp1 = A.leaf_sph[~ch1];
    if (COUNT && !pair_visit) cnt.leaf_fetches += (lf0 ? 1 : 0) + (lf1 ? 1 : 0);
    // the hit is the least t, ties going to the least particle ID, whatever the test order
    auto test = [&](int leaf, float4 sp) {
        float t = sphere_t(r.o, r.d, sp, tmin, tmax);
        if (COUNT) cnt.tests++;
        if (t >= 0.0f) {
            if (t < t_best) {
                t_best = t;
                best = leaf;
            } else if (t == t_best && best >= 0 && A.leaf_id[leaf] < A.leaf_id[best]) {
                best = leaf;
            }
        }
    };
    if (lf0) {
        test(~ch0, sp0);
        tn0 = INFINITY;
    }
    if (lf1) {
        test(~ch1, sp1);
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
        if (st.sp == 0) return true;
        node = st.pop();
    }
    return false;
}

// One node visit of an any-hit ray ignoring leaf `skip`: 0 = continue, 1 = finished without a
// hit, 2 = occluded. Nodes, leaves and leaf pairs as in step_closest.
template <int S, bool COUNT, bool PAIRS = true>
__device__ __forceinline__ int step_any(const TraceArgs& A, const RayPre& r, float tmin, float tmax, int skip,
                                        int32_t& node, LaneStack<S>& st, Count& cnt) {
    const bool pair_visit = PAIRS && node < 0;
    const Visit v = fetch<COUNT, PAIRS>(A, node, cnt);
    float tn0 = slab(r, v.a.x, v.a.y, v.a.z, v.a.w, v.b.x, v.b.y, tmin, tmax);
    float tn1 = slab(r, v.b.z, v.b.w, v.c.x, v.c.y, v.c.z, v.c.w, tmin, tmax);
    const int32_t ch0 = v.d.x, ch1 = v.d.y;
    bool h0 = tn0 != INFINITY, h1 = tn1 != INFINITY;
    const bool leaf0 = ch0 < 0 && !(PAIRS && (v.d.z & 1)), leaf1 = ch1 < 0 && !(PAIRS && (v.d.z & 2));
    const bool lf0 = h0 && leaf0 && ~ch0 != skip, lf1 = h1 && leaf1 && ~ch1 != skip;
    float4 sp0 = make_float4(0.f, 0.f, 0.f, 0.f), sp1 = sp0;
    if (lf0) sp0 = A.leaf_sph[~ch0];
    if (lf1) sp1 = A.leaf_sph[~ch1];
    if (COUNT && !pair_visit) cnt.leaf_fetches += (lf0 ? 1 : 0) + (lf1 ? 1 : 0);
    if (h0 && leaf0) {
        if (lf0) {
            if (COUNT) cnt.tests++;
            if (sphere_t(r.o, r.d, sp0, tmin, tmax) >= 0.0f) return 2;
        }
        h0 = false;
    }
    if (h1 && leaf1) {
        if (lf1) {
            if (COUNT) cnt.tests++;
            if (sphere_t(r.o, r.d, sp1, tmin, tmax) >= 0.0f) return 2;
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
        if (st.sp == 0) return 1;
        node = st.pop();
    }
    return 0;
}

// One node visit of either kind of ray, for wavefronts that hold both: a closest-hit ray (ao
// false; t_best its current bound, as in step_closest) or an any-hit ray (ao true; t_best holds
// its tmax and is never lowered, leaf `skip` is ignored, as in step_any). Both kinds run one
// instruction stream, where a lane-divergent choice between the two steps ran both bodies with
// half the lanes masked off. The arithmetic, the visit order and the tie rule are those of the
// two steps, so the results are bit-identical. Returns 0 to continue, 1 when the ray is finished
// (for any-hit: without a hit), 2 for an any-hit ray's hit.
template <int S, bool COUNT, bool PAIRS = true>
__device__ __forceinline__ int step_mixed(const TraceArgs& A, const RayPre& r, float tmin, float tmax, float& t_best,
                                          int& best, int32_t& node, LaneStack<S>& st, Count& cnt, int skip, bool ao) {
    const bool pair_visit = PAIRS && node < 0;
    const Visit v = fetch<COUNT, PAIRS>(A, node, cnt);
    float tn0 = slab(r, v.a.x, v.a.y, v.a.z, v.a.w, v.b.x, v.b.y, tmin, t_best);
    float tn1 = slab(r, v.b.z, v.b.w, v.c.x, v.c.y, v.c.z, v.c.w, tmin, t_best);
    const int32_t ch0 = v.d.x, ch1 = v.d.y;
    const bool leaf0 = ch0 < 0 && !(PAIRS && (v.d.z & 1)), leaf1 = ch1 < 0 && !(PAIRS && (v.d.z & 2));
    const bool lf0 = tn0 != INFINITY && leaf0 && ~ch0 != skip, lf1 = tn1 != INFINITY && leaf1 && ~ch1 != skip;
    float4 sp0 = make_float4(0.f, 0.f, 0.f, 0.f), sp1 = sp0;
    if (lf0) sp0 = A.leaf_sph[~ch0];
    if (lf1) sp1 = A.leaf_sph[~ch1];
    if (COUNT && !pair_visit) cnt.leaf_fetches += (lf0 ? 1 : 0) + (lf1 ? 1 : 0);
    bool hit = false;
    auto test = [&](int leaf, float4 sp) {
        const float t = sphere_t(r.o, r.d, sp, tmin, tmax);
        if (COUNT) cnt.tests++;
        if (t >= 0.0f) {
            if (ao) {
                hit = true;
            } else if (t < t_best) {
                t_best = t;
                best = leaf;
            } else if (t == t_best && best >= 0 && A.leaf_id[leaf] < A.leaf_id[best]) {
                best = leaf;
            }
        }
    };
    if (lf0) test(~ch0, sp0);
    if (hit) return 2;
    if (lf1) test(~ch1, sp1);
    if (hit) return 2;
    // leaf children are never descended (a skipped leaf neither)
    if (leaf0) tn0 = INFINITY;
    if (leaf1) tn1 = INFINITY;
    const bool h0 = tn0 <= t_best, h1 = tn1 <= t_best;
    if (h0 && h1) {
        const bool first0 = tn0 <= tn1;
        st.push(first0 ? ch1 : ch0);
        node = first0 ? ch0 : ch1;
    } else if (h0) {
        node = ch0;
    } else if (h1) {
        node = ch1;
    } else {
        if (st.sp == 0) return 1;
        node = st.pop();
    }
    return 0;
}

// Closest hit: returns the leaf index of the nearest reported intersection (or -1) and its t.
template <int S, bool COUNT, bool PAIRS = true>
__device__ int trace_closest(const TraceArgs& A, const RayPre& r, float tmin, float tmax, float& t_best,
                             LaneStack<S>& st, Count& cnt) {
    int best = -1;
    t_best = tmax;
    int32_t node = 0;
    st.sp = 0;
    while (!step_closest<S, COUNT, PAIRS>(A, r, tmin, tmax, t_best, best, node, st, cnt)) {
    }
    return best;
}

// Any hit in (tmin, tmax) ignoring leaf `skip`.
template <int S, bool COUNT, bool PAIRS = true>
__device__ bool trace_any(const TraceArgs& A, const RayPre& r, float tmin, float tmax, int skip, LaneStack<S>& st,
                          Count& cnt) {
    int32_t node = 0;
    st.sp = 0;
    int res;
    while ((res = step_any<S, COUNT, PAIRS>(A, r, tmin, tmax, skip, node, st, cnt)) == 0) {
    }
    return res == 2;
}

// Slab test of one child box returning the entry distance (+inf when missed) and the exit.
__device__ __forceinline__ float slab_io(const RayPre& r, float lx, float ly, float lz, float hx, float hy, float hz,
                                         float tmin, float tmax, float& tf_out) {
    float x0 = __builtin_fmaf(lx, r.inv.x, -r.oil.x), x1 = __builtin_fmaf(hx, r.inv.x, -r.oih.x);
    float y0 = __builtin_fmaf(ly, r.inv.y, -r.oil.y), y1 = __builtin_fmaf(hy, r.inv.y, -r.oih.y);
    float z0 = __builtin_fmaf(lz, r.inv.z, -r.oil.z), z1 = __builtin_fmaf(hz, r.inv.z, -r.oih.z);
    float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
    float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
    tf_out = tf;
    return tn <= tf ? tn : INFINITY;
}

// Treelet P-k-d traversal (the interval logic of pkd_intersect, plugins/optix_owl/cuda_resources/
// pkdPrograms.cu:45-159, under a top tree of boxes like the LBVH's). A reference is a top node
// (>= 0) or the P-k-d node at position p (~p): the node's sphere is tested when the ray's
// interval [t0, t1] meets the node's slab (split coordinate -+ the treelet's largest radius,
// widened by the padding every slab test here carries); the near child takes
// [t0, min(slab exit, t1, best)], the far child [max(t0, slab entry), min(t1, best)]. Closed
// intervals, so that equal-t hits still reach the lowest particle id. Before its exact test a
// sphere must also meet the ray in its own box. A pushed P-k-d entry keeps
// its entry distance (two stack words: t0 below the reference) and resumes with t1 = the best
// hit at the time it is popped. A treelet is entered at its root over its box's [entry, exit].
// Closest hit (ANY false): returns the position of the hit (or -1) and its t in t_best. ANY:
// returns 1 at the first hit of a sphere other than position `skip`, else 0.
template <int S, bool COUNT, bool ANY>
__device__ int trace_tl(const TraceArgs& A, const RayPre& r, float tmin, float tmax, int skip, float& t_best,
                        LaneStack<S>& st, Count& cnt) {
    int best = -1;
    t_best = tmax;
    st.sp = 0;
    int32_t node = 0;
    float t0 = 0.0f, t1 = 0.0f;
    uint32_t cur_tl = 0xffffffffu, m = 0;
    float rmax = 0.0f;
    const uint32_t lg = A.tl_log2;
    const uint32_t msk = (1u << lg) - 1u;
    for (;;) {
        bool pop = false;
        if (node >= 0) {
            const float4* q = reinterpret_cast<const float4*>(A.nodes + node);
            const float4 a = q[0], b = q[1], c = q[2];
            const int4 d = reinterpret_cast<const int4*>(q)[3];
            if (COUNT) cnt.visits++;
            float tf0, tf1;
            const float tn0 = slab_io(r, a.x, a.y, a.z, a.w, b.x, b.y, tmin, t_best, tf0);
            const float tn1 = slab_io(r, b.z, b.w, c.x, c.y, c.z, c.w, tmin, t_best, tf1);
            const bool h0 = tn0 != INFINITY && d.x != INT32_MIN, h1 = tn1 != INFINITY && d.y != INT32_MIN;
            // a treelet child ~t becomes the P-k-d root at position t << lg
            const int32_t r0 = d.x >= 0 ? d.x : (int32_t)~((uint32_t)(~d.x) << lg);
            const int32_t r1 = d.y >= 0 ? d.y : (int32_t)~((uint32_t)(~d.y) << lg);
            if (h0 && h1) {
                const bool first0 = tn0 <= tn1;
                const int32_t rn = first0 ? r0 : r1, rf = first0 ? r1 : r0;
                const float tnf = first0 ? tn1 : tn0;
                if (rf < 0) st.push(__float_as_int(tnf));
                st.push(rf);
                node = rn;
                t0 = first0 ? tn0 : tn1;
                t1 = first0 ? tf0 : tf1;
            } else if (h0) {
                node = r0;
                t0 = tn0;
                t1 = tf0;
            } else if (h1) {
                node = r1;
                t0 = tn1;
                t1 = tf1;
            } else {
                pop = true;
            }
        } else {
            const uint32_t p = ~(uint32_t)node;
            const uint32_t tl = p >> lg;
            if (tl != cur_tl) {
                cur_tl = tl;
                rmax = A.tl_rmax[tl];
                const uint32_t tb = tl << lg;
                m = A.n_sph - tb < msk + 1u ? A.n_sph - tb : msk + 1u;
            }
            const uint32_t i = p & msk, bs = p - i;
            const float4 sp = A.leaf_sph[p];
            const int dd = A.pkd_dim[p];
            if (COUNT) cnt.visits++;
            const float sc = dd == 0 ? sp.x : (dd == 1 ? sp.y : sp.z);
            const float inv = dd == 0 ? r.inv.x : (dd == 1 ? r.inv.y : r.inv.z);
            const float oil = dd == 0 ? r.oil.x : (dd == 1 ? r.oil.y : r.oil.z);
            const float oih = dd == 0 ? r.oih.x : (dd == 1 ? r.oih.y : r.oih.z);
            const float tlo = __builtin_fmaf(sc - rmax, inv, -oil), thi = __builtin_fmaf(sc + rmax, inv, -oih);
            const float tnr = fminf(tlo, thi), tfr = fmaxf(tlo, thi);
            float lim = fminf(t1, t_best);
            // a sphere in the node's interval is tested against its own box first (centre -+
            // radius: the leaf box an LBVH traversal tests before its sphere), a filter that
            // changes no result: most spheres in a slab lie off the ray in the other two axes
            if (fmaxf(t0, tnr) <= fminf(tfr, lim) && (!ANY || (int)p != skip) &&
                slab(r, sp.x - sp.w, sp.y - sp.w, sp.z - sp.w, sp.x + sp.w, sp.y + sp.w, sp.z + sp.w, tmin,
                     t_best) != INFINITY) {
                if (COUNT) cnt.tests++;
                const float t = sphere_t(r.o, r.d, sp, tmin, tmax);
                if (t >= 0.0f) {
                    if (ANY) return 1;
                    if (t < t_best) {
                        t_best = t;
                        best = (int)p;
                    } else if (t == t_best && best >= 0 && A.leaf_id[p] < A.leaf_id[best]) {
                        best = (int)p;
                    }
                    lim = fminf(t1, t_best);
                }
            }
            const float n_t1 = fminf(tfr, lim), f_t0 = fmaxf(t0, tnr);
            const uint32_t sg = inv < 0.0f ? 1u : 0u;
            const uint32_t nc = 2 * i + 1 + sg, fc = 2 * i + 2 - sg;
            const bool need_n = nc < m && t0 <= n_t1;
            const bool need_f = fc < m && f_t0 <= lim;
            if (need_n && need_f) {
                st.push(__float_as_int(f_t0));
                st.push(~(int32_t)(bs + fc));
                node = ~(int32_t)(bs + nc);
                t1 = n_t1;
            } else if (need_n) {
                node = ~(int32_t)(bs + nc);
                t1 = n_t1;
            } else if (need_f) {
                node = ~(int32_t)(bs + fc);
                t0 = f_t0;
                t1 = lim;
            } else {
                pop = true;
            }
        }
        if (pop) {
            for (;;) {
                if (st.sp == 0) return ANY ? 0 : best;
                const int32_t v = st.pop();
                if (v >= 0) {
                    node = v;
                    break;
                }
                const float a0 = __int_as_float(st.pop());
                if (a0 <= t_best) {
                    node = v;
                    t0 = a0;
                    t1 = t_best;
                    break;
                }
            }
        }
    }
}

// The structure a kernel traverses: TL = treelet P-k-d trees, else the LBVH.
template <int S, bool COUNT, bool PAIRS, bool TL>
__device__ __forceinline__ int trace_closest_x(const TraceArgs& A, const RayPre& r, float tmin, float tmax,
                                               float& t_best, LaneStack<S>& st, Count& cnt) {
    if (TL) return trace_tl<S, COUNT, false>(A, r, tmin, tmax, -1, t_best, st, cnt);
    return trace_closest<S, COUNT, PAIRS>(A, r, tmin, tmax, t_best, st, cnt);
}
template <int S, bool COUNT, bool PAIRS, bool TL>
__device__ __forceinline__ bool trace_any_x(const TraceArgs& A, const RayPre& r, float tmin, float tmax, int skip,
                                            LaneStack<S>& st, Count& cnt) {
    if (TL) {
        float tb;
        return trace_tl<S, COUNT, true>(A, r, tmin, tmax, skip, tb, st, cnt) != 0;
    }
    return trace_any<S, COUNT, PAIRS>(A, r, tmin, tmax, skip, st, cnt);
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

// Pixel epilogue of raygen.cu:250-272: average over samples, progressive blend, pseudo depth
// (background on a pixel that no sample hit), then every output buffer.
__device__ __forceinline__ void write_pixel_main(const TraceArgs& A, uint32_t opix, float4 col, float depth,
                                                 float best_t, int best_leaf) {
    const FrameConsts& f = A.fc;
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
}

__device__ __forceinline__ uint32_t half2_bits(float lo, float hi) {
    return (uint32_t)__half_as_ushort(__float2half(lo)) | ((uint32_t)__half_as_ushort(__float2half(hi)) << 16);
}

// Features of a pixel (8 halves): albedo, normal and t of sample 0's hit (zeros on a miss), then
// the mean AO visibility over the samples.
__device__ __forceinline__ void write_pixel(const TraceArgs& A, uint32_t opix, float4 col, float depth, float best_t,
                                            int best_leaf, float aux[8], float vis_sum) {
    write_pixel_main(A, opix, col, depth, best_t, best_leaf);
    if (A.fc.want_aux) {
        aux[7] = vis_sum / (float)A.fc.spp;
        *reinterpret_cast<uint4*>(A.out_aux + (size_t)opix * 8) = make_uint4(
            half2_bits(aux[0], aux[1]), half2_bits(aux[2], aux[3]), half2_bits(aux[4], aux[5]), half2_bits(aux[6], aux[7]));
    }
}

// The same features written in two parts, so that a kernel need not keep them live over the
// traversal of the remaining samples: halves 0-5 once sample 0's primary ray is resolved, halves
// 6-7 with the pixel.
__device__ __forceinline__ void write_aux_head(const TraceArgs& A, uint32_t opix, float a0, float a1, float a2,
                                               float n0, float n1, float n2) {
    if (A.fc.want_aux) {
        uint3* q = reinterpret_cast<uint3*>(A.out_aux + (size_t)opix * 8);
        *q = make_uint3(half2_bits(a0, a1), half2_bits(a2, n0), half2_bits(n1, n2));
    }
}
__device__ __forceinline__ void write_aux_tail(const TraceArgs& A, uint32_t opix, float t0, float vis_sum) {
    if (A.fc.want_aux)
        *reinterpret_cast<uint32_t*>(A.out_aux + (size_t)opix * 8 + 6) = half2_bits(t0, vis_sum / (float)A.fc.spp);
}

template <int S, int MODE, bool COUNT, bool TL = false>
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
        int leaf = trace_closest_x<S, COUNT, true, TL>(A, r, 1e-6f, 1e20f, t, st, cnt);
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
                V3 ad = ao_direction(N, aseed, s, f.spp);
                RayPre ra = prep(P, ad, pad);
                if (COUNT) cnt.rays++;
                bool occ = trace_any_x<S, COUNT, true, TL>(A, ra, 0.0f, f.ao_distance, leaf, st, cnt);
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
                    int lb = trace_closest_x<S, COUNT, true, TL>(A, rb, 1e-3f, 1e8f, tb, st, cnt);
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
    write_pixel(A, opix, col, depth, best_t, best_leaf, aux, vis_sum);
}

// Lane per pixel, samples in a loop (every sample count). Held to 72 VGPRs (7 wavefronts per SIMD
// instead of 4): 64-spp AO 64.3 -> 57.7 ms, 1-spp AO 1.22 -> 1.14 ms.
template <int S, int MODE, bool COUNT, bool STEAL, bool TL = false>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_trace(TraceArgs A) {
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
        const uint32_t tw = 1u << A.tw_log2, th = 64u >> A.tw_log2;
        uint32_t px = A.fc.win_x0 + tx * tw + (lane & (tw - 1)), py = A.fc.win_y0 + ty * th + (lane >> A.tw_log2);
        if (px < A.fc.win_x0 + A.fc.win_w && py < A.fc.win_y0 + A.fc.win_h)
            shade_pixel<S, MODE, COUNT, TL>(A, px, py, pad, st, cnt);
    }
    if (COUNT) {
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 2), (unsigned long long)cnt.visits);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 4), (unsigned long long)cnt.tests);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 6), (unsigned long long)cnt.rays);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 8), (unsigned long long)cnt.pair_visits);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 10), (unsigned long long)cnt.leaf_fetches);
    }
}

// One sample of a Lambert (MODE 0) or AO (MODE 1) pixel, the arithmetic of shade_pixel: the
// sample colour, its depth contribution, the hit (t, leaf), the visibility and, for the
// features, albedo and normal.
template <int S, int MODE, bool COUNT, bool PAIRS, bool TL = false>
__device__ __forceinline__ void shade_sample(const TraceArgs& A, uint32_t px, uint32_t py, uint32_t pix, uint32_t s,
                                             float jx, float jy, float pad, LaneStack<S>& st, Count& cnt,
                                             float4& sc, float& ray_depth, float& t, int& leaf, float& V,
                                             float4& alb_out, V3& n_out) {
    const FrameConsts& f = A.fc;
    V3 o, d;
    primary_ray(f, px, py, jx, jy, o, d);
    RayPre r = prep(o, d, pad);
    if (COUNT) cnt.rays++;
    leaf = trace_closest_x<S, COUNT, PAIRS, TL>(A, r, 1e-6f, 1e20f, t, st, cnt);
    ray_depth = FLT_MAX;
    V = 0.0f;
    if (leaf < 0) {
        if (MODE == 1)
            sc = make_float4(f.bg[0], f.bg[1], f.bg[2], f.bg[3]);
        else
            sc = make_float4(1.0f * 0.8f, 1.0f * 0.8f, 1.0f * 0.8f, 1.0f);
        return;
    }
    ray_depth = t;
    float4 sp = A.leaf_sph[leaf];
    uint32_t id = A.leaf_id[leaf];
    V3 P = V3{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z};
    V3 N = V3{P.x - sp.x, P.y - sp.y, P.z - sp.z};
    if (dot(N, d) > 0.0f) N = neg(N);
    N = normalize(N);
    float4 alb = A.col[id];
    alb_out = alb;
    n_out = N;
    if (MODE == 1) {
        uint32_t aseed = tea16(pix, 0xA0000000u + f.frame_idx * 4096u + s);
        V3 ad = ao_direction(N, aseed, s, f.spp);
        RayPre ra = prep(P, ad, pad);
        if (COUNT) cnt.rays++;
        bool occ = trace_any_x<S, COUNT, PAIRS, TL>(A, ra, 0.0f, f.ao_distance, leaf, st, cnt);
        V = occ ? 0.0f : 1.0f;
        sc = make_float4(alb.x * V, alb.y * V, alb.z * V, 1.0f);
    } else {
        float sh = 0.2f + 0.6f * fabsf(dot(N, d));
        sc = make_float4(alb.x * sh, alb.y * sh, alb.z * sh, 1.0f);
    }
}

// Sample-parallel persistent traversal (MODE 0 Lambert, 1 AO; spp a power of two or a multiple
// of 64). A wavefront takes an 8x8 (tile_w x 64/tile_w) pixel tile and walks it in passes; in a
// pass, the lanes hold L = min(spp, 64) consecutive samples of each of 64/L pixels, so the
// primary rays of a lane group start in one pixel and a background pixel costs one quick pass.
// Sample s of a pixel draws its jitter from the pixel's stream advanced by 2s steps
// (lcg_skip), and the group folds its samples into the pixel in sample order (colour sums,
// first strictly nearer hit, visibility, sample-0 features), so every output is bit-identical
// to the lane-per-pixel kernels. Held to 72 VGPRs (7 wavefronts per SIMD instead of 4, a few
// registers spilled): 64-spp AO 33.8 -> 30.6 ms, Lambert 22.9 -> 20.8 ms (5 wavefronts: 32.6 and
// 21.2 ms, 6: 31.1 and 21.3, 8: 30.8 and 21.0).
template <int S, int MODE, bool COUNT, bool PAIRS, bool TL = false>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_trace_sl(TraceArgs A) {
    __shared__ int32_t lds_stack[S * TB];
    __shared__ float4 lds_col[TB];  // the wavefront's sample colours of one round
    const uint32_t gtid = blockIdx.x * TB + threadIdx.x;
    LaneStack<S> st{lds_stack + threadIdx.x, A.spill + gtid, A.spill_stride, 0, A.ctr + 1};
    Count cnt;
    const FrameConsts& f = A.fc;
    const float amax = A.info->absmax;
    const float pad = (fmaxf(f.pad, amax) + amax) * (1.0f / 131072.0f);
    const uint32_t lane = __lane_id();
    const uint32_t L = f.spp >= 64u ? 64u : f.spp;  // lanes per pixel
    const uint32_t G = 64u / L;                      // pixels per pass
    const uint32_t rounds = f.spp / L;
    const uint32_t g = lane / L, sl = lane % L;
    const uint32_t base = g * L;  // first lane of this lane's pixel group
    const uint32_t tw = 1u << A.tw_log2, th = 64u >> A.tw_log2;
    for (;;) {
        uint32_t t0 = 0;
        if (lane == 0) t0 = atomicAdd(A.ctr, 1u);
        const uint32_t tile = __shfl(t0, 0, 64);
        if (tile >= A.n_tiles) break;
        const uint32_t tx = tile % A.tiles_x, ty = tile / A.tiles_x;
        for (uint32_t q = 0; q < 64u / G; ++q) {
            const uint32_t tp = q * G + g;  // pixel of this group inside the tile
            const uint32_t px = f.win_x0 + tx * tw + (tp & (tw - 1)), py = f.win_y0 + ty * th + (tp >> A.tw_log2);
            const bool valid = px < f.win_x0 + f.win_w && py < f.win_y0 + f.win_h;
            const uint32_t pix = py * f.width + px;
            const uint32_t seed0 = tea16(pix, f.frame_idx);
            float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
            // best_t is the least hit distance over the samples, the pixel's depth (FLT_MAX without a hit)
            float best_t = INFINITY, vis_sum = 0.0f;
            int best_leaf = -1;
            float aux[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            for (uint32_t rd = 0; rd < rounds; ++rd) {
                const uint32_t s = rd * L + sl;
                float4 sc = make_float4(0.f, 0.f, 0.f, 0.f);
                float ray_depth = FLT_MAX, t = INFINITY, V = 0.0f;
                int leaf = -1;
                float4 alb = make_float4(0.f, 0.f, 0.f, 0.f);
                V3 N = V3{0.f, 0.f, 0.f};
                if (valid) {
                    uint32_t seed = lcg_skip(seed0, 2u * s);
                    float jx = rnd(seed);
                    float jy = rnd(seed);
                    shade_sample<S, MODE, COUNT, PAIRS, TL>(A, px, py, pix, s, jx, jy, pad, st, cnt, sc, ray_depth,
                                                            t, leaf, V, alb, N);
                }
                // visibility is 0 or 1 per sample, so its sum is exact in any order: the count of
                // the group's unoccluded hits
                const uint64_t gmask = L == 64u ? ~0ull : ((1ull << L) - 1ull) << base;
                vis_sum += (float)__popcll(__ballot(leaf >= 0 && V > 0.0f) & gmask);
                // the group's closest hit, the first sample of least t (an exchange over lane ^ m
                // stays inside the aligned group of L lanes); an earlier round wins ties
                float rt = leaf >= 0 ? t : INFINITY;
                int rl = leaf;
                uint32_t rs = s;
                for (uint32_t m = L >> 1; m > 0; m >>= 1) {
                    const float ot = __shfl_xor(rt, (int)m, 64);
                    const int ol = __shfl_xor(rl, (int)m, 64);
                    const uint32_t os = __shfl_xor(rs, (int)m, 64);
                    if (ot < rt || (ot == rt && os < rs)) {
                        rt = ot;
                        rl = ol;
                        rs = os;
                    }
                }
                if (rl >= 0 && rt < best_t) {
                    best_t = rt;
                    best_leaf = rl;
                }
                // colours are summed in sample order by the group's first lane (the one that
                // writes the pixel), from the wavefront's slice of LDS
                float4* wcol = lds_col + (threadIdx.x - lane);
                wcol[lane] = sc;
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if (sl == 0) {
                    for (uint32_t i = 0; i < L; ++i) {
                        const float4 c = wcol[base + i];
                        col.x += c.x;
                        col.y += c.y;
                        col.z += c.z;
                        col.w += c.w;
                    }
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                if (rd == 0) {  // features of sample 0
                    const int src = (int)base;
                    const int l0 = __shfl(leaf, src, 64);
                    const float a0 = __shfl(alb.x, src, 64), a1 = __shfl(alb.y, src, 64), a2 = __shfl(alb.z, src, 64);
                    const float n0 = __shfl(N.x, src, 64), n1 = __shfl(N.y, src, 64), n2 = __shfl(N.z, src, 64);
                    const float t0s = __shfl(t, src, 64);
                    if (l0 >= 0) {
                        aux[0] = a0; aux[1] = a1; aux[2] = a2;
                        aux[3] = n0; aux[4] = n1; aux[5] = n2;
                        aux[6] = t0s;
                    }
                }
            }
            if (valid && sl == 0) {
                const uint32_t opix = (py - f.win_y0) * f.win_w + (px - f.win_x0);
                write_pixel(A, opix, col, best_t < INFINITY ? best_t : FLT_MAX, best_t, best_leaf, aux, vis_sum);
            }
        }
    }
    if (COUNT) {
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 2), (unsigned long long)cnt.visits);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 4), (unsigned long long)cnt.tests);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 6), (unsigned long long)cnt.rays);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 8), (unsigned long long)cnt.pair_visits);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 10), (unsigned long long)cnt.leaf_fetches);
    }
}

// Persistent traversal with wavefront-ballot compaction and refill (MODE 0 Lambert, 1 AO).
// Each lane owns one pixel at a time and walks its samples in order (primary ray, then for AO
// one occlusion ray). The inner loop advances every active lane by one node visit and keeps
// going while at least `refill` lanes of the wave still trace; below that the wave leaves the
// loop, finished lanes shade and set up their next ray (next AO ray, next sample, or a new pixel
// taken with one wave-aggregated atomic), and traversal resumes. Per-pixel arithmetic and
// random streams are those of k_trace, so the results are bit-identical.
enum : int { L_IDLE = 0, L_NEED = 1, L_PRIM = 2, L_AO = 3, L_PRIM_DONE = 4, L_AO_DONE = 5, L_OFF = 6 };

// Held to 5 wavefronts per SIMD (at most 102 VGPRs; it takes 95 and spills nothing). With one step
// for both kinds of ray the kernel is shorter, and the 1-spp frame's trace went from 1.02 ms at 7
// wavefronts (72 VGPRs and 96 bytes of spills per lane) to 0.92 ms (6 wavefronts, 60 bytes of
// spills: 0.91 ms). Before that step, 7 wavefronts had been the fastest (5: 1.07, 7: 1.04 ms).
template <int S, int MODE, bool COUNT>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_trace_refill(TraceArgs A,
                                                                                                uint32_t refill) {
    __shared__ int32_t lds_stack[S * TB];
    const uint32_t gtid = blockIdx.x * TB + threadIdx.x;
    LaneStack<S> st{lds_stack + threadIdx.x, A.spill + gtid, A.spill_stride, 0, A.ctr + 1};
    Count cnt;
    const FrameConsts& f = A.fc;
    const float amax = A.info->absmax;
    const float pad = (fmaxf(f.pad, amax) + amax) * (1.0f / 131072.0f);
    const uint32_t lane = __lane_id();
    const uint32_t tw = 1u << A.tw_log2, th = 64u >> A.tw_log2;
    const uint32_t tiles_x = (f.win_w + tw - 1) / tw;
    const uint32_t total_work = ((f.win_h + th - 1) / th) * tiles_x * 64u;  // work indices: whole tiles
    int state = L_IDLE;
    bool exhausted = false;
    // pixel state
    uint32_t px = 0, py = 0, pix = 0, opix = 0, s = 0, seed = 0;
    float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
    // best_t is the least hit distance over the samples, the pixel's depth (FLT_MAX without a hit)
    float best_t = INFINITY, vis_sum = 0.0f;
    int best_leaf_px = -1;
    float aux6 = 0.f;  // t of sample 0's hit (the other features are written when it is found)
    // ray state
    RayPre r;
    float tmin = 0.f, tmax = 0.f, t_best = 0.f;
    int best = -1, skip = -1;
    int32_t node = 0;
    float4 alb = make_float4(0.f, 0.f, 0.f, 0.f);
    float prim_t = 0.f;
    int prim_leaf = -1;
    for (;;) {
        // 1. idle lanes take pixels (work index -> 8x8 tile order inside the window)
        uint64_t idle = __ballot(state == L_IDLE);
        if (idle && !exhausted) {
            int leader = __ffsll((unsigned long long)idle) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(A.ctr, (uint32_t)__popcll(idle));
            base = __shfl(base, leader, 64);
            if (base + (uint32_t)__popcll(idle) >= total_work) exhausted = true;
            if (state == L_IDLE) {
                uint32_t w = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                uint32_t tile = w >> 6, k = w & 63u;
                uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
                uint32_t lx = tx * tw + (k & (tw - 1)), ly = ty * th + (k >> A.tw_log2);
                // tiles past the window edge map to no pixel; the lane simply asks again
                if (w < total_work && lx < f.win_w && ly < f.win_h) {
                    px = f.win_x0 + lx;
                    py = f.win_y0 + ly;
                    pix = py * f.width + px;
                    opix = ly * f.win_w + lx;
                    s = 0;
                    seed = tea16(pix, f.frame_idx);
                    col = make_float4(0.f, 0.f, 0.f, 0.f);
                    best_t = INFINITY;
                    best_leaf_px = -1;
                    vis_sum = 0.0f;
                    aux6 = 0.f;
                    state = L_NEED;
                } else if (w >= total_work) {
                    state = L_OFF;
                }
            }
        }
        if (exhausted && state == L_IDLE) state = L_OFF;
        // 2. new primary rays
        if (state == L_NEED) {
            float jx = rnd(seed);
            float jy = rnd(seed);
            V3 o, d;
            primary_ray(f, px, py, jx, jy, o, d);
            r = prep(o, d, pad);
            tmin = 1e-6f;
            tmax = 1e20f;
            t_best = tmax;
            best = -1;
            node = 0;
            st.sp = 0;
            if (COUNT) cnt.rays++;
            state = L_PRIM;
        }
        uint64_t act = __ballot(state == L_PRIM || state == L_AO);
        uint64_t live = __ballot(state != L_OFF);
        if (!live) break;
        // 3. traversal: one node visit per active lane per iteration
        while (act) {
            if (state == L_PRIM || state == L_AO) {  // one step for both kinds (primary and AO lanes mix here)
                const bool ao = state == L_AO;
                const int res = step_mixed<S, COUNT>(A, r, tmin, tmax, t_best, best, node, st, cnt, ao ? skip : -1, ao);
                if (res) {
                    if (ao) {
                        state = L_AO_DONE;
                        best = res;  // 2 = occluded
                    } else {
                        state = L_PRIM_DONE;
                    }
                }
            }
            act = __ballot(state == L_PRIM || state == L_AO);
            if ((uint32_t)__popcll(act) < refill) break;
        }
        // 4. shading of finished rays
        bool sample_done = false;
        if (state == L_PRIM_DONE) {
            if (best < 0) {
                float4 sc = MODE == 1 ? make_float4(f.bg[0], f.bg[1], f.bg[2], f.bg[3])
                                      : make_float4(1.0f * 0.8f, 1.0f * 0.8f, 1.0f * 0.8f, 1.0f);
                col.x += sc.x; col.y += sc.y; col.z += sc.z; col.w += sc.w;
                if (s == 0) write_aux_head(A, opix, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f);
                sample_done = true;
            } else {
                float t = t_best;
                float4 sp = A.leaf_sph[best];
                V3 o = r.o, d = r.d;
                V3 Pp = V3{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z};
                V3 N = V3{Pp.x - sp.x, Pp.y - sp.y, Pp.z - sp.z};
                if (dot(N, d) > 0.0f) N = neg(N);
                N = normalize(N);
                alb = A.col[A.leaf_id[best]];
                if (t < best_t) {
                    best_t = t;
                    best_leaf_px = best;
                }
                if (s == 0) {
                    write_aux_head(A, opix, alb.x, alb.y, alb.z, N.x, N.y, N.z);
                    aux6 = t;
                }
                if (MODE == 1) {
                    uint32_t aseed = tea16(pix, 0xA0000000u + f.frame_idx * 4096u + s);
                    V3 ad = ao_direction(N, aseed, s, f.spp);
                    prim_t = t;
                    prim_leaf = best;
                    r = prep(Pp, ad, pad);
                    tmin = 0.0f;
                    tmax = f.ao_distance;
                    t_best = tmax;  // an any-hit ray's bound in step_mixed
                    skip = best;
                    node = 0;
                    st.sp = 0;
                    if (COUNT) cnt.rays++;
                    state = L_AO;
                } else {
                    float sh = 0.2f + 0.6f * fabsf(dot(N, d));
                    col.x += alb.x * sh; col.y += alb.y * sh; col.z += alb.z * sh; col.w += 1.0f;
                    sample_done = true;
                }
            }
        } else if (state == L_AO_DONE) {
            float V = best == 2 ? 0.0f : 1.0f;
            col.x += alb.x * V; col.y += alb.y * V; col.z += alb.z * V; col.w += 1.0f;
            vis_sum += V;
            sample_done = true;
            (void)prim_t;
            (void)prim_leaf;
        }
        if (sample_done) {
            if (++s < f.spp) {
                state = L_NEED;
            } else {
                write_pixel_main(A, opix, col, best_t < INFINITY ? best_t : FLT_MAX, best_t, best_leaf_px);
                write_aux_tail(A, opix, aux6, vis_sum);
                state = L_IDLE;
            }
        }
    }
    if (COUNT) {
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 2), (unsigned long long)cnt.visits);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 4), (unsigned long long)cnt.tests);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 6), (unsigned long long)cnt.rays);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 8), (unsigned long long)cnt.pair_visits);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.ctr + 10), (unsigned long long)cnt.leaf_fetches);
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

template <int S, bool PAIRS>
void launch_sl_p(const TraceArgs& A, unsigned grid, bool count, int mode, hipStream_t s) {
    if (mode == 1) {
        if (count)
            hipLaunchKernelGGL((k_trace_sl<S, 1, true, PAIRS>), dim3(grid), dim3(TB), 0, s, A);
        else
            hipLaunchKernelGGL((k_trace_sl<S, 1, false, PAIRS>), dim3(grid), dim3(TB), 0, s, A);
    } else {
        if (count)
            hipLaunchKernelGGL((k_trace_sl<S, 0, true, PAIRS>), dim3(grid), dim3(TB), 0, s, A);
        else
            hipLaunchKernelGGL((k_trace_sl<S, 0, false, PAIRS>), dim3(grid), dim3(TB), 0, s, A);
    }
}
// The sample-lanes kernel is the one whose rays are coherent enough for the cost of the
// leaf-pair branch to show: a tree built without pairs gets the variant without it.
template <int S>
void launch_sl(const TraceArgs& A, unsigned grid, bool count, int mode, bool pairs, hipStream_t s) {
    if (pairs)
        launch_sl_p<S, true>(A, grid, count, mode, s);
    else
        launch_sl_p<S, false>(A, grid, count, mode, s);
}

// Kernels over the treelet P-k-d trees.
template <int MODE, bool COUNT>
void launch_tl_m(const TraceArgs& A, unsigned grid, bool sl, hipStream_t s) {
    if (sl && MODE != 2)
        hipLaunchKernelGGL((k_trace_sl<16, MODE == 2 ? 0 : MODE, COUNT, false, true>), dim3(grid), dim3(TB), 0, s, A);
    else
        hipLaunchKernelGGL((k_trace<16, MODE, COUNT, true, true>), dim3(grid), dim3(TB), 0, s, A);
}
void launch_tl(const TraceArgs& A, unsigned grid, bool sl, bool count, int mode, hipStream_t s) {
    if (mode == 1) {
        if (count) launch_tl_m<1, true>(A, grid, sl, s); else launch_tl_m<1, false>(A, grid, sl, s);
    } else if (mode == 2) {
        if (count) launch_tl_m<2, true>(A, grid, false, s); else launch_tl_m<2, false>(A, grid, false, s);
    } else {
        if (count) launch_tl_m<0, true>(A, grid, sl, s); else launch_tl_m<0, false>(A, grid, sl, s);
    }
}

template <int S>
void launch_refill(const TraceArgs& A, unsigned grid, uint32_t refill, bool count, int mode, hipStream_t s) {
    if (mode == 1) {
        if (count)
            hipLaunchKernelGGL((k_trace_refill<S, 1, true>), dim3(grid), dim3(TB), 0, s, A, refill);
        else
            hipLaunchKernelGGL((k_trace_refill<S, 1, false>), dim3(grid), dim3(TB), 0, s, A, refill);
    } else {
        if (count)
            hipLaunchKernelGGL((k_trace_refill<S, 0, true>), dim3(grid), dim3(TB), 0, s, A, refill);
        else
            hipLaunchKernelGGL((k_trace_refill<S, 0, false>), dim3(grid), dim3(TB), 0, s, A, refill);
    }
}

// Batched trace: the jobs (timestep, camera) of a series, each with its own timestep's structure
// in its constants, share one persistent launch. The global tile index runs over the concatenated
// tiles of the jobs (prefix[j] = first tile of job j); each wavefront steals tiles and shades them
// with its job's constants and output buffers.
// Held to 72 VGPRs (7 wavefronts per SIMD instead of 4) like the single-frame kernels: the batched
// series (python bench.py --config series) went from 149.0 to 135.7 ms per launch.
template <int S, int MODE, bool TL = false>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_trace_batch(const TraceArgs* __restrict__ jobs,
                                                    const uint32_t* __restrict__ prefix, uint32_t n_jobs,
                                                    uint32_t* ctr) {
    __shared__ int32_t lds_stack[S * TB];
    const uint32_t gtid = blockIdx.x * TB + threadIdx.x;
    LaneStack<S> st{lds_stack + threadIdx.x, jobs[0].spill + gtid, jobs[0].spill_stride, 0, ctr + 1};
    Count cnt;
    const uint32_t lane = __lane_id();
    const uint32_t total = prefix[n_jobs];
    for (;;) {
        uint32_t t0 = 0;
        if (lane == 0) t0 = atomicAdd(ctr, 1u);
        const uint32_t tile = __shfl(t0, 0, 64);
        if (tile >= total) break;
        // the job holding the tile: the last j with prefix[j] <= tile (binary search: a launch
        // over a whole series holds hundreds of jobs)
        uint32_t j = 0, hi = n_jobs - 1;
        while (j < hi) {
            const uint32_t mid = (j + hi + 1) >> 1;
            if (prefix[mid] <= tile)
                j = mid;
            else
                hi = mid - 1;
        }
        const TraceArgs& A = jobs[j];
        const float amax = A.info->absmax;
        const float pad = (fmaxf(A.fc.pad, amax) + amax) * (1.0f / 131072.0f);
        const uint32_t lt = tile - prefix[j];
        const uint32_t tw = 1u << A.tw_log2, th = 64u >> A.tw_log2;
        const uint32_t tx = lt % A.tiles_x, ty = lt / A.tiles_x;
        const uint32_t px = A.fc.win_x0 + tx * tw + (lane & (tw - 1)), py = A.fc.win_y0 + ty * th + (lane >> A.tw_log2);
        if (px < A.fc.win_x0 + A.fc.win_w && py < A.fc.win_y0 + A.fc.win_h)
            shade_pixel<S, MODE, false, TL>(A, px, py, pad, st, cnt);
    }
}

}  // namespace

// One batched launch over jobs that share the mode.
void launch_trace_batch(mmhrt_ctx_s* c, const void* jobs, const uint32_t* prefix, uint32_t n_jobs, uint32_t* ctr,
                        int mode, bool treelets) {
    unsigned grid = (unsigned)std::max<int64_t>(1, c->cu_count * c->opt_blocks_per_cu);
    const TraceArgs* J = static_cast<const TraceArgs*>(jobs);
    int S = (int)c->opt_stack_lds;
    launch_fill_u32(c->stream, ctr, 0u, 16);
    if (treelets) {  // treelet P-k-d trees: 16-entry LDS stacks
        if (mode == 1)
            hipLaunchKernelGGL((k_trace_batch<16, 1, true>), dim3(grid), dim3(TB), 0, c->stream, J, prefix, n_jobs, ctr);
        else
            hipLaunchKernelGGL((k_trace_batch<16, 0, true>), dim3(grid), dim3(TB), 0, c->stream, J, prefix, n_jobs, ctr);
        return;
    }
#define MMHRT_BATCH(SV)                                                                                     \
    if (mode == 1)                                                                                          \
        hipLaunchKernelGGL((k_trace_batch<SV, 1>), dim3(grid), dim3(TB), 0, c->stream, J, prefix, n_jobs, ctr); \
    else                                                                                                    \
        hipLaunchKernelGGL((k_trace_batch<SV, 0>), dim3(grid), dim3(TB), 0, c->stream, J, prefix, n_jobs, ctr);
    switch (S) {
    case 8: MMHRT_BATCH(8) break;
    case 24: MMHRT_BATCH(24) break;
    case 32: MMHRT_BATCH(32) break;
    default: MMHRT_BATCH(16) break;
    }
#undef MMHRT_BATCH
}

// Job constants for the batched kernel, written into a host buffer of batch_job_size() bytes.
size_t batch_job_size() { return sizeof(TraceArgs); }
int batch_fill_job(mmhrt_ctx_s* c, const FrameConsts& fc, size_t out_offset, void* dst) {
    uint32_t W = fc.win_w, H = fc.win_h;
    (void)H;
    TraceArgs A;
    A.nodes = c->nodes.p;
    A.leaf_sph = c->leaf_sph.p;
    A.leaf_id = c->leaf_id.p;
    A.col = c->ts_col.p;  // the timesteps' own colours, not the context's
    A.info = c->info.p;
    A.spill = c->spill.p;
    A.spill_stride = (uint32_t)(std::max<int64_t>(1, c->cu_count * c->opt_blocks_per_cu) * TB);
    A.ctr = nullptr;
    A.tw_log2 = tile_log2(c->opt_tile_w);
    A.tiles_x = (W + (1u << A.tw_log2) - 1) >> A.tw_log2;
    A.n_tiles = A.tiles_x * ((fc.win_h + (64u >> A.tw_log2) - 1) / (64u >> A.tw_log2));
    A.out_col = c->bt_col.p + out_offset;
    A.out_depth = c->bt_depth.p + out_offset;
    A.out_t = c->bt_t.p + out_offset;
    A.out_id = c->bt_id.p + out_offset;
    A.out_rgba8 = c->bt_rgba8.p + out_offset;
    A.out_aux = nullptr;
    A.pkd_dim = c->accel == 1 ? c->pkd_dim.p : nullptr;
    A.tl_rmax = c->accel == 1 ? c->tl_rmax.p : nullptr;
    A.n_sph = (uint32_t)c->n;
    A.tl_log2 = 0;
    while ((1u << A.tl_log2) < c->pkd_M) ++A.tl_log2;
    A.fc = fc;
    A.fc.want_aux = 0;
    A.fc.accumulate = 0;
    std::memcpy(dst, &A, sizeof A);
    return (int)A.n_tiles;
}

// tile width option -> log2 (1, 2, 4, 8, 16, 32 or 64 pixels wide; 64 pixels per tile)
uint32_t tile_log2(int64_t tw) {
    uint32_t l = 3;
    for (uint32_t k = 0; k <= 6; ++k)
        if ((int64_t)(1u << k) == tw) l = k;
    return l;
}

namespace {
// Kernel choice for one frame (the same for both entry points): sample-parallel lanes for Lambert
// and AO with spp a power of two up to 64 or a multiple of 64, the ballot-refill kernel when a
// refill threshold is set, else the lane-per-pixel kernel.
void dispatch_trace(mmhrt_ctx_s* c, const TraceArgs& A, unsigned grid, bool count) {
    const FrameConsts& fc = A.fc;
    int S = (int)c->opt_stack_lds;
    if (S != 8 && S != 16 && S != 24 && S != 32) S = 16;
    bool steal = c->opt_persistent != 0;
    const int64_t rq = c->opt_refill < 0 ? (fc.spp == 1 ? 16 : 0) : c->opt_refill;
    uint32_t refill = (uint32_t)std::max<int64_t>(0, std::min<int64_t>(64, rq));
    const bool sl_shape = fc.spp > 1 && ((fc.spp >= 64 && fc.spp % 64 == 0) || (fc.spp < 64 && 64 % fc.spp == 0));
    if (c->accel == 1) {
        // treelet P-k-d trees: the sample-lanes kernel where the samples fill the lanes, else
        // lane per pixel with work stealing (16-entry LDS stacks)
        const bool sl = c->opt_sample_lanes != 0 && fc.mode != 2 && sl_shape;
        launch_tl(A, grid, sl, count, fc.mode, c->stream);
        return;
    }
    if (c->opt_sample_lanes != 0 && refill == 0 && fc.mode != 2 && sl_shape) {
        switch (S) {
        case 8: launch_sl<8>(A, grid, count, fc.mode, c->bvh_pairs, c->stream); break;
        case 24: launch_sl<24>(A, grid, count, fc.mode, c->bvh_pairs, c->stream); break;
        case 32: launch_sl<32>(A, grid, count, fc.mode, c->bvh_pairs, c->stream); break;
        default: launch_sl<16>(A, grid, count, fc.mode, c->bvh_pairs, c->stream); break;
        }
    } else if (refill > 0 && fc.mode != 2) {
        switch (S) {
        case 8: launch_refill<8>(A, grid, refill, count, fc.mode, c->stream); break;
        case 24: launch_refill<24>(A, grid, refill, count, fc.mode, c->stream); break;
        case 32: launch_refill<32>(A, grid, refill, count, fc.mode, c->stream); break;
        default: launch_refill<16>(A, grid, refill, count, fc.mode, c->stream); break;
        }
    } else {
        switch (S) {
        case 8: launch1<8>(A, grid, steal, count, fc.mode, c->stream); break;
        case 24: launch1<24>(A, grid, steal, count, fc.mode, c->stream); break;
        case 32: launch1<32>(A, grid, steal, count, fc.mode, c->stream); break;
        default: launch1<16>(A, grid, steal, count, fc.mode, c->stream); break;
        }
    }
}

void fill_scene_args(mmhrt_ctx_s* c, const FrameConsts& fc, TraceArgs& A) {
    A.nodes = c->nodes.p;
    A.leaf_sph = c->leaf_sph.p;
    A.leaf_id = c->leaf_id.p;
    A.col = c->col.p;
    A.info = c->info.p;
    A.pkd_dim = c->accel == 1 ? c->pkd_dim.p : nullptr;
    A.tl_rmax = c->accel == 1 ? c->tl_rmax.p : nullptr;
    A.n_sph = (uint32_t)c->n;
    uint32_t lg = 0;
    while ((1u << lg) < c->pkd_M) ++lg;
    A.tl_log2 = lg;
    A.tw_log2 = tile_log2(c->opt_tile_w);
    A.tiles_x = (fc.win_w + (1u << A.tw_log2) - 1) >> A.tw_log2;
    A.n_tiles = A.tiles_x * ((fc.win_h + (64u >> A.tw_log2) - 1) / (64u >> A.tw_log2));
    A.fc = fc;
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
    if ((e = ensure(c, c->counters, TRACE_CTR_WORDS, "counters"))) return e;
    unsigned grid = (unsigned)(c->cu_count * c->opt_blocks_per_cu);
    if (grid == 0) grid = 1;
    uint32_t lanes = grid * TB;
    if ((e = ensure(c, c->spill, (size_t)lanes * TRACE_SPILL_PER_LANE, "traversal spill"))) return e;
    c->fb_w = W;
    c->fb_h = H;
    c->fb_x0 = fc.win_x0;
    c->fb_y0 = fc.win_y0;
    c->fb_full_w = fc.width;
    c->fb_full_h = fc.height;
    FrameOutputs out{c->fb_col.p, c->fb_depth.p, c->fb_t.p, c->fb_id.p, c->fb_rgba8.p,
                     fc.want_aux ? c->fb_aux.p : nullptr};
    return render_frame_into(c, fc, out, c->counters.p, c->spill.p, lanes, c->opt_count != 0);
}

// The frame into the given device outputs: the persistent grid is capped at `lanes`, whose
// traversal stacks spill into `spill` (lanes x TRACE_SPILL_PER_LANE words, lane-interleaved);
// `ctr` (TRACE_CTR_WORDS words) takes the tile counter, the overflow flag and the instrumentation
// counters and is cleared here by a fill kernel (safe under graph capture). Allocates nothing.
int render_frame_into(mmhrt_ctx_s* c, const FrameConsts& fc, const FrameOutputs& out, uint32_t* ctr, int32_t* spill,
                      uint32_t lanes, bool count) {
    unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(c->cu_count * c->opt_blocks_per_cu, lanes / TB));
    TraceArgs A;
    fill_scene_args(c, fc, A);
    A.ctr = ctr;
    A.spill = spill;
    A.spill_stride = grid * TB;
    A.out_col = out.color;
    A.out_depth = out.depth;
    A.out_t = out.t;
    A.out_id = out.hit_id;
    A.out_rgba8 = out.rgba8;
    A.out_aux = fc.want_aux ? out.aux : nullptr;
    launch_fill_u32(c->stream, ctr, 0u, TRACE_CTR_WORDS);
    rec(c, 6);
    dispatch_trace(c, A, grid, count);
    rec(c, 7);
    hipError_t he = hipGetLastError();
    if (he != hipSuccess) return hip_fail(c, he, "trace launch");
    return 0;
}

}  // namespace mmhrt
