/*
 * oracle.c — CPU oracle for MegaMol's particle-sphere ray-tracing path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker that tests, the smoke check and the benchmark's
 * CPU-baseline leg use. It is never linked into, loaded by or called from the product.
 *
 * What each part restates (paths relative to the MegaMol source tree):
 *   RNG            plugins/optix_hpg/src/optix/random.h:31-62 (TEA<16>, LCG, rnd)
 *   frame state    plugins/optix_hpg/src/optix/Renderer.cpp:136-172, core/include/mmcore/view/Camera.h:350-397
 *   primary ray    plugins/optix_hpg/cuda_resources/raygen.cu:182,241-245; camera.h:11-16
 *   intersection   plugins/optix_hpg/cuda_resources/sphere.cuh:14-41
 *   shading        plugins/optix_hpg/cuda_resources/raygen.cu:122-142 (maxBounces == 0)
 *   pixel resolve  plugins/optix_hpg/cuda_resources/raygen.cu:237-268
 *   ingest         plugins/optix_hpg/src/optix/SphereGeometry.cpp:94-146 with the Accessor
 *                  semantics of plugins/geometry_calls/include/geometry_calls/SimpleSphericalParticles.h:85-175
 *   TF colouring   plugins/datatools/src/AddParticleColors.cpp:19-117
 * Build-owned parts with no reference counterpart (AO estimator, LBVH) follow the
 * published methods named at each function.
 *
 * Compile with -ffp-contract=off: the FP32 sequences below must not be fused.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

double orc_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Run fn(arg, t, T) on T threads (T <= 1: inline). */
typedef struct par_slot {
    void (*fn)(void*, int, int);
    void* arg;
    int t, T;
} par_slot;
static void* par_entry(void* a) {
    par_slot* s = a;
    s->fn(s->arg, s->t, s->T);
    return NULL;
}
static void par_run(int T, void (*fn)(void*, int, int), void* arg) {
    if (T <= 1) {
        fn(arg, 0, 1);
        return;
    }
    pthread_t th[256];
    par_slot sl[256];
    for (int t = 0; t < T; ++t) {
        sl[t] = (par_slot){fn, arg, t, T};
        pthread_create(&th[t], NULL, par_entry, &sl[t]);
    }
    for (int t = 0; t < T; ++t) pthread_join(th[t], NULL);
}
static int threads_for(int threads, uint64_t work) {
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads > 256) threads = 256;
    if (work < 65536) threads = 1;
    return threads < 1 ? 1 : threads;
}
#define PAR_RANGE(n, t, T, a, b) uint64_t a = (uint64_t)(((unsigned __int128)(n) * (t)) / (T)), \
                                          b = (uint64_t)(((unsigned __int128)(n) * ((t) + 1)) / (T))

/* ------------------------------------------------------------------ RNG */
/* TEA encryption (Wheeler & Needham 1994) with an arbitrary key and cycle count. */
void orc_tea_encrypt(uint32_t v[2], const uint32_t k[4], int cycles) {
    uint32_t v0 = v[0], v1 = v[1], s0 = 0;
    for (int n = 0; n < cycles; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + k[0]) ^ (v1 + s0) ^ ((v1 >> 5) + k[1]);
        v1 += ((v0 << 4) + k[2]) ^ (v0 + s0) ^ ((v0 >> 5) + k[3]);
    }
    v[0] = v0;
    v[1] = v1;
}
/* random.h:31-44: tea<16> is TEA with the key {a341316c, c8013ea4, ad90777d, 7e95761e} and 16
 * cycles, returning the first word (Zafar, Olano, Curtis 2010 "GPU random numbers via the TEA"). */
uint32_t orc_tea16(uint32_t val0, uint32_t val1) {
    static const uint32_t key[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};
    uint32_t v[2] = {val0, val1};
    orc_tea_encrypt(v, key, 16);
    return v[0];
}
/* LCG with the Numerical Recipes constants, 24-bit output, random.h:47-52. */
uint32_t orc_lcg(uint32_t* prev) {
    *prev = 1664525u * *prev + 1013904223u;
    return *prev & 0x00FFFFFFu;
}
/* random.h:60-62: float in [0,1). */
float orc_rnd(uint32_t* prev) { return (float)orc_lcg(prev) / (float)0x01000000; }

/* ------------------------------------------------------------ camera */
static void cross3(const float a[3], const float b[3], float o[3]) {
    /* glm::cross component order */
    o[0] = a[1] * b[2] - b[1] * a[2];
    o[1] = a[2] * b[0] - b[2] * a[0];
    o[2] = a[0] * b[1] - b[0] * a[1];
}
static float dot3(const float a[3], const float b[3]) {
    float x = a[0] * b[0], y = a[1] * b[1], z = a[2] * b[2];
    return (x + y) + z;
}
static void normalize3(float v[3]) {
    /* glm::normalize: v * inversesqrt(dot(v,v)), inversesqrt = 1 / sqrt */
    float inv = 1.0f / sqrtf(dot3(v, v));
    v[0] = v[0] * inv;
    v[1] = v[1] * inv;
    v[2] = v[2] * inv;
}

void orc_frame_state_from_camera(const orc_camera* cam, orc_frame_state* fs) {
    /* Renderer.cpp:163-171 */
    for (int i = 0; i < 3; ++i) {
        fs->center[i] = cam->position[i];
        fs->front[i] = cam->direction[i];
        fs->up[i] = cam->up[i];
    }
    cross3(cam->direction, cam->up, fs->right);
    /* Renderer.cpp:148-157: near clip hard-coded to 100 */
    float hfov = 0.5f * cam->fovy;
    fs->th = tanf(hfov) * 100.0f;
    fs->rw = fs->th * cam->aspect;
    fs->nearp = 100.0f;
    /* Renderer.cpp:138-146 + glm::frustum entries used by Camera.h:392-397 */
    float n = cam->near_plane, f = cam->far_plane;
    fs->depth_a = -(f + n) / (f - n);
    fs->depth_b = -1.0f;
    fs->depth_c = -(2.0f * f * n) / (f - n);
}

/* raygen.cu:241-245 and camera.h:11-16; rx, ry are the two rnd() draws of the sample. */
void orc_primary_ray(const orc_frame_state* fs, uint32_t w, uint32_t h, uint32_t x, uint32_t y, float rx, float ry,
                     float o[3], float d[3]) {
    float u = -fs->rw + ((fs->rw + fs->rw) * ((float)x + rx)) / (float)w;
    float v = -(fs->th + ((-fs->th - fs->th) * ((float)y + ry)) / (float)h);
    for (int i = 0; i < 3; ++i) {
        o[i] = fs->center[i];
        d[i] = (fs->nearp * fs->front[i] + u * fs->right[i]) + v * fs->up[i];
    }
    normalize3(d);
}

/* ------------------------------------------------------ intersection */
/* sphere.cuh:14-41 (Ray Tracing Gems ch. 7). Returns 1 and *t_out on a hit with tmin < t < tmax. */
int orc_intersect_f32(const float o[3], const float d[3], float tmin, float tmax, const float sph[4], float* t_out) {
    float oc[3] = {o[0] - sph[0], o[1] - sph[1], o[2] - sph[2]};
    float sq = sph[3] * sph[3];
    float noc[3] = {-oc[0], -oc[1], -oc[2]};
    float b = dot3(noc, d);
    float tmp[3] = {oc[0] + b * d[0], oc[1] + b * d[1], oc[2] + b * d[2]};
    float delta = sq - dot3(tmp, tmp);
    if (delta < 0.0f) return 0;
    float c = dot3(oc, oc) - sq;
    float q = b + copysignf(sqrtf(delta), b);
    float ta = c / q, tb = q;
    float t = fminf(ta, tb);
    if (t < 0.0f) return 0;
    if (t > tmin && t < tmax) {
        *t_out = t;
        return 1;
    }
    return 0;
}

/* FP64 ray-sphere with the same acceptance rules (nearest root < 0 -> miss). disc_out receives
 * the discriminant normalised by r^2 (0 = tangent). */
int orc_intersect_f64(const double o[3], const double d[3], double tmin, const double sph[4], double* t_out,
                      double* disc_out) {
    double oc[3] = {o[0] - sph[0], o[1] - sph[1], o[2] - sph[2]};
    double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double h = -(oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2]);
    double p[3] = {oc[0] + (h / a) * d[0], oc[1] + (h / a) * d[1], oc[2] + (h / a) * d[2]};
    double r2 = sph[3] * sph[3];
    double disc = r2 - (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    if (disc_out) *disc_out = disc / r2;
    if (disc < 0.0) return 0;
    double cc = (oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2]) - r2;
    double q = h + copysign(sqrt(a * disc), h);
    double t0 = cc / q, t1 = q / a;
    double t = t0 < t1 ? t0 : t1;
    if (!(t >= 0.0)) return 0;
    if (t > tmin) {
        *t_out = t;
        return 1;
    }
    return 0;
}

/* ----------------------------------------------------------- ingest */
static const uint32_t VSIZE[5] = {0, 12, 16, 6, 24};      /* SimpleSphericalParticles.cpp:7 */
static const uint32_t CSIZE[8] = {0, 3, 4, 12, 16, 4, 8, 8}; /* SimpleSphericalParticles.cpp:9 */

/* AddParticleColors.cpp:19-36 */
static float tf_lerp(float a, float b, float inter) { return a * (1.0f - inter) + b * inter; }
static void tf_sample_base(const float* tf, uint32_t n, int base, float rest, float out[4]) {
    if (base < 0 || n == 0) {
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        return;
    }
    uint32_t last = n - 1;
    if ((uint32_t)base >= last) {
        for (int c = 0; c < 4; ++c) out[c] = tf[last * 4 + c];
        return;
    }
    for (int c = 0; c < 4; ++c) out[c] = tf_lerp(tf[base * 4 + c], tf[(base + 1) * 4 + c], rest);
}
/* AddParticleColors.cpp:72-94: fac = 1/(max-min+1e-8), val = (i-min)*fac*size, modf split. */
void orc_tf_sample(const float* tf, uint32_t tf_n, float vmin, float vmax, float value, float out[4]) {
    float fac = 1.0f / (vmax - vmin + 1e-8f);
    float val = (value - vmin) * fac * (float)tf_n;
    float main_part = truncf(val);
    float rest = val - main_part; /* modf: exact for finite val */
    /* the reference's static_cast<int> is undefined beyond int range and for NaN: values at or
     * past the table end take the last entry, NaN and values below -1 take zero */
    if (main_part > (float)tf_n) main_part = (float)tf_n;
    if (!(main_part >= -1.0f)) main_part = -1.0f;
    tf_sample_base(tf, tf_n, (int)main_part, rest, out);
}

static double rd_f64(const uint8_t* p) { double v; memcpy(&v, p, 8); return v; }
static float rd_f32(const uint8_t* p) { float v; memcpy(&v, p, 4); return v; }
static uint16_t rd_u16(const uint8_t* p) { uint16_t v; memcpy(&v, p, 2); return v; }

/* SphereGeometry.cpp:94-146 with Accessor Get_f = static_cast<float> (Accessor.h:78-90):
 * integer colours are NOT normalised; global colours are /255 (SphereGeometry.cpp:179-181). */
int orc_convert_lists(const orc_list* lists, uint32_t n_lists, const float* tf, uint32_t tf_n, float tf_min,
                      float tf_max, int tf_auto_range, float* xyzr_out, float* rgba_out, uint64_t* n_out) {
    float vmin = tf_min, vmax = tf_max;
    if (tf_auto_range) { /* AddParticleColors.cpp:59-64 */
        vmin = FLT_MAX;
        vmax = -FLT_MAX;
        for (uint32_t l = 0; l < n_lists; ++l) {
            vmin = fminf(vmin, lists[l].col_min);
            vmax = fmaxf(vmax, lists[l].col_max);
        }
    }
    uint64_t k = 0;
    for (uint32_t l = 0; l < n_lists; ++l) {
        const orc_list* L = &lists[l];
        if (L->count == 0 || L->vert_type <= 0 || L->vert_type > 4) continue;
        uint64_t vs = L->vert_stride ? L->vert_stride : VSIZE[L->vert_type];
        uint64_t cs = L->col_stride ? L->col_stride : (L->col_type >= 0 && L->col_type < 8 ? CSIZE[L->col_type] : 0);
        int is_int = (L->col_type == 5 || L->col_type == 7);
        int has_color = L->col_type != 0 && !is_int;
        int use_tf = is_int && tf && tf_n > 0;
        for (uint64_t i = 0; i < L->count; ++i, ++k) {
            const uint8_t* v = (const uint8_t*)L->vert + i * vs;
            float* o = xyzr_out + 4 * k;
            switch (L->vert_type) {
            case 1: o[0] = rd_f32(v); o[1] = rd_f32(v + 4); o[2] = rd_f32(v + 8); o[3] = L->global_radius; break;
            case 2: o[0] = rd_f32(v); o[1] = rd_f32(v + 4); o[2] = rd_f32(v + 8); o[3] = rd_f32(v + 12); break;
            case 3: o[0] = (float)rd_u16(v); o[1] = (float)rd_u16(v + 2); o[2] = (float)rd_u16(v + 4); o[3] = L->global_radius; break;
            case 4: o[0] = (float)rd_f64(v); o[1] = (float)rd_f64(v + 8); o[2] = (float)rd_f64(v + 16); o[3] = L->global_radius; break;
            }
            float* c = rgba_out + 4 * k;
            const uint8_t* cp = (const uint8_t*)L->col + i * cs;
            if (use_tf) {
                float val = L->col_type == 5 ? rd_f32(cp) : (float)rd_f64(cp);
                orc_tf_sample(tf, tf_n, vmin, vmax, val, c);
            } else if (!has_color) {
                for (int q = 0; q < 4; ++q) c[q] = (float)L->global_rgba[q] / 255.0f;
            } else {
                switch (L->col_type) {
                case 1: c[0] = cp[0]; c[1] = cp[1]; c[2] = cp[2]; c[3] = 255.0f; break;
                case 2: c[0] = cp[0]; c[1] = cp[1]; c[2] = cp[2]; c[3] = cp[3]; break;
                case 3: c[0] = rd_f32(cp); c[1] = rd_f32(cp + 4); c[2] = rd_f32(cp + 8); c[3] = 1.0f; break;
                case 4: c[0] = rd_f32(cp); c[1] = rd_f32(cp + 4); c[2] = rd_f32(cp + 8); c[3] = rd_f32(cp + 12); break;
                case 6: c[0] = rd_u16(cp); c[1] = rd_u16(cp + 2); c[2] = rd_u16(cp + 4); c[3] = rd_u16(cp + 6); break;
                }
            }
        }
    }
    *n_out = k;
    return 0;
}

/* ------------------------------------------------------ pixel helpers */
static uint32_t to_rgba8(const float c[4]) {
    uint32_t r = 0;
    for (int i = 0; i < 4; ++i) {
        float v = fminf(fmaxf(c[i], 0.0f), 1.0f);
        uint32_t b = (uint32_t)(v * 255.0f + 0.5f);
        r |= b << (8 * i);
    }
    return r;
}
static uint16_t f2h(float f) { /* round-to-nearest-even float -> IEEE half */
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t mant = x & 0x7fffffu;
    int32_t exp = (int32_t)((x >> 23) & 0xff);
    if (exp == 0xff) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0));
    int32_t e = exp - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        mant |= 0x800000u;
        uint32_t shift = (uint32_t)(14 - e);
        uint32_t h = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1);
        uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (h & 1))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1))) h++;
    return (uint16_t)(sign | h);
}

/* Scene access abstraction: brute force or LBVH. */
typedef struct scene_view {
    const float* xyzr; /* brute force: original order */
    uint64_t n;
    const orc_bvh* bvh; /* NULL = brute force (or the treelet P-k-d trees below) */
    float pad;          /* slab padding for the LBVH and P-k-d traversals */
    uint64_t node_visits, sphere_tests;
    const orc_pkd* pkd; /* treelet P-k-d trees (when bvh is NULL) */
} scene_view;

static int closest_brute(scene_view* s, const float o[3], const float d[3], float tmin, float tmax, float* t_out,
                         int32_t* id_out) {
    float best = tmax;
    int32_t bid = -1;
    for (uint64_t i = 0; i < s->n; ++i) {
        float t;
        if (orc_intersect_f32(o, d, tmin, tmax, s->xyzr + 4 * i, &t)) {
            if (t < best || (t == best && (int32_t)i < bid)) {
                best = t;
                bid = (int32_t)i;
            }
        }
    }
    s->sphere_tests += s->n;
    *t_out = best;
    *id_out = bid;
    return bid >= 0;
}
static int any_brute(scene_view* s, const float o[3], const float d[3], float tmin, float tmax, int32_t skip) {
    for (uint64_t i = 0; i < s->n; ++i) {
        float t;
        if ((int32_t)i == skip) continue;
        s->sphere_tests++;
        if (orc_intersect_f32(o, d, tmin, tmax, s->xyzr + 4 * i, &t)) return 1;
    }
    return 0;
}

/* Robust slab test (padding per Ize 2013, "Robust BVH ray traversal"): the box is widened by
 * pad on every side, which covers the rounding of the FP32 sphere test and of this test. */
static int slab(const float* box, const float o[3], const float inv[3], float pad, float tmin, float tmax, float* tn) {
    float t0 = -FLT_MAX, t1 = FLT_MAX;
    for (int a = 0; a < 3; ++a) {
        float lo = (box[a] - pad - o[a]) * inv[a];
        float hi = (box[3 + a] + pad - o[a]) * inv[a];
        float n = fminf(lo, hi), f = fmaxf(lo, hi);
        t0 = fmaxf(t0, n);
        t1 = fminf(t1, f);
    }
    *tn = t0;
    return t0 <= t1 && t1 >= tmin && t0 <= tmax;
}

static void safe_inv(const float d[3], float inv[3]) {
    for (int a = 0; a < 3; ++a) {
        float v = d[a];
        if (fabsf(v) < 1e-30f) v = copysignf(1e-30f, v);
        inv[a] = 1.0f / v;
    }
}

static int closest_bvh(scene_view* s, const float o[3], const float d[3], float tmin, float tmax, float* t_out,
                       int32_t* id_out) {
    const orc_bvh* b = s->bvh;
    float best = tmax;
    int32_t bid = -1;
    if (b->n == 1) {
        float t;
        s->sphere_tests++;
        if (orc_intersect_f32(o, d, tmin, tmax, b->leaf_sph, &t)) {
            best = t;
            bid = (int32_t)b->leaf_id[0];
        }
        *t_out = best;
        *id_out = bid;
        return bid >= 0;
    }
    float inv[3];
    safe_inv(d, inv);
    int32_t stack[256];
    int sp = 0;
    int32_t node = 0;
    for (;;) {
        s->node_visits++;
        const float* bx = b->node_box + 12 * (size_t)node;
        const int32_t* ch = b->node_child + 2 * (size_t)node;
        float tn[2];
        int hit[2];
        for (int c = 0; c < 2; ++c) hit[c] = slab(bx + 6 * c, o, inv, s->pad, tmin, best, &tn[c]);
        int32_t next[2];
        int nn = 0;
        int order0 = (hit[0] && hit[1] && tn[1] < tn[0]) ? 1 : 0;
        for (int k = 0; k < 2; ++k) {
            int c = k ^ order0;
            if (!hit[c]) continue;
            if (tn[c] > best) continue;
            int32_t r = ch[c];
            if (r < 0) {
                uint32_t leaf = (uint32_t)(~r);
                float t;
                s->sphere_tests++;
                if (orc_intersect_f32(o, d, tmin, tmax, b->leaf_sph + 4 * (size_t)leaf, &t)) {
                    int32_t id = (int32_t)b->leaf_id[leaf];
                    if (t < best || (t == best && id < bid)) {
                        best = t;
                        bid = id;
                    }
                }
            } else {
                next[nn++] = r;
            }
        }
        if (nn == 2) {
            stack[sp++] = next[1];
            node = next[0];
        } else if (nn == 1) {
            node = next[0];
        } else {
            /* pop, skipping entries that can no longer hold a closer hit is not needed for
             * correctness; the slab test at the child rejects them. */
            if (sp == 0) break;
            node = stack[--sp];
        }
    }
    *t_out = best;
    *id_out = bid;
    return bid >= 0;
}

static int any_bvh(scene_view* s, const float o[3], const float d[3], float tmin, float tmax, int32_t skip) {
    const orc_bvh* b = s->bvh;
    if (b->n == 1) {
        float t;
        if ((int32_t)b->leaf_id[0] == skip) return 0;
        s->sphere_tests++;
        return orc_intersect_f32(o, d, tmin, tmax, b->leaf_sph, &t);
    }
    float inv[3];
    safe_inv(d, inv);
    int32_t stack[256];
    int sp = 0;
    int32_t node = 0;
    for (;;) {
        s->node_visits++;
        const float* bx = b->node_box + 12 * (size_t)node;
        const int32_t* ch = b->node_child + 2 * (size_t)node;
        int32_t next[2];
        int nn = 0;
        for (int c = 0; c < 2; ++c) {
            float tn;
            if (!slab(bx + 6 * c, o, inv, s->pad, tmin, tmax, &tn)) continue;
            int32_t r = ch[c];
            if (r < 0) {
                uint32_t leaf = (uint32_t)(~r);
                if ((int32_t)b->leaf_id[leaf] == skip) continue;
                float t;
                s->sphere_tests++;
                if (orc_intersect_f32(o, d, tmin, tmax, b->leaf_sph + 4 * (size_t)leaf, &t)) return 1;
            } else {
                next[nn++] = r;
            }
        }
        if (nn == 2) {
            stack[sp++] = next[1];
            node = next[0];
        } else if (nn == 1) {
            node = next[0];
        } else {
            if (sp == 0) return 0;
            node = stack[--sp];
        }
    }
}

/* ------------------------------------------------ treelet P-k-d traversal */
/* Box slab test returning the entry and exit distances (box widened by pad as in slab()). */
static int slab2(const float* box, const float o[3], const float inv[3], float pad, float tmin, float tmax, float* tn,
                 float* tf) {
    float t0 = -FLT_MAX, t1 = FLT_MAX;
    for (int a = 0; a < 3; ++a) {
        float lo = (box[a] - pad - o[a]) * inv[a];
        float hi = (box[3 + a] + pad - o[a]) * inv[a];
        t0 = fmaxf(t0, fminf(lo, hi));
        t1 = fminf(t1, fmaxf(lo, hi));
    }
    *tn = t0;
    *tf = t1;
    return t0 <= t1 && t1 >= tmin && t0 <= tmax;
}

/* One treelet's P-k-d tree for a ray entering its box over [t0, t1]: the interval logic of
 * pkd_intersect (plugins/optix_owl/cuda_resources/pkdPrograms.cu:45-159, treeletsPrograms.cu
 * treelets_intersect): the node's sphere is tested when the ray's interval meets the node's slab
 * (split coordinate -+ the largest radius, here the treelet's, widened by pad); the near child
 * takes [t0, min(slab exit, t1, hit)], the far child [max(t0, slab entry), min(t1, hit)]. Closed
 * intervals throughout (a subtree is skipped only when its interval is empty or starts beyond the
 * best hit), so that equal-t hits still reach the lowest particle id. any: stop at the first hit
 * of a sphere other than `skip` (returns 1). */
static int pkd_treelet(scene_view* s, uint32_t t, const float o[3], const float d[3], const float inv[3], float tmin,
                       float tmax, float t0, float t1, float* best, int32_t* bid, int any, int32_t skip) {
    const orc_pkd* k = s->pkd;
    const uint64_t base = (uint64_t)t * k->M;
    const uint32_t m = (uint32_t)((k->n - base) < k->M ? (k->n - base) : k->M);
    const float rmax = k->tl_rmax[t];
    struct {
        uint32_t node;
        float t0, t1;
    } st[64];
    int sp = 0;
    uint32_t i = 0;
    for (;;) {
        for (;;) {
            const uint64_t p = base + i;
            const float* sph = k->sph + 4 * p;
            const int dd = k->dim[p];
            s->node_visits++;
            float lo = (sph[dd] - rmax - s->pad - o[dd]) * inv[dd];
            float hi = (sph[dd] + rmax + s->pad - o[dd]) * inv[dd];
            float tnr = fminf(lo, hi), tfr = fmaxf(lo, hi);
            if (fmaxf(t0, tnr) <= fminf(fminf(tfr, t1), *best)) {
                int32_t id = (int32_t)k->id[p];
                float th;
                if (!(any && id == skip)) {
                    s->sphere_tests++;
                    if (orc_intersect_f32(o, d, tmin, tmax, sph, &th)) {
                        if (any) return 1;
                        if (th < *best || (th == *best && id < *bid)) {
                            *best = th;
                            *bid = id;
                        }
                    }
                }
            }
            float n_t1 = fminf(fminf(tfr, t1), *best);
            float f_t0 = fmaxf(t0, tnr), f_t1 = fminf(t1, *best);
            int sgn = inv[dd] < 0.0f;
            uint32_t nc = 2 * i + 1 + (uint32_t)sgn, fc = 2 * i + 2 - (uint32_t)sgn;
            int need_n = nc < m && t0 <= n_t1;
            int need_f = fc < m && f_t0 <= f_t1;
            if (!need_n && !need_f) break;
            if (need_n && need_f) {
                st[sp].node = fc;
                st[sp].t0 = f_t0;
                st[sp].t1 = f_t1;
                ++sp;
                i = nc;
                t1 = n_t1;
            } else if (need_n) {
                i = nc;
                t1 = n_t1;
            } else {
                i = fc;
                t0 = f_t0;
                t1 = f_t1;
            }
        }
        for (;;) {
            if (sp == 0) return 0;
            --sp;
            t0 = st[sp].t0;
            t1 = fminf(st[sp].t1, *best);
            i = st[sp].node;
            if (t0 <= t1) break;
        }
    }
}

/* The top tree of boxes over the treelets, front to back like closest_bvh; a treelet child is
 * traversed as soon as its box is reached. */
static int trace_pkd(scene_view* s, const float o[3], const float d[3], float tmin, float tmax, float* t_out,
                     int32_t* id_out, int any, int32_t skip) {
    const orc_pkd* k = s->pkd;
    float inv[3];
    safe_inv(d, inv);
    float best = tmax;
    int32_t bid = -1;
    int32_t stack[64];
    int sp = 0;
    int32_t node = 0;
    for (;;) {
        s->node_visits++;
        const float* bx = k->top_box + 12 * (size_t)node;
        const int32_t* ch = k->top_child + 2 * (size_t)node;
        float tn[2], tf[2];
        int hit[2];
        for (int c = 0; c < 2; ++c)
            hit[c] = ch[c] != INT32_MIN && slab2(bx + 6 * c, o, inv, s->pad, tmin, best, &tn[c], &tf[c]);
        int32_t next[2];
        int nn = 0;
        int order0 = (hit[0] && hit[1] && tn[1] < tn[0]) ? 1 : 0;
        for (int q = 0; q < 2; ++q) {
            int c = q ^ order0;
            if (!hit[c] || tn[c] > best) continue;
            if (ch[c] < 0) {
                if (pkd_treelet(s, (uint32_t)(~ch[c]), o, d, inv, tmin, tmax, fmaxf(tn[c], tmin), tf[c], &best, &bid,
                                any, skip))
                    return 1;
            } else {
                next[nn++] = ch[c];
            }
        }
        if (nn == 2) {
            stack[sp++] = next[1];
            node = next[0];
        } else if (nn == 1) {
            node = next[0];
        } else {
            if (sp == 0) break;
            node = stack[--sp];
        }
    }
    if (any) return 0;
    *t_out = best;
    *id_out = bid;
    return bid >= 0;
}

/* closest hit dispatcher; the tie rule "equal t -> lowest particle id" is applied by every path */
static int closest(scene_view* s, const float o[3], const float d[3], float tmin, float tmax, float* t, int32_t* id) {
    if (s->pkd) return trace_pkd(s, o, d, tmin, tmax, t, id, 0, -1);
    return s->bvh ? closest_bvh(s, o, d, tmin, tmax, t, id) : closest_brute(s, o, d, tmin, tmax, t, id);
}
static int anyhit(scene_view* s, const float o[3], const float d[3], float tmin, float tmax, int32_t skip) {
    if (s->pkd) return trace_pkd(s, o, d, tmin, tmax, NULL, NULL, 1, skip);
    return s->bvh ? any_bvh(s, o, d, tmin, tmax, skip) : any_brute(s, o, d, tmin, tmax, skip);
}

static const float* sphere_of(const scene_view* s, int32_t id, const float* xyzr_orig) {
    (void)s;
    return xyzr_orig + 4 * (size_t)id;
}

/* sin and cos on [-pi/4, pi/4] by their Taylor series through x^11 / x^12 (error below 1e-9),
 * one fixed evaluation order, so that the device code (rt_math.hpp) gives the same bits. */
static void sincos_q(float x, float* s, float* c) {
    float x2 = x * x;
    *s = x * (1.0f + x2 * (-1.0f / 6.0f + x2 * (1.0f / 120.0f + x2 * (-1.0f / 5040.0f + x2 * (1.0f / 362880.0f +
                                                                                         x2 * (-1.0f / 39916800.0f))))));
    *c = 1.0f + x2 * (-0.5f + x2 * (1.0f / 24.0f + x2 * (-1.0f / 720.0f + x2 * (1.0f / 40320.0f +
                                                                               x2 * (-1.0f / 3628800.0f +
                                                                                     x2 * (1.0f / 479001600.0f))))));
}

/* The AO sample direction, cosine-weighted about the unit normal n. Sample smp of spp takes the
 * stratum (smp % g, smp / g) of a g x g grid over the unit square, g = floor(sqrt(spp)) (samples
 * past g*g are not stratified), jittered by two draws of the sample's own stream. The square maps
 * to the disk by the concentric map (Shirley & Chiu 1997, "A low distortion map between disk and
 * square"), the disk to the hemisphere by Malley's method (z = sqrt(1 - r^2)), and the hemisphere
 * to world space through the branch-free orthonormal basis of Duff et al. (2017, "Building an
 * orthonormal basis, revisited"). Only +,*,/,sqrt and a polynomial sin/cos are used, so that
 * CPU and GPU agree bit for bit. */
static void ao_direction(const float n[3], uint32_t* seed, uint32_t smp, uint32_t spp, float dir[3]) {
    uint32_t g = (uint32_t)sqrtf((float)spp);
    while (g * g > spp) --g;
    while ((g + 1) * (g + 1) <= spp) ++g;
    float r1 = orc_rnd(seed);
    float r2 = orc_rnd(seed);
    float u1 = r1, u2 = r2;
    if (smp < g * g) {
        float gf = (float)g;
        u1 = ((float)(smp % g) + r1) / gf;
        u2 = ((float)(smp / g) + r2) / gf;
    }
    float a = 2.0f * u1 - 1.0f, b = 2.0f * u2 - 1.0f;
    float dx = 0.0f, dy = 0.0f;
    if (a != 0.0f || b != 0.0f) {
        const float q = 0.785398163397448f; /* pi / 4 */
        float r, sn, cs;
        if (fabsf(a) > fabsf(b)) {
            r = a;
            sincos_q(q * (b / a), &sn, &cs);
            dx = r * cs;
            dy = r * sn;
        } else {
            r = b;
            sincos_q(q * (a / b), &sn, &cs); /* phi = pi/2 - (pi/4) a/b */
            dx = r * sn;
            dy = r * cs;
        }
    }
    float dz = sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy));
    float sg = copysignf(1.0f, n[2]);
    float ia = -1.0f / (sg + n[2]);
    float bb = n[0] * n[1] * ia;
    float t[3] = {1.0f + sg * n[0] * n[0] * ia, sg * bb, -sg * n[0]};
    float bt[3] = {bb, sg + n[1] * n[1] * ia, -n[1]};
    for (int k = 0; k < 3; ++k) dir[k] = (dx * t[k] + dy * bt[k]) + dz * n[k];
    float l2 = dot3(dir, dir);
    if (!(l2 > 1e-12f)) {
        dir[0] = n[0];
        dir[1] = n[1];
        dir[2] = n[2];
    } else {
        float inv = 1.0f / sqrtf(l2);
        dir[0] = dir[0] * inv;
        dir[1] = dir[1] * inv;
        dir[2] = dir[2] * inv;
    }
}

/* exported for the checks of the AO sampler */
void orc_ao_direction(const float n[3], uint32_t seed, uint32_t smp, uint32_t spp, float dir[3]) {
    ao_direction(n, &seed, smp, spp, dir);
}

typedef struct pixel_out {
    float col[4];
    float depth, t;
    int32_t id;
    float aux[8];
} pixel_out;

/* One pixel: raygen.cu:182-268 for mode 0; the AO mode keeps the same primary-ray stream and
 * adds one AO ray per sample from its own TEA-seeded stream. */
static void render_pixel(scene_view* s, const float* xyzr, const float* rgba, const orc_frame_state* fs,
                         const orc_params* p, uint32_t x, uint32_t y, pixel_out* po) {
    uint32_t pix = y * p->width + x;
    uint32_t seed = orc_tea16(pix, p->frame_idx);
    float col[4] = {0, 0, 0, 0};
    float depth = FLT_MAX;
    float best_t = INFINITY;
    int32_t best_id = -1;
    float aux[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    float vis_sum = 0.0f;
    for (uint32_t smp = 0; smp < p->spp; ++smp) {
        float rx = orc_rnd(&seed);
        float ry = orc_rnd(&seed);
        float o[3], d[3];
        orc_primary_ray(fs, p->width, p->height, x, y, rx, ry, o, d);
        float t;
        int32_t id;
        float sc[4];
        float ray_depth = FLT_MAX;
        if (!closest(s, o, d, 1e-6f, 1e20f, &t, &id)) {
            if (p->mode == 1) {
                sc[0] = p->background[0]; sc[1] = p->background[1]; sc[2] = p->background[2]; sc[3] = p->background[3];
            } else {
                sc[0] = sc[1] = sc[2] = 1.0f * 0.8f; /* raygen.cu:131-133 */
                sc[3] = 1.0f;
            }
        } else {
            ray_depth = t; /* set_depth, utils_device.h:217-226 */
            const float* sp = sphere_of(s, id, xyzr);
            float P[3], N[3];
            for (int a = 0; a < 3; ++a) {
                P[a] = o[a] + t * d[a];
                N[a] = P[a] - sp[a];
            }
            if (dot3(N, d) > 0.0f) {
                N[0] = -N[0]; N[1] = -N[1]; N[2] = -N[2];
            }
            normalize3(N);
            const float* alb = rgba + 4 * (size_t)id;
            if (p->mode == 1) {
                uint32_t aseed = orc_tea16(pix, 0xA0000000u + p->frame_idx * 4096u + smp);
                float ad[3];
                ao_direction(N, &aseed, smp, p->spp, ad);
                int occ = anyhit(s, P, ad, 0.0f, p->ao_distance, id);
                float V = occ ? 0.0f : 1.0f;
                sc[0] = alb[0] * V; sc[1] = alb[1] * V; sc[2] = alb[2] * V; sc[3] = 1.0f;
                vis_sum += V;
            } else if (p->mode == 2 && p->max_bounces > 0) {
                /* raygen.cu:122-155: col *= albedo per hit; miss -> col * 0.8; depth cap -> 0.1 */
                float col3[3] = {alb[0], alb[1], alb[2]};
                float ro[3] = {o[0], o[1], o[2]}, rd[3] = {d[0], d[1], d[2]}, rt = t, n[3] = {N[0], N[1], N[2]};
                for (int depth_i = 0;; ++depth_i) {
                    if (depth_i >= p->max_bounces) {
                        sc[0] = sc[1] = sc[2] = 0.1f;
                        break;
                    }
                    float so[3], u[3], sd[3];
                    for (int a = 0; a < 3; ++a) so[a] = ro[a] + rt * rd[a];
                    /* random_in_unit_sphere (raygen.cu:71-77), draws taken x, y, z in order */
                    do {
                        float rx = orc_rnd(&seed);
                        float ry = orc_rnd(&seed);
                        float rz = orc_rnd(&seed);
                        u[0] = 2.0f * rx - 1.0f;
                        u[1] = 2.0f * ry - 1.0f;
                        u[2] = 2.0f * rz - 1.0f;
                    } while (dot3(u, u) >= 1.0f);
                    for (int a = 0; a < 3; ++a) sd[a] = n[a] + u[a];
                    normalize3(sd);
                    float tb;
                    int32_t ib;
                    if (!closest(s, so, sd, 1e-3f, 1e8f, &tb, &ib)) {
                        for (int a = 0; a < 3; ++a) sc[a] = col3[a] * 0.8f;
                        break;
                    }
                    const float* sb = sphere_of(s, ib, xyzr);
                    const float* ab = rgba + 4 * (size_t)ib;
                    float Pb[3], Nb[3];
                    for (int a = 0; a < 3; ++a) {
                        Pb[a] = so[a] + tb * sd[a];
                        Nb[a] = Pb[a] - sb[a];
                    }
                    if (dot3(Nb, sd) > 0.0f) {
                        Nb[0] = -Nb[0]; Nb[1] = -Nb[1]; Nb[2] = -Nb[2];
                    }
                    normalize3(Nb);
                    for (int a = 0; a < 3; ++a) {
                        col3[a] = col3[a] * ab[a];
                        ro[a] = so[a];
                        rd[a] = sd[a];
                        n[a] = Nb[a];
                    }
                    rt = tb;
                }
                sc[3] = 1.0f;
            } else {
                float sh = 0.2f + 0.6f * fabsf(dot3(N, d)); /* raygen.cu:141 */
                sc[0] = alb[0] * sh; sc[1] = alb[1] * sh; sc[2] = alb[2] * sh; sc[3] = 1.0f;
            }
            if (t < best_t) {
                best_t = t;
                best_id = id;
            }
            if (smp == 0) {
                aux[0] = alb[0]; aux[1] = alb[1]; aux[2] = alb[2];
                aux[3] = N[0]; aux[4] = N[1]; aux[5] = N[2];
                aux[6] = t;
            }
        }
        for (int c = 0; c < 4; ++c) col[c] += sc[c];
        depth = fminf(depth, ray_depth);
    }
    for (int c = 0; c < 4; ++c) col[c] /= (float)p->spp;
    if (depth < FLT_MAX) {
        depth = (fs->depth_c / depth) - fs->depth_a;
        depth = 0.5f * (depth + 1.0f);
    } else {
        depth = 1.0f;
        for (int c = 0; c < 4; ++c) col[c] = p->background[c];
    }
    aux[7] = vis_sum / (float)p->spp;
    memcpy(po->col, col, sizeof col);
    po->depth = depth;
    po->t = best_t;
    po->id = best_id;
    memcpy(po->aux, aux, sizeof aux);
}

static void store_pixel(const pixel_out* po, size_t i, uint32_t* rgba8, float* depth, float* color, float* t_out,
                        int32_t* id_out, uint16_t* aux) {
    if (rgba8) rgba8[i] = to_rgba8(po->col);
    if (depth) depth[i] = po->depth;
    if (color) memcpy(color + 4 * i, po->col, 16);
    if (t_out) t_out[i] = po->t;
    if (id_out) id_out[i] = po->id;
    if (aux)
        for (int c = 0; c < 8; ++c) aux[8 * i + c] = f2h(po->aux[c]);
}

/* Progressive blend of raygen.cu:254-268 over a whole frame: a pixel that some sample of the new
 * frame hit becomes lerp(old, new, 1 / (frame_idx + 1)) with lerp(a, b, t) = a + t * (b - a)
 * (utils_device.h:152-154); a pixel that no sample hit takes the background (the override comes
 * after the blend). frame_idx 0 keeps the new colour. rgba8 (may be NULL) receives the packed
 * result. */
void orc_blend_frames(const float* old_col, const float* new_col, const int32_t* hit_id, uint64_t n_px,
                      uint32_t frame_idx, const float bg[4], float* out_col, uint32_t* rgba8) {
    const float a = 1.0f / (float)(frame_idx + 1);
    for (uint64_t i = 0; i < n_px; ++i) {
        float c[4];
        for (int k = 0; k < 4; ++k) {
            float nv = new_col[4 * i + k];
            c[k] = frame_idx > 0 ? old_col[4 * i + k] + a * (nv - old_col[4 * i + k]) : nv;
            if (hit_id[i] < 0) c[k] = bg[k];
        }
        memcpy(out_col + 4 * i, c, sizeof c);
        if (rgba8) rgba8[i] = to_rgba8(c);
    }
}

/* RGBA8 packing of the framebuffer (clamp to [0,1], round half up). */
void orc_pack_rgba8(const float* col, uint64_t n_px, uint32_t* out) {
    for (uint64_t i = 0; i < n_px; ++i) out[i] = to_rgba8(col + 4 * i);
}

int orc_render_brute(const float* xyzr, const float* rgba, uint64_t n, const orc_camera* cam, const orc_params* p,
                     uint32_t* rgba8, float* depth, float* color, float* t_out, int32_t* id_out, uint16_t* aux) {
    orc_frame_state fs;
    orc_frame_state_from_camera(cam, &fs);
    scene_view s = {xyzr, n, NULL, 0.0f, 0, 0, NULL};
    for (uint32_t y = 0; y < p->height; ++y)
        for (uint32_t x = 0; x < p->width; ++x) {
            pixel_out po;
            render_pixel(&s, xyzr, rgba, &fs, p, x, y, &po);
            store_pixel(&po, (size_t)y * p->width + x, rgba8, depth, color, t_out, id_out, aux);
        }
    return 0;
}

/* FP64 primary hit per pixel (first sample's jittered ray, generated in FP32 as the reference
 * does): nearest id, its t, the normalised discriminant and the second-nearest t (ties). */
int orc_primary_f64(const float* xyzr, uint64_t n, const orc_camera* cam, const orc_params* p, uint32_t x0,
                    uint32_t y0, uint32_t x1, uint32_t y1, int32_t* id_out, double* t_out, double* disc_out,
                    double* t2_out) {
    orc_frame_state fs;
    orc_frame_state_from_camera(cam, &fs);
    size_t k = 0;
    for (uint32_t y = y0; y < y1; ++y)
        for (uint32_t x = x0; x < x1; ++x, ++k) {
            uint32_t seed = orc_tea16(y * p->width + x, p->frame_idx);
            float rx = orc_rnd(&seed), ry = orc_rnd(&seed);
            float of[3], df[3];
            orc_primary_ray(&fs, p->width, p->height, x, y, rx, ry, of, df);
            double o[3] = {of[0], of[1], of[2]}, d[3] = {df[0], df[1], df[2]};
            double best = INFINITY, second = INFINITY, bdisc = 0.0;
            int32_t bid = -1;
            for (uint64_t i = 0; i < n; ++i) {
                double sp[4] = {xyzr[4 * i], xyzr[4 * i + 1], xyzr[4 * i + 2], xyzr[4 * i + 3]};
                double t, disc;
                if (orc_intersect_f64(o, d, 1e-6, sp, &t, &disc) && t < 1e20) {
                    if (t < best) {
                        second = best;
                        best = t;
                        bid = (int32_t)i;
                        bdisc = disc;
                    } else if (t < second) {
                        second = t;
                    }
                }
            }
            id_out[k] = bid;
            t_out[k] = best;
            disc_out[k] = bdisc;
            t2_out[k] = second;
        }
    return 0;
}

/* The same FP64 nearest hit for the pixels of k rectangles (rects: x0, y0, x1, y1 each; outputs
 * rectangle by rectangle, row-major inside each), over very large scenes: threads split the
 * spheres into contiguous ranges, and a sphere is tested against a rectangle's rays only if it
 * meets the cone from the camera that holds all of them (a conservative bound: axis a, half
 * angle th; a sphere at axial distance s and distance q from the axis can meet the cone only if
 * q <= max(s, 0) tan th + r / cos th). Per-thread results merge in range order, so equal t goes
 * to the lowest index as in orc_primary_f64. */
typedef struct f64_rect {
    uint32_t x0, y0, x1, y1;
    size_t first; /* first output slot */
    double axis[3], tan_th, inv_cos;
} f64_rect;
typedef struct f64_job {
    const float* xyzr;
    uint64_t n;
    double o[3];
    const double* dirs; /* 3 per output pixel */
    const f64_rect* rects;
    uint32_t k;
    size_t npx;
    double *best, *second, *disc; /* per thread: npx each */
    int32_t* id;
} f64_job;

static void f64_worker(void* a, int t, int T) {
    f64_job* J = a;
    PAR_RANGE(J->n, t, T, i0, i1);
    double* best = J->best + (size_t)t * J->npx;
    double* second = J->second + (size_t)t * J->npx;
    double* disc = J->disc + (size_t)t * J->npx;
    int32_t* id = J->id + (size_t)t * J->npx;
    for (size_t q = 0; q < J->npx; ++q) {
        best[q] = second[q] = INFINITY;
        disc[q] = 0.0;
        id[q] = -1;
    }
    for (uint64_t i = i0; i < i1; ++i) {
        double sp[4] = {J->xyzr[4 * i], J->xyzr[4 * i + 1], J->xyzr[4 * i + 2], J->xyzr[4 * i + 3]};
        double w[3] = {sp[0] - J->o[0], sp[1] - J->o[1], sp[2] - J->o[2]};
        double ww = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        for (uint32_t r = 0; r < J->k; ++r) {
            const f64_rect* R = &J->rects[r];
            double s = w[0] * R->axis[0] + w[1] * R->axis[1] + w[2] * R->axis[2];
            double q2 = ww - s * s;
            double lim = (s > 0.0 ? s : 0.0) * R->tan_th + fabs(sp[3]) * R->inv_cos;
            lim = lim * (1.0 + 1e-9) + 1e-9 * sqrt(ww) + 1e-12;
            if (q2 > lim * lim) continue;
            size_t q = R->first;
            for (uint32_t y = R->y0; y < R->y1; ++y)
                for (uint32_t x = R->x0; x < R->x1; ++x, ++q) {
                    double tt, dd;
                    if (orc_intersect_f64(J->o, J->dirs + 3 * q, 1e-6, sp, &tt, &dd) && tt < 1e20) {
                        if (tt < best[q]) {
                            second[q] = best[q];
                            best[q] = tt;
                            id[q] = (int32_t)i;
                            disc[q] = dd;
                        } else if (tt < second[q]) {
                            second[q] = tt;
                        }
                    }
                }
        }
    }
}

int orc_primary_f64_rects(const float* xyzr, uint64_t n, const orc_camera* cam, const orc_params* p,
                          const uint32_t* rects, uint32_t k, int threads, int32_t* id_out, double* t_out,
                          double* disc_out, double* t2_out) {
    orc_frame_state fs;
    orc_frame_state_from_camera(cam, &fs);
    f64_rect* R = calloc(k ? k : 1, sizeof *R);
    size_t npx = 0;
    for (uint32_t r = 0; r < k; ++r) {
        R[r].x0 = rects[4 * r];
        R[r].y0 = rects[4 * r + 1];
        R[r].x1 = rects[4 * r + 2];
        R[r].y1 = rects[4 * r + 3];
        if (R[r].x1 > p->width || R[r].y1 > p->height || R[r].x0 > R[r].x1 || R[r].y0 > R[r].y1) {
            free(R);
            return -1;
        }
        R[r].first = npx;
        npx += (size_t)(R[r].x1 - R[r].x0) * (R[r].y1 - R[r].y0);
    }
    double* dirs = malloc((npx ? npx : 1) * 3 * sizeof(double));
    double o[3] = {fs.center[0], fs.center[1], fs.center[2]};
    for (uint32_t r = 0; r < k; ++r) {
        double ax[3] = {0, 0, 0};
        size_t q = R[r].first;
        for (uint32_t y = R[r].y0; y < R[r].y1; ++y)
            for (uint32_t x = R[r].x0; x < R[r].x1; ++x, ++q) {
                /* the first sample's jittered ray, generated in FP32 as the reference does */
                uint32_t seed = orc_tea16(y * p->width + x, p->frame_idx);
                float rx = orc_rnd(&seed), ry = orc_rnd(&seed);
                float of[3], df[3];
                orc_primary_ray(&fs, p->width, p->height, x, y, rx, ry, of, df);
                for (int a = 0; a < 3; ++a) {
                    dirs[3 * q + a] = df[a];
                    ax[a] += df[a];
                }
            }
        double len = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
        double cmin = 1.0;
        for (int a = 0; a < 3; ++a) R[r].axis[a] = len > 0 ? ax[a] / len : 0.0;
        for (size_t j = R[r].first; j < q; ++j) {
            const double* d = dirs + 3 * j;
            double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            double c = (d[0] * R[r].axis[0] + d[1] * R[r].axis[1] + d[2] * R[r].axis[2]) / dl;
            if (c < cmin) cmin = c;
        }
        if (len == 0.0 || cmin < 0.05) { /* degenerate or very wide: no culling */
            R[r].tan_th = 1e300;
            R[r].inv_cos = 1e300;
        } else {
            cmin = cmin * (1.0 - 1e-12) - 1e-12;
            R[r].tan_th = sqrt(fmax(0.0, 1.0 - cmin * cmin)) / cmin;
            R[r].inv_cos = 1.0 / cmin;
        }
    }
    int T = threads_for(threads, n);
    f64_job J = {xyzr, n, {o[0], o[1], o[2]}, dirs, R, k, npx, NULL, NULL, NULL, NULL};
    size_t tot = (size_t)T * (npx ? npx : 1);
    J.best = malloc(tot * sizeof(double));
    J.second = malloc(tot * sizeof(double));
    J.disc = malloc(tot * sizeof(double));
    J.id = malloc(tot * sizeof(int32_t));
    par_run(T, f64_worker, &J);
    for (size_t q = 0; q < npx; ++q) {
        double b = INFINITY, s = INFINITY, dd = 0.0;
        int32_t id = -1;
        for (int t = 0; t < T; ++t) { /* ranges in index order: a later range wins only if strictly nearer */
            size_t j = (size_t)t * npx + q;
            if (J.best[j] < b) {
                s = fmin(b, J.second[j]);
                b = J.best[j];
                id = J.id[j];
                dd = J.disc[j];
            } else {
                s = fmin(s, J.best[j]);
            }
        }
        id_out[q] = id;
        t_out[q] = b;
        disc_out[q] = dd;
        t2_out[q] = s;
    }
    free(J.best);
    free(J.second);
    free(J.disc);
    free(J.id);
    free(dirs);
    free(R);
    return 0;
}

/* --------------------------------------------------------------- LBVH */
/* Morton codes (Morton 1966) over the centre bounds with one uniform scale, then a stable sort
 * by key, the Karras (2012, "Maximizing parallelism in the construction of BVHs, octrees and
 * k-d trees") hierarchy with the index tie-break for equal keys, and bottom-up boxes. */
static uint32_t expand10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
static uint64_t expand21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

static int delta_k(const uint64_t* keys, uint64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= (int64_t)n) return -1;
    uint64_t a = keys[i], b = keys[j];
    if (a == b) return 64 + __builtin_clz((uint32_t)i ^ (uint32_t)j);
    return __builtin_clzll(a ^ b);
}

typedef struct build_job {
    const float* xyzr;
    uint64_t n, m;
    int bits, shift;
    float tlo[256][3], thi[256][3], tamax[256];
    float lo[3], scale;
    uint32_t maxq;
    uint64_t *key, *key2;
    uint32_t *idx, *idx2;
    size_t (*hist)[256]; /* per thread digit counts, then scatter offsets */
    orc_bvh* out;
    int32_t* parent;
    atomic_uchar* seen;
} build_job;

static void bj_bounds(void* a, int t, int T) {
    build_job* J = a;
    PAR_RANGE(J->n, t, T, i0, i1);
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float am = 0.0f;
    for (uint64_t i = i0; i < i1; ++i)
        for (int ax = 0; ax < 3; ++ax) {
            lo[ax] = fminf(lo[ax], J->xyzr[4 * i + ax]);
            hi[ax] = fmaxf(hi[ax], J->xyzr[4 * i + ax]);
            am = fmaxf(am, fabsf(J->xyzr[4 * i + ax]) + fabsf(J->xyzr[4 * i + 3]));
        }
    memcpy(J->tlo[t], lo, sizeof lo);
    memcpy(J->thi[t], hi, sizeof hi);
    J->tamax[t] = am;
}
static void bj_keys(void* a, int t, int T) {
    build_job* J = a;
    PAR_RANGE(J->n, t, T, i0, i1);
    for (uint64_t i = i0; i < i1; ++i) {
        uint32_t q[3];
        for (int ax = 0; ax < 3; ++ax) {
            float f = (J->xyzr[4 * i + ax] - J->lo[ax]) * J->scale;
            uint32_t u = (uint32_t)f;
            q[ax] = u > J->maxq ? J->maxq : u;
        }
        J->key[i] = J->bits == 63 ? ((expand21(q[0]) << 2) | (expand21(q[1]) << 1) | expand21(q[2]))
                                  : (uint64_t)((expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]));
        J->idx[i] = (uint32_t)i;
    }
}
static void bj_hist(void* a, int t, int T) {
    build_job* J = a;
    PAR_RANGE(J->n, t, T, i0, i1);
    size_t* h = J->hist[t];
    memset(h, 0, 256 * sizeof(size_t));
    for (uint64_t i = i0; i < i1; ++i) h[(J->key[i] >> J->shift) & 0xff]++;
}
static void bj_scatter(void* a, int t, int T) {
    build_job* J = a;
    PAR_RANGE(J->n, t, T, i0, i1);
    size_t* off = J->hist[t];
    for (uint64_t i = i0; i < i1; ++i) {
        size_t dst = off[(J->key[i] >> J->shift) & 0xff]++;
        J->key2[dst] = J->key[i];
        J->idx2[dst] = J->idx[i];
    }
}
static void bj_gather(void* a, int t, int T) {
    build_job* J = a;
    PAR_RANGE(J->n, t, T, k0, k1);
    for (uint64_t k = k0; k < k1; ++k)
        memcpy(J->out->leaf_sph + 4 * k, J->xyzr + 4 * (size_t)J->out->leaf_id[k], 16);
}
/* Karras (2012) node i: direction, range end by exponential + binary search, split by binary
 * search over the common prefix of the augmented keys. */
static void bj_nodes(void* a, int t, int T) {
    build_job* J = a;
    const uint64_t* key = J->out->keys;
    uint64_t n = J->n, m = J->m;
    PAR_RANGE(m, t, T, i0, i1);
    for (int64_t i = (int64_t)i0; i < (int64_t)i1; ++i) {
        int dir = (delta_k(key, n, i, i + 1) - delta_k(key, n, i, i - 1)) >= 0 ? 1 : -1;
        int dmin = delta_k(key, n, i, i - dir);
        int64_t lmax = 2;
        while (delta_k(key, n, i, i + lmax * dir) > dmin) lmax *= 2;
        int64_t l = 0;
        for (int64_t s = lmax / 2; s >= 1; s /= 2)
            if (delta_k(key, n, i, i + (l + s) * dir) > dmin) l += s;
        int64_t j = i + l * dir;
        int dnode = delta_k(key, n, i, j);
        int64_t s = 0, w = l;
        do {
            w = (w + 1) >> 1;
            if (delta_k(key, n, i, i + (s + w) * dir) > dnode) s += w;
        } while (w > 1);
        int64_t gamma = i + s * dir + (dir < 0 ? -1 : 0);
        int64_t first = i < j ? i : j, last = i < j ? j : i;
        int32_t left = (first == gamma) ? ~(int32_t)gamma : (int32_t)gamma;
        int32_t right = (last == gamma + 1) ? ~(int32_t)(gamma + 1) : (int32_t)(gamma + 1);
        J->out->node_child[2 * i] = left;
        J->out->node_child[2 * i + 1] = right;
        J->parent[left < 0 ? m + (uint64_t)(~left) : (uint64_t)left] = (int32_t)i;
        J->parent[right < 0 ? m + (uint64_t)(~right) : (uint64_t)right] = (int32_t)i;
    }
}
/* Bottom-up boxes: every leaf climbs; the first arrival at a node stops, the second (which
 * finds the other child's box stored) computes the node's box and climbs on. */
static void bj_boxes(void* a, int t, int T) {
    build_job* J = a;
    orc_bvh* out = J->out;
    uint64_t m = J->m;
    PAR_RANGE(J->n, t, T, k0, k1);
    for (uint64_t k = k0; k < k1; ++k) {
        const float* sp = out->leaf_sph + 4 * k;
        float b[6] = {sp[0] - sp[3], sp[1] - sp[3], sp[2] - sp[3], sp[0] + sp[3], sp[1] + sp[3], sp[2] + sp[3]};
        int32_t child = ~(int32_t)k;
        int32_t p = J->parent[m + k];
        while (p >= 0) {
            int slot = out->node_child[2 * p] == child ? 0 : 1;
            memcpy(out->node_box + 12 * (size_t)p + 6 * slot, b, sizeof b);
            if (atomic_fetch_add_explicit(&J->seen[p], 1, memory_order_acq_rel) == 0) break;
            const float* c0 = out->node_box + 12 * (size_t)p;
            for (int ax = 0; ax < 3; ++ax) {
                b[ax] = fminf(c0[ax], c0[6 + ax]);
                b[3 + ax] = fmaxf(c0[3 + ax], c0[9 + ax]);
            }
            child = p;
            p = J->parent[p];
        }
    }
}

int orc_bvh_build_mt(const float* xyzr, uint64_t n, int bits, int threads, orc_bvh* out) {
    memset(out, 0, sizeof *out);
    if (n == 0 || n > 0x7fffffffull) return -1;
    int T = threads_for(threads, n);
    build_job* J = calloc(1, sizeof *J);
    J->xyzr = xyzr;
    J->n = n;
    J->m = n - 1;
    J->bits = bits;
    J->out = out;
    out->n = n;
    out->xyzr_in = xyzr;
    par_run(T, bj_bounds, J);
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    out->amax = 0.0f;
    for (int t = 0; t < T; ++t) {
        for (int ax = 0; ax < 3; ++ax) {
            lo[ax] = fminf(lo[ax], J->tlo[t][ax]);
            hi[ax] = fmaxf(hi[ax], J->thi[t][ax]);
        }
        out->amax = fmaxf(out->amax, J->tamax[t]);
    }
    float ext = fmaxf(fmaxf(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
    float cells = bits == 63 ? 2097152.0f : 1024.0f;
    memcpy(J->lo, lo, sizeof lo);
    J->maxq = bits == 63 ? 2097151u : 1023u;
    J->scale = ext > 0.0f ? cells / ext : 0.0f;
    J->key = malloc(n * 8);
    J->key2 = malloc(n * 8);
    J->idx = malloc(n * 4);
    J->idx2 = malloc(n * 4);
    par_run(T, bj_keys, J);
    /* stable LSD radix sort, 8-bit digits: per-thread counts over contiguous ranges, offsets in
     * (digit, thread) order, so equal keys keep their input order */
    J->hist = malloc((size_t)T * sizeof *J->hist);
    int passes = bits == 63 ? 8 : 4;
    for (int ps = 0; ps < passes; ++ps) {
        J->shift = 8 * ps;
        par_run(T, bj_hist, J);
        size_t run = 0;
        for (int d = 0; d < 256; ++d)
            for (int t = 0; t < T; ++t) {
                size_t c = J->hist[t][d];
                J->hist[t][d] = run;
                run += c;
            }
        par_run(T, bj_scatter, J);
        uint64_t* tk = J->key; J->key = J->key2; J->key2 = tk;
        uint32_t* ti = J->idx; J->idx = J->idx2; J->idx2 = ti;
    }
    free(J->hist);
    free(J->key2);
    free(J->idx2);
    out->keys = J->key;
    out->leaf_id = J->idx;
    out->leaf_sph = malloc(n * 16);
    par_run(T, bj_gather, J);
    if (n > 1) {
        uint64_t m = n - 1;
        out->node_box = malloc(m * 12 * sizeof(float));
        out->node_child = malloc(m * 2 * sizeof(int32_t));
        J->parent = malloc((2 * n - 1) * sizeof(int32_t)); /* [0,m) internal, [m, m+n) leaves */
        J->parent[0] = -1;
        par_run(T, bj_nodes, J);
        J->seen = calloc(m, sizeof(atomic_uchar));
        par_run(T, bj_boxes, J);
        free((void*)J->seen);
        free(J->parent);
    }
    free(J);
    return 0;
}

int orc_bvh_build(const float* xyzr, uint64_t n, int bits, orc_bvh* out) {
    return orc_bvh_build_mt(xyzr, n, bits, 1, out);
}

void orc_bvh_free(orc_bvh* b) {
    free(b->node_box);
    free(b->node_child);
    free(b->leaf_id);
    free(b->keys);
    free(b->leaf_sph);
    memset(b, 0, sizeof *b);
}

int orc_bvh_copy_out(const orc_bvh* b, float* node_box, int32_t* node_child, uint32_t* leaf_id, uint64_t* keys) {
    uint64_t m = b->n > 0 ? b->n - 1 : 0;
    if (node_box && m) memcpy(node_box, b->node_box, m * 12 * sizeof(float));
    if (node_child && m) memcpy(node_child, b->node_child, m * 2 * sizeof(int32_t));
    if (leaf_id) memcpy(leaf_id, b->leaf_id, b->n * 4);
    if (keys) memcpy(keys, b->keys, b->n * 8);
    return 0;
}

/* Multi-threaded LBVH renderer: the "mmospray-equivalent" CPU path. Same camera, RNG, FP32
 * intersection, shading and AO estimator as above; 16x16 pixel tiles handed out by an atomic
 * counter to `threads` workers (OSPRay uses TBB tiles, AbstractOSPRayRenderer.cpp:126-132). */
typedef struct mt_job {
    const orc_bvh* b;
    const orc_pkd* k; /* set instead of b: the treelet P-k-d trees */
    const float* xyzr;
    const float* rgba;
    const orc_frame_state* fs;
    const orc_params* p;
    float pad;
    atomic_uint next;
    uint32_t tiles_x, tiles, row0, row1;
    uint32_t* rgba8;
    float* depth;
    float* color;
    float* t_out;
    int32_t* id_out;
    uint16_t* aux;
    atomic_ullong visits, tests;
} mt_job;

static void* mt_worker(void* arg) {
    mt_job* J = arg;
    scene_view s = {J->xyzr, J->k ? J->k->n : J->b->n, J->b, J->pad, 0, 0, J->k};
    for (;;) {
        uint32_t tile = atomic_fetch_add(&J->next, 1u);
        if (tile >= J->tiles) break;
        uint32_t tx = tile % J->tiles_x, ty = tile / J->tiles_x;
        for (uint32_t y = J->row0 + ty * 16; y < J->row0 + ty * 16 + 16 && y < J->row1; ++y)
            for (uint32_t x = tx * 16; x < tx * 16 + 16 && x < J->p->width; ++x) {
                pixel_out po;
                render_pixel(&s, J->xyzr, J->rgba, J->fs, J->p, x, y, &po);
                store_pixel(&po, (size_t)y * J->p->width + x, J->rgba8, J->depth, J->color, J->t_out, J->id_out,
                            J->aux);
            }
    }
    atomic_fetch_add(&J->visits, s.node_visits);
    atomic_fetch_add(&J->tests, s.sphere_tests);
    return NULL;
}

static int render_mt(const orc_bvh* b, const orc_pkd* pk, const float* rgba, const orc_camera* cam,
                     const orc_params* p, int threads, uint32_t* rgba8, float* depth, float* color, float* t_out,
                     int32_t* id_out, uint16_t* aux, uint64_t* counters) {
    orc_frame_state fs;
    orc_frame_state_from_camera(cam, &fs);
    /* particle positions in original order (for shading lookups): the build's input */
    const float* xyzr = pk ? pk->xyzr_in : b->xyzr_in;
    const float amax = pk ? pk->amax : b->amax;
    float cmax = fmaxf(fmaxf(fabsf(cam->position[0]), fabsf(cam->position[1])), fabsf(cam->position[2]));
    mt_job J;
    memset(&J, 0, sizeof J);
    J.b = pk ? NULL : b;
    J.k = pk;
    J.xyzr = xyzr;
    J.rgba = rgba;
    J.fs = &fs;
    J.p = p;
    J.pad = (fmaxf(cmax, amax) + amax) * (1.0f / 131072.0f);
    J.row0 = p->row_end > p->row_begin ? p->row_begin : 0;
    J.row1 = p->row_end > p->row_begin ? p->row_end : p->height;
    J.tiles_x = (p->width + 15) / 16;
    J.tiles = J.tiles_x * ((J.row1 - J.row0 + 15) / 16);
    atomic_init(&J.next, 0u);
    J.rgba8 = rgba8;
    J.depth = depth;
    J.color = color;
    J.t_out = t_out;
    J.id_out = id_out;
    J.aux = aux;
    atomic_init(&J.visits, 0ull);
    atomic_init(&J.tests, 0ull);
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads > 256) threads = 256;
    pthread_t th[256];
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, mt_worker, &J);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    if (counters) {
        counters[0] = atomic_load(&J.visits);
        counters[1] = atomic_load(&J.tests);
    }
    
    return 0;
}

int orc_render_bvh(const orc_bvh* b, const float* rgba, const orc_camera* cam, const orc_params* p, int threads,
                   uint32_t* rgba8, float* depth, float* color, float* t_out, int32_t* id_out, uint16_t* aux,
                   uint64_t* counters) {
    return render_mt(b, NULL, rgba, cam, p, threads, rgba8, depth, color, t_out, id_out, aux, counters);
}

int orc_render_pkd(const orc_pkd* k, const float* rgba, const orc_camera* cam, const orc_params* p, int threads,
                   uint32_t* rgba8, float* depth, float* color, float* t_out, int32_t* id_out, uint16_t* aux,
                   uint64_t* counters) {
    return render_mt(NULL, k, rgba, cam, p, threads, rgba8, depth, color, t_out, id_out, aux, counters);
}

/* ------------------------------------------------- treelet P-k-d build */
/* Treelets: the Morton order cut into runs of M spheres (the reference cuts at spatial midpoints
 * until a run has at most `threshold` particles, TreeletsCreate.h:55-123; Morton runs are cells of
 * the same kind of midpoint subdivision, and a fixed M lets the device address a treelet by
 * shifting). Each treelet holds a left-balanced k-d tree over its spheres in heap order
 * (PKDCreate.h:86-160): the node's split dimension is the widest axis of its cell (arg max of
 * the extents, ties to the lower axis), and its sphere is the one of rank |left subtree| among the
 * subtree's spheres ordered by the coordinate on that axis (ties: the lower Morton position), so
 * the left subtree holds the spheres below it and the right subtree those above; the children's
 * cells are the node's cell cut at its coordinate. The reference reaches the same partition with
 * two heaps (makeHeap / trickle); the order here fixes which of equal coordinates goes where. The
 * root cell is the treelet's sphere bounds. */
static uint32_t ford(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

uint32_t orc_pkd_subtree_size(uint32_t i, uint32_t m) {
    uint64_t first = i, width = 1, cnt = 0;
    while (first < m) {
        uint64_t end = first + width;
        cnt += end <= m ? width : m - first;
        first = 2 * first + 1;
        width *= 2;
    }
    return (uint32_t)cnt;
}

static int widest_axis(const float box[6]) {
    float e0 = box[3] - box[0], e1 = box[4] - box[1], e2 = box[5] - box[2];
    int d = 0;
    float e = e0;
    if (e1 > e) {
        d = 1;
        e = e1;
    }
    if (e2 > e) d = 2;
    return d;
}

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}

/* node i over the local indices elems[0..cnt) of the treelet's spheres sph (Morton order) */
static void pkd_node(const float* sph, const uint32_t* ids, uint64_t* elems, uint32_t cnt, uint32_t i, uint32_t m,
                     const float box[6], float* out_sph, uint32_t* out_id, uint8_t* out_dim) {
    if (cnt == 0) return;
    int d = widest_axis(box);
    for (uint32_t q = 0; q < cnt; ++q) {
        uint32_t j = (uint32_t)elems[q];
        elems[q] = ((uint64_t)ford(sph[4 * (size_t)j + d]) << 32) | j;
    }
    qsort(elems, cnt, sizeof *elems, cmp_u64);
    uint32_t L = 2 * i + 1 < m ? orc_pkd_subtree_size(2 * i + 1, m) : 0;
    uint32_t j = (uint32_t)elems[L];
    memcpy(out_sph + 4 * (size_t)i, sph + 4 * (size_t)j, 16);
    out_id[i] = ids[j];
    out_dim[i] = (uint8_t)d;
    float lb[6], rb[6];
    memcpy(lb, box, sizeof lb);
    memcpy(rb, box, sizeof rb);
    lb[3 + d] = rb[d] = sph[4 * (size_t)j + d];
    for (uint32_t q = 0; q < cnt; ++q) elems[q] &= 0xffffffffull;
    pkd_node(sph, ids, elems, L, 2 * i + 1, m, lb, out_sph, out_id, out_dim);
    pkd_node(sph, ids, elems + L + 1, cnt - L - 1, 2 * i + 2, m, rb, out_sph, out_id, out_dim);
}

typedef struct pkd_job {
    const orc_bvh* b;
    orc_pkd* k;
} pkd_job;

static void pkd_worker(void* a, int t, int T) {
    pkd_job* J = a;
    orc_pkd* k = J->k;
    uint64_t* elems = malloc((size_t)k->M * sizeof *elems);
    PAR_RANGE(k->T, t, T, t0, t1);
    for (uint64_t tl = t0; tl < t1; ++tl) {
        uint64_t base = tl * k->M;
        uint32_t m = (uint32_t)((k->n - base) < k->M ? (k->n - base) : k->M);
        const float* sph = J->b->leaf_sph + 4 * base;
        float box[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
        float rmax = 0.0f;
        for (uint32_t j = 0; j < m; ++j) {
            const float* q = sph + 4 * (size_t)j;
            for (int ax = 0; ax < 3; ++ax) {
                box[ax] = fminf(box[ax], q[ax] - q[3]);
                box[3 + ax] = fmaxf(box[3 + ax], q[ax] + q[3]);
            }
            rmax = fmaxf(rmax, q[3]);
            elems[j] = j;
        }
        memcpy(k->tl_box + 6 * tl, box, sizeof box);
        k->tl_rmax[tl] = rmax;
        pkd_node(sph, J->b->leaf_id + base, elems, m, 0, m, box, k->sph + 4 * base, k->id + base, k->dim + base);
    }
    free(elems);
}

/* The box that no ray reaches (every coordinate 3e38): an absent child of the top tree. */
#define PKD_ABSENT 3.0e38f

int orc_pkd_build(const orc_bvh* b, uint32_t M, int threads, orc_pkd* out) {
    memset(out, 0, sizeof *out);
    if (b->n == 0 || M < 2 || (M & (M - 1))) return -1;
    out->n = b->n;
    out->xyzr_in = b->xyzr_in;
    out->amax = b->amax;
    out->M = M;
    out->T = (uint32_t)((b->n + M - 1) / M);
    uint32_t L = 2;
    while (L < out->T) L *= 2;
    out->top_leaves = L;
    out->sph = malloc(b->n * 16);
    out->id = malloc(b->n * 4);
    out->dim = malloc(b->n);
    out->tl_box = malloc((size_t)out->T * 24);
    out->tl_rmax = malloc((size_t)out->T * 4);
    out->top_box = malloc((size_t)(L - 1) * 48);
    out->top_child = malloc((size_t)(L - 1) * 8);
    pkd_job J = {b, out};
    par_run(threads_for(threads, b->n), pkd_worker, &J);
    /* boxes of all 2L-1 heap nodes of the top tree, bottom up: a leaf is its treelet's box (or
     * absent), a node the union of its present children */
    size_t nn = 2 * (size_t)L - 1;
    float* box = malloc(nn * 24);
    uint8_t* present = malloc(nn);
    for (size_t h = nn; h-- > 0;) {
        float* bx = box + 6 * h;
        if (h >= L - 1) {
            size_t t = h - (L - 1);
            present[h] = t < out->T;
            if (present[h]) memcpy(bx, out->tl_box + 6 * t, 24);
        } else {
            size_t c0 = 2 * h + 1, c1 = 2 * h + 2;
            present[h] = present[c0] || present[c1];
            if (present[c0] && present[c1]) {
                for (int a = 0; a < 3; ++a) {
                    bx[a] = fminf(box[6 * c0 + a], box[6 * c1 + a]);
                    bx[3 + a] = fmaxf(box[6 * c0 + 3 + a], box[6 * c1 + 3 + a]);
                }
            } else if (present[h]) {
                memcpy(bx, box + 6 * (present[c0] ? c0 : c1), 24);
            }
        }
        if (!present[h])
            for (int a = 0; a < 6; ++a) bx[a] = PKD_ABSENT;
    }
    for (size_t h = 0; h < L - 1; ++h)
        for (int c = 0; c < 2; ++c) {
            size_t ch = 2 * h + 1 + (size_t)c;
            memcpy(out->top_box + 12 * h + 6 * c, box + 6 * ch, 24);
            out->top_child[2 * h + c] =
                !present[ch] ? INT32_MIN : ch >= L - 1 ? ~(int32_t)(ch - (L - 1)) : (int32_t)ch;
        }
    free(box);
    free(present);
    return 0;
}

void orc_pkd_free(orc_pkd* k) {
    free(k->sph);
    free(k->id);
    free(k->dim);
    free(k->tl_box);
    free(k->tl_rmax);
    free(k->top_box);
    free(k->top_child);
    memset(k, 0, sizeof *k);
}
