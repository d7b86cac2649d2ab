// rt_math.hpp — FP32 arithmetic of the sphere path, device side.
//
// Every expression keeps the operation order of the reference (MegaMol optix_hpg, glm vector
// semantics) and the library is compiled with -ffp-contract=off, so no multiply-add is fused
// behind our back: hit IDs, t and colours are bit-identical to a plain FP32 CPU evaluation.
// Where a fused multiply-add is wanted (the slab test, whose rounding is absorbed by the box
// padding) it is written explicitly with __builtin_fmaf.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mmhrt {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 mul(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
// glm::dot on vec3: (x*x' + y*y') + z*z'
__device__ __forceinline__ float dot(V3 a, V3 b) {
    float px = a.x * b.x, py = a.y * b.y, pz = a.z * b.z;
    return (px + py) + pz;
}
// glm::normalize: v * (1 / sqrt(dot(v, v))), correctly rounded sqrt and division
__device__ __forceinline__ V3 normalize(V3 v) {
    float inv = 1.0f / __builtin_sqrtf(dot(v, v));
    return V3{v.x * inv, v.y * inv, v.z * inv};
}

// TEA, 16 rounds (plugins/optix_hpg/src/optix/random.h:31-44).
__device__ __forceinline__ uint32_t tea16(uint32_t a, uint32_t b) {
    uint32_t s = 0;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        s += 0x9e3779b9u;
        a += ((b << 4) + 0xa341316cu) ^ (b + s) ^ ((b >> 5) + 0xc8013ea4u);
        b += ((a << 4) + 0xad90777du) ^ (a + s) ^ ((a >> 5) + 0x7e95761eu);
    }
    return a;
}
// LCG (Numerical Recipes constants) -> float in [0,1) with 24 bits (random.h:47-62).
__device__ __forceinline__ float rnd(uint32_t& st) {
    st = st * 1664525u + 1013904223u;
    return (float)(st & 0x00FFFFFFu) / 16777216.0f;
}
// The LCG state after k steps from st, by binary powers of the affine map x -> a x + c
// (F. B. Brown, "Random number generation with arbitrary strides", 1994): the same 32-bit
// state that k calls of rnd() reach.
__device__ __forceinline__ uint32_t lcg_skip(uint32_t st, uint32_t k) {
    uint32_t A = 1u, C = 0u;                 // accumulated map x -> A x + C
    uint32_t a = 1664525u, c = 1013904223u;  // the map raised to the current power of two
    while (k) {
        if (k & 1u) {
            A = A * a;
            C = C * a + c;
        }
        c = c * (a + 1u);
        a = a * a;
        k >>= 1;
    }
    return A * st + C;
}

// Per-frame constants derived on the host from the camera (Renderer.cpp:136-172).
struct FrameConsts {
    float front[3], right[3], up[3], center[3];
    float th, rw, nearp;
    float depth_a, depth_c;
    float pad;          // slab padding (absolute, world units)
    float ao_distance;
    float bg[4];
    uint32_t width, height, spp, frame_idx;
    int32_t mode, accumulate, want_aux, max_bounces;
    uint32_t win_x0, win_y0, win_w, win_h;  // rendered pixel window (the whole frame by default)
};

// Primary ray of one sample: raygen.cu:241-245 + camera.h:11-16.
__device__ __forceinline__ void primary_ray(const FrameConsts& f, uint32_t px, uint32_t py, float jx, float jy,
                                            V3& o, V3& d) {
    float u = -f.rw + ((f.rw + f.rw) * ((float)px + jx)) / (float)f.width;
    float v = -(f.th + ((-f.th - f.th) * ((float)py + jy)) / (float)f.height);
    V3 dir;
    dir.x = (f.nearp * f.front[0] + u * f.right[0]) + v * f.up[0];
    dir.y = (f.nearp * f.front[1] + u * f.right[1]) + v * f.up[1];
    dir.z = (f.nearp * f.front[2] + u * f.right[2]) + v * f.up[2];
    o = V3{f.center[0], f.center[1], f.center[2]};
    d = normalize(dir);
}

// Ray-sphere test of plugins/optix_hpg/cuda_resources/sphere.cuh:14-41: returns t of the
// reported intersection or a negative value for "no report". (tmin, tmax) is the ray interval.
__device__ __forceinline__ float sphere_t(V3 o, V3 d, float4 s, float tmin, float tmax) {
    V3 oc = V3{o.x - s.x, o.y - s.y, o.z - s.z};
    float rr = s.w * s.w;
    float b = dot(neg(oc), d);
    V3 tmp = V3{oc.x + b * d.x, oc.y + b * d.y, oc.z + b * d.z};
    float delta = rr - dot(tmp, tmp);
    if (delta < 0.0f) return -1.0f;
    float c = dot(oc, oc) - rr;
    float q = b + __builtin_copysignf(__builtin_sqrtf(delta), b);
    float t = fminf(c / q, q);
    if (t < 0.0f) return -1.0f;
    return (t > tmin && t < tmax) ? t : -1.0f;
}

// sin and cos on [-pi/4, pi/4] by their Taylor series (terms through x^11 / x^12: error below
// 1e-9, under half an ulp of the result), evaluated with plain multiplies and adds in a fixed
// order so that CPU and GPU produce the same bits.
__device__ __forceinline__ void sincos_q(float x, float& s, float& c) {
    float x2 = x * x;
    s = x * (1.0f + x2 * (-1.0f / 6.0f + x2 * (1.0f / 120.0f + x2 * (-1.0f / 5040.0f + x2 * (1.0f / 362880.0f +
                                                                                        x2 * (-1.0f / 39916800.0f))))));
    c = 1.0f + x2 * (-0.5f + x2 * (1.0f / 24.0f + x2 * (-1.0f / 720.0f + x2 * (1.0f / 40320.0f +
                                                                              x2 * (-1.0f / 3628800.0f +
                                                                                    x2 * (1.0f / 479001600.0f))))));
}

// AO sample direction, cosine-weighted about the unit normal n. Sample s of spp takes the
// stratum (s % g, s / g) of a g x g grid over the unit square (g = floor(sqrt(spp)); samples past
// g*g are not stratified), jittered by two draws of its own stream; the square maps to the disk
// by the concentric map (Shirley & Chiu 1997), the disk to the hemisphere by Malley's method, and
// the hemisphere to world space through the orthonormal basis of Duff et al. (2017).
__device__ __forceinline__ V3 ao_direction(V3 n, uint32_t& st, uint32_t s, uint32_t spp) {
    uint32_t g = (uint32_t)__builtin_sqrtf((float)spp);
    while (g * g > spp) --g;
    while ((g + 1) * (g + 1) <= spp) ++g;
    float r1 = rnd(st);
    float r2 = rnd(st);
    float u1 = r1, u2 = r2;
    if (s < g * g) {
        float gf = (float)g;
        u1 = ((float)(s % g) + r1) / gf;
        u2 = ((float)(s / g) + r2) / gf;
    }
    float a = 2.0f * u1 - 1.0f, b = 2.0f * u2 - 1.0f;
    float dx = 0.0f, dy = 0.0f;
    if (a != 0.0f || b != 0.0f) {
        const float q = 0.785398163397448f;  // pi / 4
        float r, sn, cs;
        if (fabsf(a) > fabsf(b)) {
            r = a;
            sincos_q(q * (b / a), sn, cs);
            dx = r * cs;
            dy = r * sn;
        } else {
            r = b;
            sincos_q(q * (a / b), sn, cs);  // phi = pi/2 - q * a/b: cos phi = sin, sin phi = cos
            dx = r * sn;
            dy = r * cs;
        }
    }
    float dz = __builtin_sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy));
    float sg = __builtin_copysignf(1.0f, n.z);
    float ia = -1.0f / (sg + n.z);
    float bb = n.x * n.y * ia;
    V3 t = V3{1.0f + sg * n.x * n.x * ia, sg * bb, -sg * n.x};
    V3 bt = V3{bb, sg + n.y * n.y * ia, -n.y};
    V3 d = V3{(dx * t.x + dy * bt.x) + dz * n.x, (dx * t.y + dy * bt.y) + dz * n.y, (dx * t.z + dy * bt.z) + dz * n.z};
    float l2 = dot(d, d);
    if (!(l2 > 1e-12f)) return n;
    float inv = 1.0f / __builtin_sqrtf(l2);
    return V3{d.x * inv, d.y * inv, d.z * inv};
}

__device__ __forceinline__ uint32_t pack_rgba8(float4 c) {
    float v[4] = {c.x, c.y, c.z, c.w};
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float q = fminf(fmaxf(v[i], 0.0f), 1.0f);
        r |= ((uint32_t)(q * 255.0f + 0.5f)) << (8 * i);
    }
    return r;
}

}  // namespace mmhrt
