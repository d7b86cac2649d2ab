// ingest.hip — particle-list conversion on the device.
//
// Replaces the single-threaded host loop of SphereGeometry::assertData
// (plugins/optix_hpg/src/optix/SphereGeometry.cpp:118-136, four to eight virtual Accessor::Get_f
// calls per particle) and, for intensity lists, AddParticleColors::manipulateData
// (plugins/datatools/src/AddParticleColors.cpp:74-94). The raw, possibly interleaved list bytes
// (an MMPLD frame keeps vertex and colour in one record, MMPLDDataSource.cpp:197-198) are
// copied to the device once and decoded here, one thread per particle.
//
// Value semantics follow the Accessor classes (SimpleSphericalParticles.h:85-175): every field
// is static_cast<float>; 8- and 16-bit colours are NOT normalised; the alpha of UINT8_RGB is
// 255 and that of FLOAT_RGB is 1; lists without per-particle colour use global_rgba / 255.
#include <algorithm>

#include "ctx.hpp"

namespace mmhrt {
namespace {

template <bool ALIGNED>
__device__ __forceinline__ float ld_f32(const uint8_t* p) {
    if constexpr (ALIGNED) {
        return *reinterpret_cast<const float*>(p);
    } else {
        uint32_t u = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
        return __uint_as_float(u);
    }
}
__device__ __forceinline__ double ld_f64(const uint8_t* p) {
    uint64_t u = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) u |= (uint64_t)p[i] << (8 * i);
    return __longlong_as_double((long long)u);
}
__device__ __forceinline__ uint16_t ld_u16(const uint8_t* p) { return (uint16_t)(p[0] | (p[1] << 8)); }

// AddParticleColors.cpp:19-36 (sample_tf, lerp) and :91-94 (val, modf split).
__device__ __forceinline__ float4 tf_lookup(const float4* __restrict__ tf, uint32_t n, float vmin, float fac,
                                            float value) {
    float val = (value - vmin) * fac * (float)n;
    float main_part = truncf(val);
    float rest = val - main_part;
    // keep the float->int conversion defined: NaN and large negatives map below 0 (zero colour),
    // anything at or beyond the table end maps to the last entry
    if (main_part > (float)n) main_part = (float)n;
    if (!(main_part >= -1.0f)) main_part = -1.0f;
    int base = (int)main_part;
    if (base < 0 || n == 0) return make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t last = n - 1;
    if ((uint32_t)base >= last) return tf[last];
    float4 a = tf[base], b = tf[base + 1];
    float inv = 1.0f - rest;
    return make_float4(a.x * inv + b.x * rest, a.y * inv + b.y * rest, a.z * inv + b.z * rest,
                       a.w * inv + b.w * rest);
}

template <bool ALIGNED>
__global__ __launch_bounds__(256) void k_convert(ListDesc d, const float4* __restrict__ tf, uint32_t tf_n,
                                                 float tf_vmin, float tf_fac, float4* __restrict__ sph,
                                                 float4* __restrict__ col) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.count) return;
    const uint8_t* v = d.vert + i * d.vstride;
    float4 s;
    switch (d.vtype) {
    case 1:
        s = make_float4(ld_f32<ALIGNED>(v), ld_f32<ALIGNED>(v + 4), ld_f32<ALIGNED>(v + 8), d.gradius);
        break;
    case 2:
        s = make_float4(ld_f32<ALIGNED>(v), ld_f32<ALIGNED>(v + 4), ld_f32<ALIGNED>(v + 8), ld_f32<ALIGNED>(v + 12));
        break;
    case 3:
        s = make_float4((float)ld_u16(v), (float)ld_u16(v + 2), (float)ld_u16(v + 4), d.gradius);
        break;
    default:  // 4: DOUBLE_XYZ
        s = make_float4((float)ld_f64(v), (float)ld_f64(v + 8), (float)ld_f64(v + 16), d.gradius);
        break;
    }
    float4 c = make_float4(d.gcol[0], d.gcol[1], d.gcol[2], d.gcol[3]);
    const uint8_t* cp = d.col + i * d.cstride;
    switch (d.ctype) {
    case 1: c = make_float4((float)cp[0], (float)cp[1], (float)cp[2], 255.0f); break;
    case 2: c = make_float4((float)cp[0], (float)cp[1], (float)cp[2], (float)cp[3]); break;
    case 3: c = make_float4(ld_f32<false>(cp), ld_f32<false>(cp + 4), ld_f32<false>(cp + 8), 1.0f); break;
    case 4:
        c = make_float4(ld_f32<false>(cp), ld_f32<false>(cp + 4), ld_f32<false>(cp + 8), ld_f32<false>(cp + 12));
        break;
    case 5:
        if (d.use_tf) c = tf_lookup(tf, tf_n, tf_vmin, tf_fac, ld_f32<false>(cp));
        break;
    case 6:
        c = make_float4((float)ld_u16(cp), (float)ld_u16(cp + 2), (float)ld_u16(cp + 4), (float)ld_u16(cp + 6));
        break;
    case 7:
        if (d.use_tf) c = tf_lookup(tf, tf_n, tf_vmin, tf_fac, (float)ld_f64(cp));
        break;
    default: break;
    }
    sph[d.out_offset + i] = s;
    col[d.out_offset + i] = c;
}

__global__ __launch_bounds__(256) void k_copy_packed(const float4* __restrict__ xyzr, const float4* __restrict__ rgba,
                                                     uint64_t n, float4 gcol, float4* __restrict__ sph,
                                                     float4* __restrict__ col) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sph[i] = xyzr[i];
    col[i] = rgba ? rgba[i] : gcol;
}

__global__ __launch_bounds__(256) void k_fill_f4(float4* __restrict__ p, float4 v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace

void launch_fill_f4(hipStream_t s, float4* p, float4 v, uint64_t n) {
    if (n == 0) return;
    unsigned g = (unsigned)std::min<uint64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill_f4, dim3(g), dim3(256), 0, s, p, v, n);
}

int launch_convert_list(mmhrt_ctx_s* c, const ListDesc& d, const float* tf_dev, uint32_t tf_n, float tf_vmin,
                        float tf_fac) {
    if (d.count == 0) return 0;
    unsigned g = (unsigned)((d.count + 255) / 256);
    bool aligned = ((uintptr_t)d.vert % 4 == 0) && (d.vstride % 4 == 0);
    if (aligned)
        hipLaunchKernelGGL(k_convert<true>, dim3(g), dim3(256), 0, c->stream, d, (const float4*)tf_dev, tf_n, tf_vmin,
                           tf_fac, c->sph.p, c->col.p);
    else
        hipLaunchKernelGGL(k_convert<false>, dim3(g), dim3(256), 0, c->stream, d, (const float4*)tf_dev, tf_n, tf_vmin,
                           tf_fac, c->sph.p, c->col.p);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(c, e, "convert launch");
}

int launch_copy_packed(mmhrt_ctx_s* c, const float4* xyzr, const float4* rgba, uint64_t n, float4 gcol) {
    if (n == 0) return 0;
    unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_copy_packed, dim3(g), dim3(256), 0, c->stream, xyzr, rgba, n, gcol, c->sph.p, c->col.p);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(c, e, "copy launch");
}

}  // namespace mmhrt
