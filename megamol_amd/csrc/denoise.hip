// denoise.hip — learned denoiser for 1-spp AO frames: a 4-level U-Net (3x3 convolutions,
// 32/48/64/96 channels, ReLU, skip concatenation, 2x2 average pooling, nearest upsampling),
// trained and run with hand-written BF16 MFMA kernels that accumulate in FP32.
//
// No counterpart in the reference (MegaMol's optix_hpg path has no denoiser); the renderer
// itself produces the training pairs (1-spp features against high-spp frames).
//
// Tensors are NHWC bf16 with channel counts padded to multiples of 16. A 3x3 convolution is an
// implicit GEMM: M = pixels, N = output channels, K = 9 taps x input channels. One wavefront
// computes 64 consecutive pixels of one row against all output channels with
// v_mfma_f32_16x16x16_bf16 (4 pixel blocks x N/16 channel blocks of 16x16 accumulators).
// The input of a decoder convolution is the concatenation [skip (full res), coarser (half res,
// nearest-upsampled)]; the kernel reads both tensors directly, so neither the upsampled nor the
// concatenated tensor is ever stored.
//
// Lane layout of v_mfma_f32_16x16x16_bf16 (64 lanes): A holds row l%16, k = 4*(l/16)..+3;
// B holds column l%16, k = 4*(l/16)..+3; the 4 results of a lane are rows 4*(l/16)..+3 of
// column l%16. v_mfma_f32_16x16x32_bf16 is the same with 8 k values per lane (k = 8*(l/16)..+7).
// The forward pass uses the latter over K = 9 taps x CI input channels taken 32 at a time (a
// lane's 8 values are 8 consecutive channels of one tap, one 16-byte load). The LDS-staged and
// DMA forward kernels take the weights as A and 16 pixels as B, so that a lane's 4 results are 4
// adjacent channels of one pixel (one store).
#include <math.h>

#include <algorithm>
#include <random>
#include <vector>

#include "denoise.hpp"

namespace mmhrt {
namespace dn {
namespace {

typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even: one v_cvt_pk_bf16_f32 (gfx950), which rounds as the mode register says
// (nearest even, denormals kept) and keeps NaN a NaN; a shift-and-add rounding costs 6 VALU
// instructions per value, which dominated the convolution epilogues
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // loaded and stored as one value (never a memcpy)
typedef float f32x2_t __attribute__((ext_vector_type(2)));
// two values in one conversion: lo in bits 0-15, hi in bits 16-31
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t){lo, hi}, bf16x2_t));
}

// Epilogue of 16x16 MFMA output blocks: lane (q, lr) holds channel co = 16 nb + lr of pixels
// x0 + 16 pb + 4q + e, e < 4; bias added, optional ReLU, bf16 or fp32 out. The values are final
// before the first store, so no store waits for anything. (Transposing within lane quads to
// store 4 channels at once measured slower: the shuffles and registers cost more than the
// narrow stores.)
template <int PB, int NB>
__device__ __forceinline__ void store_blocks(const ConvArgs& a, const f32x4 (&acc)[PB][NB], const float (&bias)[NB],
                                             int nb0, int NBT, int b, int y, int x0, int q, int lr) {
    const size_t rowo = ((size_t)b * a.H + y) * a.W;
    const bool f32out = a.Yf != nullptr;
    const int yc = a.YfC ? a.YfC : a.CO;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        if (nb0 + j >= NBT) break;
        const int co = (nb0 + j) * 16 + lr;
#pragma unroll
        for (int pb = 0; pb < PB; ++pb) {
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v[e] = acc[pb][j][e] + bias[j];
                if (a.relu) v[e] = fmaxf(v[e], 0.0f);
            }
            if (f32out) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int x = x0 + pb * 16 + q * 4 + e;
                    if (x < a.W && co < yc) a.Yf[(rowo + x) * yc + co] = v[e];
                }
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int x = x0 + pb * 16 + q * 4 + e;
                    if (x < a.W) a.Y[(rowo + x) * a.CO + co] = f2bf(v[e]);
                }
            }
        }
    }
}

// Epilogue of one output row for blocks computed as weights x activations (the weight fragment
// as the MFMA's A operand): lane (q, lr) holds channels 16 nb + 4q + e, e < 4, of pixel
// x0 + 16 pb + lr, so its four values are adjacent in NHWC and leave in one 8-byte (bf16) or
// 16-byte (fp32) store: PB * NB stores per lane (pixels as the A operand would give one channel
// of 4 pixels per lane and 4x the stores). The accumulators started from the bias (the first
// MFMA's C operand), so nothing is added here. ReLU on bf16 output is one packed 16-bit integer
// max per channel pair after rounding: a bf16 with the sign bit set is a negative int16, and
// rounding keeps the sign, so max(round(v), 0) = round(max(v, 0)). Channel counts written are
// multiples of 4 (16 NB, or 4 for the fp32 output layer). Every store is a buffer store over the
// row (bytes [0, W * channels * size)), and a pixel or channel outside it gets an offset past the
// end, which the hardware drops. No store sits in a branch, so every path issues exactly PB * NB
// of them, and a wait for loads issued before them can count them instead of waiting for the
// stores too. row_ok false drops the whole row.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t v) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, v), (s16x2){0, 0}));
}
template <int PB, int NB, bool F32, bool RELU>
__device__ __forceinline__ void store_row_t(const ConvArgs& a, const f32x4 (&acc)[PB][NB], int b, int y, bool row_ok,
                                            int x0, int q, int lr, int c0) {
    const int yc = F32 ? (a.YfC ? a.YfC : a.CO) : a.CO;
    const size_t rowo = ((size_t)b * a.H + (row_ok ? y : 0)) * a.W;
    const uint64_t ba = F32 ? (uint64_t)(a.Yf + rowo * yc) : (uint64_t)(a.Y + rowo * a.CO);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)ba);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(ba >> 32));
    void* base = (void*)(((uint64_t)hi << 32) | lo);
    constexpr int S = F32 ? 4 : 2;  // bytes per value
    const int row_bytes = __builtin_amdgcn_readfirstlane(row_ok ? a.W * yc * S : 0);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(base, 0, row_bytes, 0x00020000);
    constexpr int kOut = 0x40000000;  // an offset past any row
#pragma unroll
    for (int pb = 0; pb < PB; ++pb) {
        const int x = x0 + pb * 16 + lr;
        const int px = x < a.W ? x * yc * S : kOut;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int co = c0 + j * 16 + q * 4;
            const int off = co < yc ? px + co * S : kOut;
            const f32x4 v = acc[pb][j];
            if constexpr (F32) {
                float w[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) w[e] = RELU ? fmaxf(v[e], 0.0f) : v[e];
                __builtin_amdgcn_raw_buffer_store_b128(
                    (u32x4){__float_as_uint(w[0]), __float_as_uint(w[1]), __float_as_uint(w[2]), __float_as_uint(w[3])},
                    rs, off, 0, 0);
            } else {
                uint32_t p0 = f2bf2(v[0], v[1]), p1 = f2bf2(v[2], v[3]);
                if constexpr (RELU) {
                    p0 = relu_bf16x2(p0);
                    p1 = relu_bf16x2(p1);
                }
                __builtin_amdgcn_raw_buffer_store_b64((u32x2){p0, p1}, rs, off, 0, 0);
            }
        }
    }
}
template <int PB, int NB, bool F32>
__device__ __forceinline__ void store_row_t(const ConvArgs& a, const f32x4 (&acc)[PB][NB], int b, int y, bool row_ok,
                                            int x0, int q, int lr, int c0 = 0) {
    if (a.relu)
        store_row_t<PB, NB, F32, true>(a, acc, b, y, row_ok, x0, q, lr, c0);
    else
        store_row_t<PB, NB, F32, false>(a, acc, b, y, row_ok, x0, q, lr, c0);
}

// bias of the channels c0 + 16 nb + 4q + e a lane holds in the layout above
template <int NB>
__device__ __forceinline__ void load_bias_t(const ConvArgs& a, int q, f32x4 (&bias)[NB], int c0 = 0) {
#pragma unroll
    for (int j = 0; j < NB; ++j) bias[j] = *reinterpret_cast<const f32x4*>(a.bias + c0 + j * 16 + q * 4);
}

union Frag {
    bf16x4 v;
    uint2 u;
    uint16_t h[4];
};
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
union Frag8 {
    bf16x8 v;
    uint4 u;
};

// ------------------------------------------------------------------ forward convolution
// X1: [B,H,W,C1], X2: [B,H/2,W/2,C2] (nearest upsampled) or null; Y: [B,H,W,CO] (bf16) or
// Yf (fp32). Wf: weights in fragment order [tap][ci/16][co/16][co%16][ci%16] (bf16).
template <int NB>
__global__ __launch_bounds__(256) void k_conv_fwd(ConvArgs a) {
    const int lane = __lane_id();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int runs_per_row = (a.W + 63) / 64;
    const int row = wave / runs_per_row;  // b*H + y
    if (row >= a.B * a.H) return;
    const int x0 = (wave % runs_per_row) * 64;
    const int b = row / a.H, y = row % a.H;
    const int nb0 = blockIdx.y * NB;  // output-channel block offset (grid.y splits CO)
    const int CC = (a.C1 + a.C2) / 16;
    const int CC1 = a.C1 / 16;
    const int NBT = a.CO / 16;
    f32x4 acc[4][NB];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int kq = lane >> 4;  // k quad
    const int lr = lane & 15;  // row (A) / column (B)
    for (int t = 0; t < 9; ++t) {
        const int yy = y + t / 3 - 1;
        const bool yok = yy >= 0 && yy < a.H;
        for (int cc = 0; cc < CC; ++cc) {
            Frag fa[4];
#pragma unroll
            for (int pb = 0; pb < 4; ++pb) {
                int xx = x0 + pb * 16 + lr + t % 3 - 1;
                fa[pb].u = make_uint2(0u, 0u);
                if (yok && xx >= 0 && xx < a.W) {
                    const uint16_t* src;
                    if (cc < CC1)
                        src = a.X1 + (((size_t)b * a.H + yy) * a.W + xx) * a.C1 + cc * 16 + kq * 4;
                    else
                        src = a.X2 + (((size_t)b * (a.H >> 1) + (yy >> 1)) * (a.W >> 1) + (xx >> 1)) * a.C2 +
                              (cc - CC1) * 16 + kq * 4;
                    fa[pb].u = *reinterpret_cast<const uint2*>(src);
                }
            }
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                int nb = nb0 + j;
                if (nb >= NBT) break;
                Frag fb;
                fb.u = *reinterpret_cast<const uint2*>(a.Wt + ((((size_t)t * CC + cc) * NBT + nb) * 16 + lr) * 16 + kq * 4);
#pragma unroll
                for (int pb = 0; pb < 4; ++pb)
                    acc[pb][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[pb].v, fb.v, acc[pb][j], 0, 0, 0);
            }
        }
    }
    // epilogue: bias, ReLU, store
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        int nb = nb0 + j;
        if (nb >= NBT) break;
        int co = nb * 16 + lr;
        float bias = a.bias[co];
#pragma unroll
        for (int pb = 0; pb < 4; ++pb)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                int x = x0 + pb * 16 + kq * 4 + e;
                if (x >= a.W) continue;
                float v = acc[pb][j][e] + bias;
                if (a.relu) v = fmaxf(v, 0.0f);
                size_t o = (((size_t)b * a.H + y) * a.W + x) * a.CO + co;
                if (a.Yf) {
                    const int yc = a.YfC ? a.YfC : a.CO;
                    if (co < yc) a.Yf[(((size_t)b * a.H + y) * a.W + x) * yc + co] = v;
                }
                else
                    a.Y[o] = f2bf(v);
            }
    }
}

// The same convolution with v_mfma_f32_16x16x32_bf16 over 32-wide chunks of k = tap*CI + ci.
// A wavefront computes PB blocks of 16 pixels (PB = 4, 2 or 1: the coarse levels use fewer
// pixels and fewer output-channel blocks per wavefront so that the launch still fills the chip).
template <int NB, int PB>
__global__ __launch_bounds__(256) void k_conv_fwd32(ConvArgs a) {
    const int lane = __lane_id();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int runs_per_row = (a.W + 16 * PB - 1) / (16 * PB);
    const int row = wave / runs_per_row;  // b*H + y
    if (row >= a.B * a.H) return;
    const int x0 = (wave % runs_per_row) * 16 * PB;
    const int b = row / a.H, y = row % a.H;
    const int nb0 = blockIdx.y * NB;
    const int CI = a.C1 + a.C2;
    const int K9 = 9 * CI;
    const int nkc = (K9 + 31) / 32;
    const int NBT = a.CO / 16;
    f32x4 acc[PB][NB];
#pragma unroll
    for (int i = 0; i < PB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int q = lane >> 4;   // which 8 of the chunk's 32 k values
    const int lr = lane & 15;  // row (A) / column (B)
    float bias[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) bias[j] = a.bias[(nb0 + j < NBT ? nb0 + j : 0) * 16 + lr];
    // fragments of chunk kc: A loads at clamped, always valid addresses, zeroed where the tap
    // falls outside the image or past CI (a guarded load would be a branch and a wait per chunk)
    auto load_a = [&](int kc, Frag8 (&fa)[PB]) {
        const int k0 = kc * 32 + q * 8;
        const bool kok = k0 < K9;
        const int kk = kok ? k0 : 0;
        const int t = kk / CI;
        const int ci0 = kk - t * CI;
        const int yy = y + t / 3 - 1;
        const int dx = t % 3 - 1;
        const bool yok = kok && yy >= 0 && yy < a.H;
        const int yc = yy < 0 ? 0 : (yy >= a.H ? a.H - 1 : yy);
        const bool two = ci0 >= a.C1;
#pragma unroll
        for (int pb = 0; pb < PB; ++pb) {
            const int xx = x0 + pb * 16 + lr + dx;
            const bool ok = yok && xx >= 0 && xx < a.W;
            const int xc = xx < 0 ? 0 : (xx >= a.W ? a.W - 1 : xx);
            const uint16_t* src =
                two ? a.X2 + (((size_t)b * (a.H >> 1) + (yc >> 1)) * (a.W >> 1) + (xc >> 1)) * a.C2 + (ci0 - a.C1)
                    : a.X1 + (((size_t)b * a.H + yc) * a.W + xc) * a.C1 + ci0;
            const uint4 v = *reinterpret_cast<const uint4*>(src);
            fa[pb].u = ok ? v : make_uint4(0u, 0u, 0u, 0u);
        }
    };
    auto load_b = [&](int kc, Frag8 (&fb)[NB]) {
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int nb = nb0 + j < NBT ? nb0 + j : NBT - 1;
            fb[j].u = *reinterpret_cast<const uint4*>(a.Wt32 + (((size_t)kc * NBT + nb) * 64 + lane) * 8);
        }
    };
    // software pipeline: chunk kc + 1's fragments load while chunk kc's products run
    Frag8 fa[PB], fb[NB];
    load_a(0, fa);
    load_b(0, fb);
    for (int kc = 0; kc < nkc; ++kc) {
        Frag8 ca[PB], cb[NB];
#pragma unroll
        for (int pb = 0; pb < PB; ++pb) ca[pb] = fa[pb];
#pragma unroll
        for (int j = 0; j < NB; ++j) cb[j] = fb[j];
        if (kc + 1 < nkc) {
            load_a(kc + 1, fa);
            load_b(kc + 1, fb);
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (nb0 + j >= NBT) break;
#pragma unroll
            for (int pb = 0; pb < PB; ++pb)
                acc[pb][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ca[pb].v, cb[j].v, acc[pb][j], 0, 0, 0);
        }
    }
    store_blocks<PB, NB>(a, acc, bias, nb0, NBT, b, y, x0, q, lr);
}

// The same convolution staged through LDS, at every resolution. A workgroup (4 wavefronts)
// computes 4 rows x 64 pixels x all output channels. For each 32-channel slice of the input it
// stages rows y0-1..y0+4, pixels x0-1..x0+64 in LDS (the decoder's [skip, upsampled]
// concatenation is resolved here, once per element), then the wavefronts run the 9 taps out of
// LDS: up to 48 output channels each takes 16 pixels of all 4 rows, wider layers one row each.
// Pixels are padded to 40 channels in LDS so that the 16 lanes of an activation-fragment read
// land on distinct banks. Weights: [tap][slice][co/16][lane][8], zero past CI.
constexpr int LT_PX = 66, LT_CP = 40;
// output rows of a tile: 8 up to 32 output channels (a wavefront's 16 pixels x 8 rows x NB blocks of
// accumulators fit its registers beside the staged slice), 4 above
__host__ __device__ constexpr int lds_tile_rows(int nb) { return nb <= 2 ? 8 : 4; }

template <int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_conv_fwd_lds(ConvArgs a) {
    // up to 3 output-channel blocks the slice's weights (9 * NB KB) sit in LDS beside the tile,
    // so that the tap loop reads LDS only; wider layers read them from global memory, one tap ahead
    constexpr bool WLDS = NB <= 3;
    constexpr int ROWS = lds_tile_rows(NB);  // output rows of a tile
    constexpr int TR = ROWS + 2;              // staged rows
    constexpr int NT = TR * LT_PX * 4;        // 16-byte pieces of a slice
    constexpr int NI = (NT + 255) / 256;
    constexpr int NWT = 9 * NB * 64;         // 16-byte weight fragments of a slice
    constexpr int NW = WLDS ? (NWT + 255) / 256 : 0;
    // both arrays are padded to whole rounds of 256 pieces, so that every staging store is
    // unconditional (a guarded store lets the compiler sink its load into the branch). With the
    // weights in LDS the tile is unpadded (LT_PX-pixel rows of 4 16-byte groups) and a pixel's
    // groups are XOR-swizzled (slot 4P + (g ^ ((P >> 1) & 3)) for tile pixel P), so that the 8
    // consecutive pixels of a fragment read phase cover all 32 banks; wider layers keep pixels
    // padded to LT_CP channels.
    __shared__ __attribute__((aligned(16))) uint16_t
        tile[WLDS ? NI * 256 * 8 : (NI * 256 / (LT_PX * 4) + 1) * LT_PX * LT_CP];
    __shared__ uint4 s_w[WLDS ? NW * 256 : 1];
    const int CI = a.C1 + a.C2;
    const int nslice = (CI + 31) / 32;
    const int runs = (a.W + 63) / 64;
    const int rblocks = (a.H + ROWS - 1) / ROWS;
    int blk = blockIdx.x;
    const int run = blk % runs;
    blk /= runs;
    const int rb = blk % rblocks;
    const int b = blk / rblocks;
    const int x0 = run * 64, y0 = rb * ROWS;
    // output channels: group blockIdx.y of NB 16-channel blocks out of NBT (a layer wider than the
    // kernel's NB runs as several groups, each with its slice of the weights in LDS)
    const int NBT = a.CO / 16;
    const int c0 = (int)blockIdx.y * NB * 16;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: the store row
    const int lane = __lane_id();
    const int q = lane >> 4, lr = lane & 15;
    // acc[row][nb] (WLDS: the wavefront's 16 pixels of each row) or acc[pb][nb] (one row); they start
    // from the bias (channels 16 nb + 4q + e of the lane's pixel)
    constexpr int NA = WLDS ? ROWS : 4;
    f32x4 acc[NA][NB];
    {
        f32x4 bias[NB];
        load_bias_t<NB>(a, q, bias, c0);
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) acc[i][j] = bias[j];
    }
    const uint4* wl4 = reinterpret_cast<const uint4*>(a.Wl);
    // A piece's place in the tile (row, pixel, 8-channel group) is fixed per thread; its source is
    // X1 (channels below C1) or the half-resolution X2 (nearest upsampled), read with buffer loads:
    // a piece outside the image or past CI gets an offset past its tensor and reads zeros, so no
    // load is guarded and no mask is needed.
    int where[NI];  // tile row | pixel << 8 | group << 16, or -1 past the tile
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        const int i = threadIdx.x + k * 256;
        where[k] = i < NT ? (((i >> 2) / LT_PX) | (((i >> 2) % LT_PX) << 8) | ((i & 3) << 16)) : -1;
    }
    const __amdgpu_buffer_rsrc_t r1 =
        __builtin_amdgcn_make_buffer_rsrc((void*)a.X1, 0, a.B * a.H * a.W * a.C1 * 2, 0x00020000);
    const __amdgpu_buffer_rsrc_t r2 = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a.X2 ? a.X2 : a.X1), 0, a.X2 ? a.B * (a.H >> 1) * (a.W >> 1) * a.C2 * 2 : 0, 0x00020000);
    constexpr int kOut = (int)0x80000000;
    // stage a slice into registers: 6 rows x 66 pixels x 4 groups of 8 channels and, WLDS, its
    // weights; every load is issued before any is stored
    uint4 v[NI];
    u32x4 wv[NW > 0 ? NW : 1];
    auto off1 = [&](int yy, int xx, int ci0) { return (((b * a.H + yy) * a.W + xx) * a.C1 + ci0) * 2; };
    auto off2 = [&](int yy, int xx, int ci0) {
        return (((b * (a.H >> 1) + (yy >> 1)) * (a.W >> 1) + (xx >> 1)) * a.C2 + ci0 - a.C1) * 2;
    };
    auto stage = [&](int sl) {
        const bool from1 = sl * 32 < a.C1, from2 = sl * 32 + 32 > a.C1 && a.C2 > 0;  // uniform
        if (!(from1 && from2)) {  // the whole slice from one tensor: one load per piece
            const __amdgpu_buffer_rsrc_t rs = from1 ? r1 : r2;
#pragma unroll
            for (int k = 0; k < NI; ++k) {
                const int w = where[k];
                const int yy = y0 - 1 + (w & 0xff), xx = x0 - 1 + ((w >> 8) & 0xff);
                const int ci0 = sl * 32 + ((w >> 16) & 3) * 8;
                const bool in = w >= 0 && (unsigned)yy < (unsigned)a.H && (unsigned)xx < (unsigned)a.W && ci0 < CI;
                const int o = in ? (from1 ? off1(yy, xx, ci0) : off2(yy, xx, ci0)) : kOut;
                v[k] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, o, 0, 0));
            }
        } else {  // a slice across C1 (C1 not a multiple of 32): both loads issued, then combined
            uint4 v2[NI];
#pragma unroll
            for (int k = 0; k < NI; ++k) {
                const int w = where[k];
                const int yy = y0 - 1 + (w & 0xff), xx = x0 - 1 + ((w >> 8) & 0xff);
                const int ci0 = sl * 32 + ((w >> 16) & 3) * 8;
                const bool in = w >= 0 && (unsigned)yy < (unsigned)a.H && (unsigned)xx < (unsigned)a.W;
                v[k] = __builtin_bit_cast(
                    uint4, __builtin_amdgcn_raw_buffer_load_b128(r1, in && ci0 < a.C1 ? off1(yy, xx, ci0) : kOut, 0, 0));
                v2[k] = __builtin_bit_cast(
                    uint4, __builtin_amdgcn_raw_buffer_load_b128(
                               r2, in && ci0 >= a.C1 && ci0 < CI ? off2(yy, xx, ci0) : kOut, 0, 0));
            }
#pragma unroll
            for (int k = 0; k < NI; ++k)
                v[k] = make_uint4(v[k].x | v2[k].x, v[k].y | v2[k].y, v[k].z | v2[k].z, v[k].w | v2[k].w);
        }
        const u32x4* w4 = reinterpret_cast<const u32x4*>(a.Wl);
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            const int i = threadIdx.x + k * 256;
            const int ii = i < NWT ? i : 0;
            const int t = ii / (NB * 64), r = ii - t * (NB * 64);  // r = nb * 64 + lane
            wv[k] = w4[((size_t)t * nslice + sl) * (NBT * 64) + (int)blockIdx.y * NB * 64 + r];
        }
    };
    // with the weights in LDS the next slice's loads are issued right after a slice is stored and
    // land while its taps run; wider layers stream their weights from global memory during the
    // taps, whose waits would then include those loads, so they stage each slice just before
    // storing it
    if constexpr (WLDS) stage(0);
    for (int sl = 0; sl < nslice; ++sl) {
        if constexpr (!WLDS) stage(sl);
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int i = threadIdx.x + k * 256;  // pieces past NT land in the padding
            const int g8 = i & 3, P = i >> 2;      // P = tile row * LT_PX + pixel
            if constexpr (WLDS)
                reinterpret_cast<uint4*>(tile)[P * 4 + (g8 ^ ((P >> 1) & 3))] = v[k];
            else
                *reinterpret_cast<uint4*>(tile + P * LT_CP + g8 * 8) = v[k];
        }
#pragma unroll
        for (int k = 0; k < NW; ++k) reinterpret_cast<u32x4*>(s_w)[threadIdx.x + k * 256] = wv[k];
        __syncthreads();
        if constexpr (WLDS)
            if (sl + 1 < nslice) stage(sl + 1);
        if constexpr (WLDS) {
            // wavefront w computes pixels x0 + 16w .. + 15 of all ROWS rows (acc[row][nb]): the
            // fragment of one input row and tap column feeds every output row it reaches (up to 3),
            // so a 4-row slice takes 18 activation and 9 NB weight fragment reads for 36 NB MFMAs
            // (one row x 64 pixels per wavefront took 36 and 9 NB), an 8-row one 30 and 9 NB for 72 NB
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                Frag8 cw[3][NB];
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int j = 0; j < NB; ++j) cw[dy][j].u = s_w[((dy * 3 + dx) * NB + j) * 64 + lane];
#pragma unroll
                for (int ry = 0; ry < TR; ++ry) {
                    const int P = ry * LT_PX + wave * 16 + lr + dx;
                    Frag8 fa;
                    fa.u = reinterpret_cast<const uint4*>(tile)[P * 4 + (q ^ ((P >> 1) & 3))];
#pragma unroll
                    for (int orow = 0; orow < ROWS; ++orow) {
                        const int dy = ry - orow;  // input row y0 - 1 + ry = output row y0 + orow + dy - 1
                        if (dy < 0 || dy > 2) continue;
#pragma unroll
                        for (int j = 0; j < NB; ++j)
                            acc[orow][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cw[dy][j].v, fa.v, acc[orow][j], 0, 0, 0);
                    }
                }
            }
        } else {
            // wavefront w computes row y0 + w (acc[pb][nb]: pixels x0 + 16 pb ..), its weights
            // read from global memory one tap ahead; the tap loop stays rolled (unrolled, its
            // registers would cost these kernels half their occupancy)
            auto load_w = [&](int t, Frag8 (&fb)[NB]) {
#pragma unroll
                for (int j = 0; j < NB; ++j)
                    fb[j].u = wl4[(((size_t)t * nslice + sl) * NBT + (int)blockIdx.y * NB + j) * 64 + lane];
            };
            Frag8 fw[NB];
            load_w(0, fw);
#pragma unroll 1
            for (int t = 0; t < 9; ++t) {
                Frag8 cw[NB];
#pragma unroll
                for (int j = 0; j < NB; ++j) cw[j] = fw[j];
                if (t < 8) load_w(t + 1, fw);
                const int ry = wave + t / 3;  // tile row of input row y + t/3 - 1
                const int dx = t % 3;         // tile pixel offset = output pixel - x0 + dx
                Frag8 fa[4];
#pragma unroll
                for (int pb = 0; pb < 4; ++pb)
                    fa[pb].u = *reinterpret_cast<const uint4*>(tile + (ry * LT_PX + pb * 16 + lr + dx) * LT_CP + q * 8);
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                    for (int pb = 0; pb < 4; ++pb)
                        acc[pb][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cw[j].v, fa[pb].v, acc[pb][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if constexpr (WLDS) {  // acc[row]: the wavefront's 16 pixels of row y0 + row
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const int y = y0 + r;
            const f32x4(&ar)[1][NB] = *reinterpret_cast<const f32x4(*)[1][NB]>(&acc[r]);
            if (a.Yf)
                store_row_t<1, NB, true>(a, ar, b, y, y < a.H, x0 + 16 * wave, q, lr, c0);
            else
                store_row_t<1, NB, false>(a, ar, b, y, y < a.H, x0 + 16 * wave, q, lr, c0);
        }
    } else {  // acc[pb]: pixels x0 + 16 pb .. of row y0 + wave
        const int y = y0 + wave;
        if (a.Yf)
            store_row_t<4, NB, true>(a, acc, b, y, y < a.H, x0, q, lr, c0);
        else
            store_row_t<4, NB, false>(a, acc, b, y, y < a.H, x0, q, lr, c0);
    }
}

// The same convolution with the input brought into LDS by direct loads (buffer_load_dwordx4 ...
// lds: 16 bytes per lane into consecutive LDS slots), double buffered: a persistent workgroup walks
// its tiles and issues the next tile's loads before it computes the current one, without holding
// the data in registers. The input is X1 only (C2 = 0) and at most 2 GB, so that the buffer
// offsets fit 31 bits (the launch checks both). The tile is unpadded, and a
// pixel's four 8-channel groups are stored XOR-swizzled (slot 4P + (g ^ ((P >> 1) & 3)) for tile
// pixel P), so that the 8 consecutive pixels of an activation-fragment read phase cover all 32 banks.
// Pieces outside the image or past CI load 16 zero bytes (out-of-range buffer loads), so every wave
// issues the same fixed number of loads per item, and the compiler can count them when it waits
// for anything else. The kernel serves single-slice layers (CI <= 32) of up to 32 output
// channels, whose weights sit in LDS for the whole launch. (A multi-slice variant that loaded
// each slice's weights into registers per item measured 2.4x slower than the LDS-staged kernel
// on the 80-channel decoder layer: the weight latency it exposes per item is not hidden.)
constexpr int DT_ROWS = 6;                      // staged rows of a 4-row tile
constexpr int DT_SLOTS = DT_ROWS * LT_PX * 4;    // 1584 16-byte slots of a tile slice
constexpr int DT_INST = (DT_SLOTS + 255) / 256 * 4;  // 28 wave-wide loads: 7 per wave
constexpr int DT_ALLOC = DT_INST * 64;           // slots allocated per buffer

template <int NB, bool F32>
__global__ __launch_bounds__(256) void k_conv_fwd_dma(ConvArgs a) {
    __shared__ uint4 tA[DT_ALLOC];
    __shared__ uint4 tB[DT_ALLOC];
    constexpr int NWI = (9 * NB * 64 + 255) / 256;
    __shared__ uint4 s_w[NWI * 256];  // padded: the weight stores are unconditional
    const int CI = a.C1 + a.C2;  // <= 32: one slice
    const int runs = (a.W + 63) / 64;
    const int rblocks = (a.H + 3) / 4;
    const int ntiles = a.B * rblocks * runs;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: the store rows
    const int lane = __lane_id();
    const int q = lane >> 4, lr = lane & 15;
    const uint4* wl4 = reinterpret_cast<const uint4*>(a.Wl);
    int tl = (int)blockIdx.x;
    if (tl >= ntiles) return;
    // s_waitcnt immediate (gfx9 layout: vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8) that
    // waits until at most kStores vector memory operations and no LDS operation are outstanding
    constexpr int kStores = 4 * NB;  // store_row_t: one store per 16-pixel block and channel block
    constexpr int kWaitLoads = (kStores & 15) | ((kStores >> 4) << 14) | (7 << 4);  // and lgkmcnt(0)
    f32x4 bias[NB];
    load_bias_t<NB>(a, q, bias);
    {  // the layer's weights, once (slot = (t * NB + nb) * 64 + lane)
        uint4 wv[NWI];
#pragma unroll
        for (int k = 0; k < NWI; ++k) {
            const int i = threadIdx.x + k * 256;
            wv[k] = wl4[i < 9 * NB * 64 ? i : 0];
        }
#pragma unroll
        for (int k = 0; k < NWI; ++k) s_w[threadIdx.x + k * 256] = wv[k];
    }
    auto tile_origin = [&](int t, int& b, int& x0, int& y0) {
        const int run = t % runs;
        const int r2 = t / runs;
        y0 = (r2 % rblocks) * 4;
        b = r2 / rblocks;
        x0 = run * 64;
    };
    // Loads of a tile into buf: wave w issues the wave-wide loads w, w + 4, ..., 7 of them, as
    // buffer loads over the input tensor. A slot's piece (tile row, pixel, 8-channel group) is the
    // same for every tile, so its byte offset from the tile's corner and its place are computed once;
    // per tile only the corner and the image bounds change. A piece outside the image or past CI
    // gets an offset past the tensor, which the hardware answers with 16 zero bytes.
    constexpr int NSLOT = DT_INST / 4;
    int rel[NSLOT];    // byte offset of the piece from pixel (y0 - 1, x0 - 1), channel 0
    int where[NSLOT];  // tile row | pixel << 8, or -1 for a slot past the tile or past CI
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) {
        const int S = (wave + 4 * k) * 64 + lane;
        const int P = S >> 2, g = (S & 3) ^ ((P >> 1) & 3);
        const int ry = P / LT_PX, pxl = P - ry * LT_PX;
        rel[k] = ((ry * a.W + pxl) * a.C1 + g * 8) * 2;
        where[k] = (P < DT_ROWS * LT_PX && g * 8 < CI) ? (ry | (pxl << 8)) : -1;
    }
    const __amdgpu_buffer_rsrc_t src =
        __builtin_amdgcn_make_buffer_rsrc((void*)a.X1, 0, a.B * a.H * a.W * a.C1 * 2, 0x00020000);
    auto issue = [&](uint4* buf, int t) {
        int b, x0, y0;
        tile_origin(t, b, x0, y0);
        const int corner = (((b * a.H + y0 - 1) * a.W) + x0 - 1) * a.C1 * 2;
#pragma unroll
        for (int k = 0; k < NSLOT; ++k) {
            const int yy = y0 - 1 + (where[k] & 0xff), xx = x0 - 1 + ((where[k] >> 8) & 0xff);
            const bool ok = where[k] >= 0 && (unsigned)yy < (unsigned)a.H && (unsigned)xx < (unsigned)a.W;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                src, (__attribute__((address_space(3))) void*)(buf + (wave + 4 * k) * 64), 16,
                ok ? corner + rel[k] : (int)0x80000000, 0, 0, 0);
        }
    };
    // the 9 taps unrolled: in a rolled loop the compiler moved the accumulators between AGPRs on every
    // iteration (24 v_accvgpr moves per tap, as many VALU instructions as the tap's address math)
    auto compute = [&](const uint4* buf, f32x4 (&acc)[4][NB]) {
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            Frag8 cw[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) cw[j].u = s_w[(t * NB + j) * 64 + lane];
            const int ry = wave + t / 3;
            const int dx = t % 3;
            Frag8 fa[4];
#pragma unroll
            for (int pb = 0; pb < 4; ++pb) {
                const int P = ry * LT_PX + pb * 16 + lr + dx;
                fa[pb].u = buf[P * 4 + (q ^ ((P >> 1) & 3))];
            }
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int pb = 0; pb < 4; ++pb)
                    acc[pb][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cw[j].v, fa[pb].v, acc[pb][j], 0, 0, 0);
        }
    };
    issue(tA, tl);
    // the first item's loads are followed by no stores: wait for them outright (vmcnt(0))
    __builtin_amdgcn_s_waitcnt(7 << 4);
    f32x4 acc[4][NB];  // every item starts from the bias, which the first tap's MFMAs take as C
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = bias[j];
    // one step: wait for this item, start the next item into the other buffer, compute this one
    // and store it. The barrier opening a step is the only one: before it each wave waits for its
    // own loads into `cur` (issued a whole step earlier; the stores issued after them have a fixed
    // count, so the wait does not include them), and after it every wave is also done reading
    // `nxt` (it read it in the previous step's compute, whose results it has already stored). The
    // buffers are distinct arrays, so the compiler knows that reads of one need not wait for loads
    // into the other.
    auto step = [&](const uint4* cur, uint4* nxt) -> bool {
        // this wave's loads into cur were followed by exactly kStores stores (before the first
        // step there are none, and the wait after the first issue below covers it); vector memory
        // operations complete in issue order, so vmcnt(kStores) means the loads have landed
        // (a plain barrier: __syncthreads' fence would also wait for the stores). The signal
        // fences keep the compiler from moving LDS accesses across it.
        __builtin_amdgcn_s_waitcnt(kWaitLoads);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const int ntl = tl + (int)gridDim.x;
        const bool more = ntl < ntiles;
        // issued unconditionally (the last item reloads itself) so that every path has the same
        // number of loads in flight and the compiler's waits can count them
        issue(nxt, more ? ntl : tl);
        compute(cur, acc);
        int b, x0, y0;
        tile_origin(tl, b, x0, y0);
        const int y = y0 + wave;
        store_row_t<4, NB, F32>(a, acc, b, y, y < a.H, x0, q, lr);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) acc[i][j] = bias[j];
        tl = ntl;
        return more;
    };
    while (true) {
        if (!step(tA, tB)) break;
        if (!step(tB, tA)) break;
    }
    // the reload issued by the last step lands in LDS: wait for it before the workgroup ends
    __syncthreads();
}

// ------------------------------------------------------------------ backward data
// dXc[p, ci] = sum_t sum_co G[p - o_t, co] * W[t][ci][co], G = dY * (Y > 0 if relu).
// dXc is [B,H,W,CI] bf16 (full-resolution concatenated input gradient).
// Wb: weights in fragment order [tap][co/16][ci/16][ci%16][co%16].
template <int NB>
__global__ __launch_bounds__(256) void k_conv_bwd_data(ConvBwdArgs a) {
    const int lane = __lane_id();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int runs_per_row = (a.W + 63) / 64;
    const int row = wave / runs_per_row;
    if (row >= a.B * a.H) return;
    const int x0 = (wave % runs_per_row) * 64;
    const int b = row / a.H, y = row % a.H;
    const int nb0 = blockIdx.y * NB;
    const int KC = a.CO / 16;  // k chunks per tap (output channels of the forward conv)
    const int NBT = a.CI / 16;
    f32x4 acc[4][NB];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int kq = lane >> 4, lr = lane & 15;
    for (int t = 0; t < 9; ++t) {
        const int yy = y - (t / 3 - 1);
        const bool yok = yy >= 0 && yy < a.H;
        for (int kc = 0; kc < KC; ++kc) {
            Frag fa[4];
#pragma unroll
            for (int pb = 0; pb < 4; ++pb) {
                int xx = x0 + pb * 16 + lr - (t % 3 - 1);
                fa[pb].u = make_uint2(0u, 0u);
                if (yok && xx >= 0 && xx < a.W) {
                    size_t o = (((size_t)b * a.H + yy) * a.W + xx) * a.CO + kc * 16 + kq * 4;
                    Frag g;
                    g.u = *reinterpret_cast<const uint2*>(a.dY + o);
                    if (a.relu) {
                        Frag yv;
                        yv.u = *reinterpret_cast<const uint2*>(a.Y + o);
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            if (!(bf2f(yv.h[e]) > 0.0f)) g.h[e] = 0;
                    }
                    fa[pb] = g;
                }
            }
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                int nb = nb0 + j;
                if (nb >= NBT) break;
                Frag fb;
                fb.u = *reinterpret_cast<const uint2*>(a.Wt + ((((size_t)t * KC + kc) * NBT + nb) * 16 + lr) * 16 + kq * 4);
#pragma unroll
                for (int pb = 0; pb < 4; ++pb)
                    acc[pb][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[pb].v, fb.v, acc[pb][j], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        int nb = nb0 + j;
        if (nb >= NBT) break;
        int ci = nb * 16 + lr;
#pragma unroll
        for (int pb = 0; pb < 4; ++pb)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                int x = x0 + pb * 16 + kq * 4 + e;
                if (x >= a.W) continue;
                a.dX[(((size_t)b * a.H + y) * a.W + x) * a.CI + ci] = f2bf(acc[pb][j][e]);
            }
    }
}

// ------------------------------------------------------------------ backward weights
// dW[t][ci][co] += sum_p X[p + o_t, ci] * G[p, co] over this block's pixel range; one wave per
// (tap, 16 input channels) and pixel split; K of the MFMA runs over 16 pixels at a time.
template <int NB>
__global__ __launch_bounds__(64) void k_conv_bwd_weight(ConvWgtArgs a) {
    const int lane = __lane_id();
    const int CC = (a.C1 + a.C2) / 16;
    const int CC1 = a.C1 / 16;
    const int t = blockIdx.x / CC, cc = blockIdx.x % CC;
    const int NBT = a.CO / 16;
    const int nb0 = blockIdx.z * NB;
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    const long P = (long)a.B * a.H * a.W;
    const long per = (P + gridDim.y - 1) / gridDim.y;
    const long p0 = (long)blockIdx.y * per;
    const long p1 = p0 + per < P ? p0 + per : P;
    f32x4 acc[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int kq = lane >> 4, lr = lane & 15;
    for (long base = p0; base < p1; base += 16) {
        Frag fa;
        Frag fg[NB];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            long p = base + kq * 4 + e;
            uint16_t xv = 0;
            bool in = p < p1;
            int b = 0, y = 0, x = 0;
            if (in) {
                x = (int)(p % a.W);
                long r = p / a.W;
                y = (int)(r % a.H);
                b = (int)(r / a.H);
                int yy = y + dy, xx = x + dx;
                if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
                    if (cc < CC1)
                        xv = a.X1[(((size_t)b * a.H + yy) * a.W + xx) * a.C1 + cc * 16 + lr];
                    else
                        xv = a.X2[(((size_t)b * (a.H >> 1) + (yy >> 1)) * (a.W >> 1) + (xx >> 1)) * a.C2 +
                                  (cc - CC1) * 16 + lr];
                }
            }
            fa.h[e] = xv;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                int nb = nb0 + j;
                uint16_t gv = 0;
                if (in && nb < NBT) {
                    size_t o = (size_t)p * a.CO + nb * 16 + lr;
                    gv = a.dY[o];
                    if (a.relu && !(bf2f(a.Y[o]) > 0.0f)) gv = 0;
                }
                fg[j].h[e] = gv;
            }
        }
        // A = X^T (rows: 16 input channels, k: 16 pixels), B = G (k: 16 pixels, columns: 16 outputs)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa.v, fg[j].v, acc[j], 0, 0, 0);
    }
    // result rows = input channels 4*kq+e, column = output channel lr
    const int CI = a.C1 + a.C2;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        int nb = nb0 + j;
        if (nb >= NBT) break;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int ci = cc * 16 + kq * 4 + e;
            atomicAdd(a.dW + ((size_t)t * CI + ci) * a.CO + nb * 16 + lr, acc[j][e]);
        }
    }
}

// bias gradient: db[co] = sum_p G[p, co]
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t* __restrict__ dY, const uint16_t* __restrict__ Y,
                                                   int relu, long P, int CO, float* db) {
    int co = blockIdx.y;
    float s = 0.0f;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
        size_t o = (size_t)p * CO + co;
        float g = bf2f(dY[o]);
        if (relu && !(bf2f(Y[o]) > 0.0f)) g = 0.0f;
        s += g;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ float ws[4];
    if (__lane_id() == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(db + co, ws[0] + ws[1] + ws[2] + ws[3]);
}

// ------------------------------------------------------------------ pooling / concat split
// 2x2 average pooling, 8 channels per thread (C is a multiple of 8): four 16-byte loads, one
// 16-byte store; each channel is ((x00 + x01) + x10) + x11, times 0.25.
__global__ void k_avgpool(const uint16_t* __restrict__ X, uint16_t* __restrict__ Y, int B, int H, int W, int C) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    long n = (long)B * Ho * Wo * C8;
    if (i >= n) return;
    int c = (int)(i % C8) * 8;
    long r = i / C8;
    int xo = (int)(r % Wo);
    r /= Wo;
    int yo = (int)(r % Ho);
    int b = (int)(r / Ho);
    const uint16_t* s = X + (((size_t)b * H + 2 * yo) * W + 2 * xo) * C + c;
    const uint4 q00 = *reinterpret_cast<const uint4*>(s);
    const uint4 q01 = *reinterpret_cast<const uint4*>(s + C);
    const uint4 q10 = *reinterpret_cast<const uint4*>(s + (size_t)W * C);
    const uint4 q11 = *reinterpret_cast<const uint4*>(s + (size_t)W * C + C);
    const uint32_t a00[4] = {q00.x, q00.y, q00.z, q00.w}, a01[4] = {q01.x, q01.y, q01.z, q01.w};
    const uint32_t a10[4] = {q10.x, q10.y, q10.z, q10.w}, a11[4] = {q11.x, q11.y, q11.z, q11.w};
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t w = 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int sh = 16 * h;
            const float v = bf2f((uint16_t)(a00[k] >> sh)) + bf2f((uint16_t)(a01[k] >> sh)) +
                            bf2f((uint16_t)(a10[k] >> sh)) + bf2f((uint16_t)(a11[k] >> sh));
            w |= (uint32_t)f2bf(0.25f * v) << sh;
        }
        o[k] = w;
    }
    *reinterpret_cast<uint4*>(Y + (((size_t)b * Ho + yo) * Wo + xo) * C + c) = make_uint4(o[0], o[1], o[2], o[3]);
}

// dX(full) += dPooled / 4 spread over each 2x2 block (accumulate into an existing gradient)
__global__ void k_avgpool_bwd(const uint16_t* __restrict__ dY, uint16_t* __restrict__ dX, int B, int H, int W, int C,
                              int accumulate) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)B * H * W * C;
    if (i >= n) return;
    int c = (int)(i % C);
    long r = i / C;
    int x = (int)(r % W);
    r /= W;
    int y = (int)(r % H);
    int b = (int)(r / H);
    float g = 0.25f * bf2f(dY[(((size_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c]);
    if (accumulate) g += bf2f(dX[i]);
    dX[i] = f2bf(g);
}

// Split the gradient of a concatenated input: channels [0,C1) go to dX1 (full res),
// channels [C1,C1+C2) are summed over 2x2 blocks into dX2 (half res, nearest-upsample adjoint).
__global__ void k_concat_split(const uint16_t* __restrict__ dXc, int B, int H, int W, int C1, int C2,
                               uint16_t* __restrict__ dX1, uint16_t* __restrict__ dX2) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int CI = C1 + C2;
    long n1 = (long)B * H * W * C1;
    long n2 = (long)B * (H / 2) * (W / 2) * C2;
    if (i < n1) {
        int c = (int)(i % C1);
        long p = i / C1;
        dX1[i] = dXc[(size_t)p * CI + c];
    }
    if (i < n2) {
        int c = (int)(i % C2);
        long r = i / C2;
        int xo = (int)(r % (W / 2));
        r /= (W / 2);
        int yo = (int)(r % (H / 2));
        int b = (int)(r / (H / 2));
        const uint16_t* s = dXc + (((size_t)b * H + 2 * yo) * W + 2 * xo) * CI + C1 + c;
        float v = bf2f(s[0]) + bf2f(s[CI]) + bf2f(s[(size_t)W * CI]) + bf2f(s[(size_t)W * CI + CI]);
        dX2[i] = f2bf(v);
    }
}

// ------------------------------------------------------------------ features / output
// Raw feature record per pixel (12 halfs): colour rgb, albedo rgb, normal xyz, t, visibility, 0.
// Network input (16 bf16): visibility, albedo rgb, normal xyz, depth t/depth_scale, colour rgb,
// hit mask, 0...
__global__ void k_pack_input(const uint16_t* __restrict__ raw, int B, int H, int W, int Hp, int Wp, float inv_depth,
                             uint16_t* __restrict__ X0) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)B * Hp * Wp;
    if (i >= n) return;
    int x = (int)(i % Wp);
    long r = i / Wp;
    int y = (int)(r % Hp);
    int b = (int)(r / Hp);
    float f[16];
    for (int k = 0; k < 16; ++k) f[k] = 0.0f;
    if (x < W && y < H) {
        const uint16_t* s = raw + (((size_t)b * H + y) * W + x) * 12;
        float v[12];
        for (int k = 0; k < 12; ++k) v[k] = __half2float(__ushort_as_half(s[k]));
        float t = v[9];
        bool hit = t > 0.0f;
        f[0] = v[10];
        f[1] = v[3]; f[2] = v[4]; f[3] = v[5];
        f[4] = v[6]; f[5] = v[7]; f[6] = v[8];
        f[7] = hit ? t * inv_depth : 0.0f;
        f[8] = v[0]; f[9] = v[1]; f[10] = v[2];
        f[11] = hit ? 1.0f : 0.0f;
    }
    uint16_t* d = X0 + i * 16;
    for (int k = 0; k < 16; ++k) d[k] = f2bf(f[k]);
}

// Frame -> raw features and network input in one pass (device path of the renderer output): raw
// = colour (as half), the 8 feature halfs, 0; the input channels are those k_pack_input makes of
// raw (colour rounded to half first, then every channel to bf16). 16-byte loads, 8/16-byte stores.
__global__ void k_frame_input(const float4* __restrict__ col, const __half* __restrict__ aux, int H, int W, int Hp,
                              int Wp, float inv_depth, uint16_t* __restrict__ raw, uint16_t* __restrict__ X0) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)Hp * Wp;
    if (i >= n) return;
    const int x = (int)(i % Wp), y = (int)(i / Wp);
    uint32_t o[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
    if (x < W && y < H) {
        const long p = (long)y * W + x;
        const float4 c = col[p];
        const uint4 q = reinterpret_cast<const uint4*>(aux)[p];  // 8 halfs: albedo 3, normal 3, t, coverage
        const uint16_t h[8] = {(uint16_t)q.x, (uint16_t)(q.x >> 16), (uint16_t)q.y, (uint16_t)(q.y >> 16),
                               (uint16_t)q.z, (uint16_t)(q.z >> 16), (uint16_t)q.w, (uint16_t)(q.w >> 16)};
        const uint16_t c0 = __half_as_ushort(__float2half(c.x)), c1 = __half_as_ushort(__float2half(c.y)),
                       c2 = __half_as_ushort(__float2half(c.z));
        if (raw) {  // raw: colour 3, aux 8, 0 (a rendered frame's compose reads the aux buffer instead)
            uint2* d = reinterpret_cast<uint2*>(raw + p * 12);
            d[0] = make_uint2((uint32_t)c0 | ((uint32_t)c1 << 16), (uint32_t)c2 | ((uint32_t)h[0] << 16));
            d[1] = make_uint2((uint32_t)h[1] | ((uint32_t)h[2] << 16), (uint32_t)h[3] | ((uint32_t)h[4] << 16));
            d[2] = make_uint2((uint32_t)h[5] | ((uint32_t)h[6] << 16), (uint32_t)h[7]);
        }
        auto hf = [](uint16_t v) { return __half2float(__ushort_as_half(v)); };
        const float t = hf(h[6]);
        const bool hit = t > 0.0f;
        float f[16];
        f[0] = hf(h[7]);
        f[1] = hf(h[0]); f[2] = hf(h[1]); f[3] = hf(h[2]);
        f[4] = hf(h[3]); f[5] = hf(h[4]); f[6] = hf(h[5]);
        f[7] = hit ? t * inv_depth : 0.0f;
        f[8] = hf(c0); f[9] = hf(c1); f[10] = hf(c2);
        f[11] = hit ? 1.0f : 0.0f;
        f[12] = f[13] = f[14] = f[15] = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) o[k] = (uint32_t)f2bf(f[2 * k]) | ((uint32_t)f2bf(f[2 * k + 1]) << 16);
    }
    uint4* dx = reinterpret_cast<uint4*>(X0 + i * 16);
    dx[0] = make_uint4(o[0], o[1], o[2], o[3]);
    dx[1] = make_uint4(o[4], o[5], o[6], o[7]);
}


// Network output (fp32 [B,Hp,Wp,16]: v in channel 0, colour residual in 1-3) -> colour
// albedo * v + residual on hit pixels, background elsewhere. With a target, also the squared
// error and d(loss)/d(out).
__global__ __launch_bounds__(256) void k_compose(const float* __restrict__ out, const uint16_t* __restrict__ feat,
                                                 int fstride, int fa, int ft, int B, int H, int W, int Hp, int Wp,
                                                 float4 bg,
                                                 const float* __restrict__ target, float inv_count,
                                                 float* __restrict__ color, uint16_t* __restrict__ dout,
                                                 float* __restrict__ loss, float4* __restrict__ fb_col,
                                                 uint32_t* __restrict__ fb_rgba8) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)B * Hp * Wp;
    float se = 0.0f;
    if (i < n) {
        int x = (int)(i % Wp);
        long r = i / Wp;
        int y = (int)(r % Hp);
        int b = (int)(r / Hp);
        float g0 = 0.0f, gr[3] = {0.0f, 0.0f, 0.0f};
        if (x < W && y < H) {
            size_t q = ((size_t)b * H + y) * W + x;
            const uint16_t* s = feat + q * fstride;
            float t = __half2float(__ushort_as_half(s[ft]));
            float al[3] = {__half2float(__ushort_as_half(s[fa])), __half2float(__ushort_as_half(s[fa + 1])),
                           __half2float(__ushort_as_half(s[fa + 2]))};
            const float4 o4 = reinterpret_cast<const float4*>(out)[i];  // visibility, colour residual
            const float v = o4.x, res[3] = {o4.y, o4.z, o4.w};
            float c[3];
            bool hit = t > 0.0f;
            for (int k = 0; k < 3; ++k)
                c[k] = hit ? al[k] * v + res[k] : (k == 0 ? bg.x : (k == 1 ? bg.y : bg.z));
            if (color)
                for (int k = 0; k < 3; ++k) color[q * 3 + k] = c[k];
            if (fb_col) {  // a rendered frame: colour into the framebuffer (alpha kept), RGBA8 beside it
                float4 f = fb_col[q];
                f.x = c[0];
                f.y = c[1];
                f.z = c[2];
                fb_col[q] = f;
                fb_rgba8[q] = pack_rgba8(f);
            }
            if (target) {
                for (int k = 0; k < 3; ++k) {
                    float d = c[k] - target[q * 3 + k];
                    se += d * d;
                    if (hit) {
                        g0 += 2.0f * d * al[k] * inv_count;
                        gr[k] = 2.0f * d * inv_count;
                    }
                }
            }
        }
        if (dout) {
            uint16_t* d = dout + i * 16;
            d[0] = f2bf(g0);
            for (int k = 1; k < 4; ++k) d[k] = f2bf(gr[k - 1]);
            for (int k = 4; k < 16; ++k) d[k] = 0;
        }
    }
    if (loss) {
        for (int off = 32; off > 0; off >>= 1) se += __shfl_xor(se, off, 64);
        __shared__ float ws[4];
        if (__lane_id() == 0) ws[threadIdx.x >> 6] = se;
        __syncthreads();
        if (threadIdx.x == 0) atomicAdd(loss, (ws[0] + ws[1] + ws[2] + ws[3]) * inv_count);
    }
}

// Denoised colour back into the frame buffers.
// ------------------------------------------------------------------ parameters
// Canonical fp32 layout per layer: W[t][ci][co] then bias[co]. Fragment layouts for the MFMA
// kernels are rebuilt from it after every update.
__global__ void k_relayout(const float* __restrict__ Wc, int CI, int CO, uint16_t* __restrict__ Wf,
                           uint16_t* __restrict__ Wb) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = 9L * CI * CO;
    if (i >= n) return;
    int co = (int)(i % CO);
    long r = i / CO;
    int ci = (int)(r % CI);
    int t = (int)(r / CI);
    uint16_t v = f2bf(Wc[i]);
    // forward: [t][ci/16][co/16][co%16][ci%16]
    int CC = CI / 16, NBT = CO / 16;
    Wf[((((size_t)t * CC + ci / 16) * NBT + co / 16) * 16 + co % 16) * 16 + ci % 16] = v;
    // backward data: [t][co/16][ci/16][ci%16][co%16]
    int KC = CO / 16, NBI = CI / 16;
    Wb[((((size_t)t * KC + co / 16) * NBI + ci / 16) * 16 + ci % 16) * 16 + co % 16] = v;
}

// Canonical W[t][ci][co] -> [t][ci/32][co/16][lane][8], zero past CI (the LDS-staged kernel).
__global__ void k_relayout_lds(const float* __restrict__ Wc, int CI, int CO, uint16_t* __restrict__ Wl) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nslice = (CI + 31) / 32, NBT = CO / 16;
    long n = 9L * nslice * NBT * 64 * 8;
    if (i >= n) return;
    const int e = (int)(i % 8);
    long r = i / 8;
    const int lane = (int)(r % 64);
    r /= 64;
    const int nb = (int)(r % NBT);
    r /= NBT;
    const int sl = (int)(r % nslice);
    const int t = (int)(r / nslice);
    const int ci = sl * 32 + 8 * (lane >> 4) + e;
    const int co = nb * 16 + (lane & 15);
    Wl[i] = ci < CI ? f2bf(Wc[((size_t)t * CI + ci) * CO + co]) : (uint16_t)0;
}

// Canonical W[t][ci][co] -> [k/32][co/16][lane][8] with k = t*CI + ci, zero past 9*CI.
__global__ void k_relayout32(const float* __restrict__ Wc, int CI, int CO, uint16_t* __restrict__ W32) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nkc = (9 * CI + 31) / 32;
    long n = (long)nkc * 32 * CO;
    if (i >= n) return;
    const int e = (int)(i % 8);
    long r = i / 8;
    const int lane = (int)(r % 64);
    r /= 64;
    const int NBT = CO / 16;
    const int nb = (int)(r % NBT);
    const int kc = (int)(r / NBT);
    const int k = kc * 32 + 8 * (lane >> 4) + e;
    const int co = nb * 16 + (lane & 15);
    W32[i] = k < 9 * CI ? f2bf(Wc[(size_t)k * CO + co]) : (uint16_t)0;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void k_zero_f32(float* p, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0f;
}

inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ------------------------------------------------------------------ host side
const LayerSpec kLayers[kNumLayers] = {
    // name, C1 (full-res input), C2 (half-res input), CO, relu
    {"enc0a", 16, 0, 32, 1}, {"enc0b", 32, 0, 32, 1},  {"enc1a", 32, 0, 48, 1}, {"enc1b", 48, 0, 48, 1},
    {"enc2a", 48, 0, 64, 1}, {"enc2b", 64, 0, 64, 1},  {"bota", 64, 0, 96, 1},  {"botb", 96, 0, 96, 1},
    {"dec2a", 64, 96, 64, 1}, {"dec2b", 64, 0, 64, 1}, {"dec1a", 48, 64, 48, 1}, {"dec1b", 48, 0, 48, 1},
    {"dec0a", 32, 48, 32, 1}, {"dec0b", 32, 0, 32, 1}, {"out", 32, 0, 16, 0},
};

size_t param_count() {
    size_t n = 0;
    for (const auto& L : kLayers) n += 9ull * (L.c1 + L.c2) * L.co + L.co;
    return n;
}

static size_t param_offset(int l) {
    size_t n = 0;
    for (int i = 0; i < l; ++i) n += 9ull * (kLayers[i].c1 + kLayers[i].c2) * kLayers[i].co + kLayers[i].co;
    return n;
}

// shape = 0 picks the kernel and launch shape from the launch size (the LDS-staged kernel for
// large launches, otherwise the 16x16x32 kernel with fewer channel and pixel blocks per wave
// as the launch shrinks); a test hook forces one: 5 = LDS-staged, else pixel blocks per wave
// (shape % 10: 4, 2 or 1) and, for 4, channel blocks per wave (shape / 10, 0 = automatic).
// Persistent grid of the LDS-DMA kernel: the workgroups that fit on the device at once (queried
// once per instance), at most one per tile.
template <int NB, bool F32>
void launch_dma(hipStream_t s, const ConvArgs& a, long tiles) {
    static int fit = 0;
    if (fit == 0) {
        int per_cu = 0, dev = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_conv_fwd_dma<NB, F32>, 256, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
        fit = per_cu * cus;
    }
    const long g = tiles < fit ? tiles : fit;
    hipLaunchKernelGGL((k_conv_fwd_dma<NB, F32>), dim3((unsigned)g), dim3(256), 0, s, a);
}

static void conv_fwd(hipStream_t s, const ConvArgs& a, bool k32, int shape) {
    long waves = (long)a.B * a.H * ((a.W + 63) / 64);
    int NBT = a.CO / 16;
    int nb = NBT <= 6 ? NBT : 4;
    dim3 g(nblk(waves * 64), (NBT + nb - 1) / nb);
    if (k32) {
        // enough wavefronts to fill 256 CUs: fewer channel blocks, then fewer pixels per wave
        const long want = 6144;
        int pbk = 4;
        // the LDS-DMA double-buffered kernel (persistent) for single-slice layers of up to 32
        // output channels: the default for them (shape 0) and shape 6
        if ((shape == 0 || shape == 6) && a.C2 == 0 && a.C1 <= 32 && NBT <= (a.Yf ? 1 : 2) &&
            (double)a.B * a.H * a.W * a.C1 * 2 < 2147483648.0) {
            const long tiles = (long)a.B * ((a.H + 3) / 4) * ((a.W + 63) / 64);
            if (a.Yf)
                launch_dma<1, true>(s, a, tiles);  // the output layer (16 channels)
            else if (NBT == 1)
                launch_dma<1, false>(s, a, tiles);
            else
                launch_dma<2, false>(s, a, tiles);
            return;
        }
        // the LDS-staged kernel for every layer (it beats the per-wavefront kernel at all four
        // resolutions); shapes 5 and 7 force it (7 without output-channel groups), other nonzero
        // shapes but 6 force the per-wavefront kernel (its buffer loads and stores take 31-bit
        // offsets: tensors under 2 GB)
        const bool small = (double)a.B * a.H * a.W * std::max(std::max(a.C1, a.CO * (a.Yf ? 2 : 1)), 1) * 2 <
                           2147483648.0;
        const bool lds = small && (shape == 5 || shape == 6 || shape == 7 || (shape == 0 && NBT <= 6));
        if (lds) {
            // 96-channel layers run as two groups of 48 output channels (grid.y), each with its slice
            // of the weights in LDS: the eighth-resolution level then fills the device and no tap
            // waits for weights from global memory (21 -> 13 us a layer). 64-channel layers measured
            // slower as two groups of 32 (27 -> 31 us: the same number of workgroups, each staging
            // the whole input). Shape 7 keeps every layer whole (weights streamed per tap).
            int nb = NBT, groups = 1;
            if (shape != 7 && NBT == 6) {
                nb = 3;
                groups = 2;
            }
            const int rows = lds_tile_rows(nb);
            const long blocks = (long)a.B * ((a.H + rows - 1) / rows) * ((a.W + 63) / 64);
            const dim3 grid((unsigned)blocks, (unsigned)groups);
            switch (nb) {
            case 1: hipLaunchKernelGGL((k_conv_fwd_lds<1>), grid, dim3(256), 0, s, a); break;
            case 2: hipLaunchKernelGGL((k_conv_fwd_lds<2>), grid, dim3(256), 0, s, a); break;
            case 3: hipLaunchKernelGGL((k_conv_fwd_lds<3>), grid, dim3(256), 0, s, a); break;
            case 4: hipLaunchKernelGGL((k_conv_fwd_lds<4>), grid, dim3(256), 0, s, a); break;
            case 5: hipLaunchKernelGGL((k_conv_fwd_lds<5>), grid, dim3(256), 0, s, a); break;
            default: hipLaunchKernelGGL((k_conv_fwd_lds<6>), grid, dim3(256), 0, s, a); break;
            }
            return;
        }
        if (shape != 0) {
            pbk = shape % 10;
            if (pbk != 1 && pbk != 2) pbk = 4;
            if (pbk != 4)
                nb = 1;
            else if (shape / 10 > 0)
                nb = std::min(shape / 10, std::min(NBT, 6));
        } else if (waves * (NBT / nb) < want) {
            nb = 1;
            if (waves * NBT < want) pbk = 2;
            if (waves * 2 * NBT < want) pbk = 1;
        }
        long wv = (long)a.B * a.H * ((a.W + 16 * pbk - 1) / (16 * pbk));
        dim3 g32(nblk(wv * 64), (NBT + nb - 1) / nb);
        if (pbk == 1) {
            hipLaunchKernelGGL((k_conv_fwd32<1, 1>), g32, dim3(256), 0, s, a);
        } else if (pbk == 2) {
            hipLaunchKernelGGL((k_conv_fwd32<1, 2>), g32, dim3(256), 0, s, a);
        } else {
            switch (nb) {
            case 1: hipLaunchKernelGGL((k_conv_fwd32<1, 4>), g32, dim3(256), 0, s, a); break;
            case 2: hipLaunchKernelGGL((k_conv_fwd32<2, 4>), g32, dim3(256), 0, s, a); break;
            case 3: hipLaunchKernelGGL((k_conv_fwd32<3, 4>), g32, dim3(256), 0, s, a); break;
            case 4: hipLaunchKernelGGL((k_conv_fwd32<4, 4>), g32, dim3(256), 0, s, a); break;
            case 5: hipLaunchKernelGGL((k_conv_fwd32<5, 4>), g32, dim3(256), 0, s, a); break;
            default: hipLaunchKernelGGL((k_conv_fwd32<6, 4>), g32, dim3(256), 0, s, a); break;
            }
        }
        return;
    }
    switch (nb) {
    case 1: hipLaunchKernelGGL(k_conv_fwd<1>, g, dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_conv_fwd<2>, g, dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_conv_fwd<3>, g, dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_conv_fwd<4>, g, dim3(256), 0, s, a); break;
    case 5: hipLaunchKernelGGL(k_conv_fwd<5>, g, dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL(k_conv_fwd<6>, g, dim3(256), 0, s, a); break;
    }
}

static void conv_bwd_data(hipStream_t s, const ConvBwdArgs& a) {
    long waves = (long)a.B * a.H * ((a.W + 63) / 64);
    int NBT = a.CI / 16;
    int nb = NBT <= 5 ? NBT : (NBT % 5 == 0 ? 5 : 4);
    dim3 g(nblk(waves * 64), (NBT + nb - 1) / nb);
    switch (nb) {
    case 1: hipLaunchKernelGGL(k_conv_bwd_data<1>, g, dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_conv_bwd_data<2>, g, dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_conv_bwd_data<3>, g, dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_conv_bwd_data<4>, g, dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL(k_conv_bwd_data<5>, g, dim3(256), 0, s, a); break;
    }
}

static void conv_bwd_weight(hipStream_t s, const ConvWgtArgs& a, int cu) {
    int CC = (a.C1 + a.C2) / 16;
    int NBT = a.CO / 16;
    int nb = NBT <= 6 ? NBT : 4;
    long P = (long)a.B * a.H * a.W;
    int splits = (int)std::max<long>(1, std::min<long>((long)cu * 8 / (9 * CC), (P + 1023) / 1024));
    dim3 g(9 * CC, splits, (NBT + nb - 1) / nb);
    switch (nb) {
    case 1: hipLaunchKernelGGL(k_conv_bwd_weight<1>, g, dim3(64), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_conv_bwd_weight<2>, g, dim3(64), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_conv_bwd_weight<3>, g, dim3(64), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_conv_bwd_weight<4>, g, dim3(64), 0, s, a); break;
    case 5: hipLaunchKernelGGL(k_conv_bwd_weight<5>, g, dim3(64), 0, s, a); break;
    default: hipLaunchKernelGGL(k_conv_bwd_weight<6>, g, dim3(64), 0, s, a); break;
    }
}


// feat: the per-pixel features, either the 12-half raw records (colour 3, albedo 3, normal 3, t,
// coverage, 0: stride 12, albedo at 3, t at 9) or a rendered frame's 8-half aux buffer (albedo 3,
// normal 3, t, coverage: stride 8, albedo at 0, t at 6)
void launch_compose(hipStream_t s, const float* out, const uint16_t* feat, int fstride, int fa, int ft, int B, int H,
                    int W, int Hp, int Wp, float4 bg, const float* target, float inv_count, float* color,
                    uint16_t* dout, float* loss, float4* fb_col, uint32_t* fb_rgba8) {
    long n = (long)B * Hp * Wp;
    hipLaunchKernelGGL(k_compose, dim3(nblk(n)), dim3(256), 0, s, out, feat, fstride, fa, ft, B, H, W, Hp, Wp, bg,
                       target, inv_count, color, dout, loss, fb_col, fb_rgba8);
}
void zero(hipStream_t s, float* p, long n) {
    hipLaunchKernelGGL(k_zero_f32, dim3((unsigned)std::min<long>(nblk(n), 4096)), dim3(256), 0, s, p, n);
}
void adam(hipStream_t s, float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
          float c1, float c2) {
    hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
}

int Denoiser::ensure_params(mmhrt_ctx_s* c) {
    size_t n = param_count();
    int e = 0;
    if ((e = mmhrt::ensure(c, params, n, "denoiser params"))) return e;
    if ((e = mmhrt::ensure(c, grads, n, "denoiser grads"))) return e;
    if ((e = mmhrt::ensure(c, adam_m, n, "denoiser adam m"))) return e;
    if ((e = mmhrt::ensure(c, adam_v, n, "denoiser adam v"))) return e;
    size_t wn = 0;
    for (const auto& L : kLayers) wn += 9ull * (L.c1 + L.c2) * L.co;
    if ((e = mmhrt::ensure(c, wf, wn, "denoiser fwd weights"))) return e;
    if ((e = mmhrt::ensure(c, wb, wn, "denoiser bwd weights"))) return e;
    size_t wn32 = 0;
    for (const auto& L : kLayers) wn32 += (size_t)((9 * (L.c1 + L.c2) + 31) / 32) * 32 * L.co;
    if ((e = mmhrt::ensure(c, wf32, wn32, "denoiser fwd weights (k32)"))) return e;
    size_t wnl = 0;
    for (const auto& L : kLayers) wnl += (size_t)9 * ((L.c1 + L.c2 + 31) / 32) * 32 * L.co;
    if ((e = mmhrt::ensure(c, wl, wnl, "denoiser fwd weights (LDS)"))) return e;
    if ((e = mmhrt::ensure(c, scalar, 4, "denoiser scalars"))) return e;
    return 0;
}

void Denoiser::relayout(hipStream_t s) {
    size_t po = 0, wo = 0, w32 = 0, wlo = 0;
    for (const auto& L : kLayers) {
        int CI = L.c1 + L.c2;
        long nw = 9L * CI * L.co;
        long n32 = (long)((9 * CI + 31) / 32) * 32 * L.co;
        hipLaunchKernelGGL(k_relayout, dim3(nblk(nw)), dim3(256), 0, s, params.p + po, CI, L.co, wf.p + wo, wb.p + wo);
        hipLaunchKernelGGL(k_relayout32, dim3(nblk(n32)), dim3(256), 0, s, params.p + po, CI, L.co, wf32.p + w32);
        long nl = 9L * ((CI + 31) / 32) * 32 * L.co;
        hipLaunchKernelGGL(k_relayout_lds, dim3(nblk(nl)), dim3(256), 0, s, params.p + po, CI, L.co, wl.p + wlo);
        po += nw + L.co;
        wo += nw;
        w32 += n32;
        wlo += nl;
    }
}

int Denoiser::ensure_acts(mmhrt_ctx_s* c, int B, int Hp, int Wp, bool train) {
    size_t P0 = (size_t)B * Hp * Wp;
    size_t need = 0;
    // activation slots: per layer output + pooled tensors + input
    size_t res[4] = {P0, P0 / 4, P0 / 16, P0 / 64};
    int lvl[kNumLayers] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    for (int l = 0; l < kNumLayers; ++l) need += res[lvl[l]] * kLayers[l].co;
    need += P0 * 16 + res[1] * 32 + res[2] * 48 + res[3] * 64;  // input + pooled
    int e = 0;
    if ((e = mmhrt::ensure(c, acts, need, "denoiser activations"))) return e;
    if ((e = mmhrt::ensure(c, out, P0 * 4, "denoiser output"))) return e;
    if (train) {
        size_t g = P0 * (16 + 288 + 32) + (P0 / 4) * 48 + (P0 / 16) * 64;
        if ((e = mmhrt::ensure(c, gbuf, g, "denoiser gradient buffers"))) return e;
    }
    // carve activation slots
    size_t o = 0;
    x0 = acts.p + o;
    o += P0 * 16;
    for (int l = 0; l < kNumLayers; ++l) {
        a[l] = acts.p + o;
        o += res[lvl[l]] * kLayers[l].co;
    }
    pool[0] = acts.p + o;
    o += res[1] * 32;
    pool[1] = acts.p + o;
    o += res[2] * 48;
    pool[2] = acts.p + o;
    o += res[3] * 64;
    return 0;
}

// Forward pass over X0 [B,Hp,Wp,16] into `out` (fp32, channels 0-3 used).
void Denoiser::forward(hipStream_t s, int B, int Hp, int Wp) {
    int Hs[4] = {Hp, Hp / 2, Hp / 4, Hp / 8}, Ws[4] = {Wp, Wp / 2, Wp / 4, Wp / 8};
    size_t wo = 0, po = 0, w32 = 0, wlo = 0;
    auto conv = [&](int l, const uint16_t* x1, const uint16_t* x2, int lv, uint16_t* y, float* yf) {
        const LayerSpec& L = kLayers[l];
        ConvArgs A{};
        A.X1 = x1; A.X2 = x2; A.C1 = L.c1; A.C2 = L.c2; A.CO = L.co;
        A.B = B; A.H = Hs[lv]; A.W = Ws[lv];
        A.Wt = wf.p + wo; A.Wt32 = wf32.p + w32; A.Wl = wl.p + wlo;
        A.bias = params.p + po + 9L * (L.c1 + L.c2) * L.co;
        A.Y = y; A.Yf = yf; A.relu = L.relu;
        A.YfC = yf ? 4 : 0;  // the output layer: only channels 0-3 (visibility, colour residual) are used
        conv_fwd(s, A, k32, conv_shape);
        wo += 9L * (L.c1 + L.c2) * L.co;
        w32 += (size_t)((9 * (L.c1 + L.c2) + 31) / 32) * 32 * L.co;
        wlo += (size_t)9 * ((L.c1 + L.c2 + 31) / 32) * 32 * L.co;
        po += 9L * (L.c1 + L.c2) * L.co + L.co;
    };
    auto pl = [&](const uint16_t* x, uint16_t* y, int lv, int C) {
        long n = (long)B * (Hs[lv] / 2) * (Ws[lv] / 2) * (C / 8);
        hipLaunchKernelGGL(k_avgpool, dim3(nblk(n)), dim3(256), 0, s, x, y, B, Hs[lv], Ws[lv], C);
    };
    conv(0, x0, nullptr, 0, a[0], nullptr);
    conv(1, a[0], nullptr, 0, a[1], nullptr);
    pl(a[1], pool[0], 0, 32);
    conv(2, pool[0], nullptr, 1, a[2], nullptr);
    conv(3, a[2], nullptr, 1, a[3], nullptr);
    pl(a[3], pool[1], 1, 48);
    conv(4, pool[1], nullptr, 2, a[4], nullptr);
    conv(5, a[4], nullptr, 2, a[5], nullptr);
    pl(a[5], pool[2], 2, 64);
    conv(6, pool[2], nullptr, 3, a[6], nullptr);
    conv(7, a[6], nullptr, 3, a[7], nullptr);
    conv(8, a[5], a[7], 2, a[8], nullptr);
    conv(9, a[8], nullptr, 2, a[9], nullptr);
    conv(10, a[3], a[9], 1, a[10], nullptr);
    conv(11, a[10], nullptr, 1, a[11], nullptr);
    conv(12, a[1], a[11], 0, a[12], nullptr);
    conv(13, a[12], nullptr, 0, a[13], nullptr);
    conv(14, a[13], nullptr, 0, nullptr, out.p);
}

// Backward pass. dOut (bf16 [B,Hp,Wp,16]) must be in the first P0*16 halfs of gbuf; weight and
// bias gradients are accumulated into `grads` (zero it first).
void Denoiser::backward(hipStream_t s, int B, int Hp, int Wp, int cu) {
    int Hs[4] = {Hp, Hp / 2, Hp / 4, Hp / 8}, Ws[4] = {Wp, Wp / 2, Wp / 4, Wp / 8};
    size_t P0 = (size_t)B * Hp * Wp;
    const int lvl[kNumLayers] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    // gbuf: dOut | three rotating temporaries (96 channels at full res) | skip gradients
    uint16_t* dout = gbuf.p;
    uint16_t* tmp[3] = {gbuf.p + P0 * 16, gbuf.p + P0 * (16 + 96), gbuf.p + P0 * (16 + 192)};
    uint16_t* dskip[3];  // gradients of a[1] (lv0, 32), a[3] (lv1, 48), a[5] (lv2, 64)
    dskip[0] = gbuf.p + P0 * (16 + 288);
    dskip[1] = dskip[0] + P0 * 32;
    dskip[2] = dskip[1] + (P0 / 4) * 48;
    const uint16_t* in1[kNumLayers] = {x0, a[0], pool[0], a[2], pool[1], a[4], pool[2], a[6],
                                       a[5], a[8], a[3], a[10], a[1], a[12], a[13]};
    const uint16_t* in2[kNumLayers] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                       a[7], nullptr, a[9], nullptr, a[11], nullptr, nullptr};
    size_t woff[kNumLayers], poff[kNumLayers];
    {
        size_t wo = 0, po = 0;
        for (int l = 0; l < kNumLayers; ++l) {
            woff[l] = wo;
            poff[l] = po;
            wo += 9ull * (kLayers[l].c1 + kLayers[l].c2) * kLayers[l].co;
            po += 9ull * (kLayers[l].c1 + kLayers[l].c2) * kLayers[l].co + kLayers[l].co;
        }
    }
    auto free_tmp = [&](const uint16_t* x, const uint16_t* y) -> uint16_t* {
        for (int i = 0; i < 3; ++i)
            if (tmp[i] != x && tmp[i] != y) return tmp[i];
        return tmp[0];
    };
    const uint16_t* cur = dout;
    for (int l = kNumLayers - 1; l >= 0; --l) {
        const LayerSpec& L = kLayers[l];
        int lv = lvl[l];
        int H = Hs[lv], W = Ws[lv];
        int CI = L.c1 + L.c2;
        const uint16_t* Y = L.relu ? a[l] : nullptr;
        ConvWgtArgs wa{};
        wa.X1 = in1[l]; wa.X2 = in2[l]; wa.C1 = L.c1; wa.C2 = L.c2; wa.CO = L.co;
        wa.B = B; wa.H = H; wa.W = W; wa.dY = cur; wa.Y = Y; wa.relu = L.relu; wa.dW = grads.p + poff[l];
        conv_bwd_weight(s, wa, cu);
        long P = (long)B * H * W;
        hipLaunchKernelGGL(k_bias_grad, dim3((unsigned)std::min<long>(nblk(P), 256), L.co), dim3(256), 0, s, cur, Y,
                           L.relu, P, L.co, grads.p + poff[l] + 9ull * CI * L.co);
        if (l == 0) break;
        uint16_t* dx = free_tmp(cur, nullptr);
        ConvBwdArgs ba{};
        ba.dY = cur; ba.Y = Y; ba.relu = L.relu; ba.CI = CI; ba.CO = L.co; ba.B = B; ba.H = H; ba.W = W;
        ba.Wt = wb.p + woff[l]; ba.dX = dx;
        conv_bwd_data(s, ba);
        if (L.c2 > 0) {  // decoder: [skip | upsampled coarser]
            int si = (l == 8) ? 2 : (l == 10 ? 1 : 0);
            uint16_t* dco = free_tmp(cur, dx);
            long n1 = (long)B * H * W * L.c1, n2 = (long)B * (H / 2) * (W / 2) * L.c2;
            hipLaunchKernelGGL(k_concat_split, dim3(nblk(std::max(n1, n2))), dim3(256), 0, s, dx, B, H, W, L.c1,
                               L.c2, dskip[si], dco);
            cur = dco;
        } else if (l == 2 || l == 4 || l == 6) {  // input was the pooled skip: add the pooling adjoint
            int si = l / 2 - 1;
            int lp = lv - 1;
            long n = (long)B * Hs[lp] * Ws[lp] * L.c1;
            hipLaunchKernelGGL(k_avgpool_bwd, dim3(nblk(n)), dim3(256), 0, s, dx, dskip[si], B, Hs[lp], Ws[lp], L.c1, 1);
            cur = dskip[si];
        } else {
            cur = dx;
        }
    }
}

}  // namespace dn
}  // namespace mmhrt

// ====================================================================== C ABI
using mmhrt::dn::Denoiser;
using mmhrt::dn::kLayers;
using mmhrt::dn::kNumLayers;

namespace {
int dn_check(mmhrt_ctx c, bool need_ready) {
    if (!c) return MMHRT_E_INVALID;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return mmhrt::hip_fail(c, e, "hipSetDevice");
    if (!c->dn) c->dn = new Denoiser();
    if (need_ready && !c->dn->ready)
        return mmhrt::fail(c, MMHRT_E_STATE, "denoiser: call mmhrt_denoiser_init or mmhrt_denoiser_set_params first");
    return MMHRT_OK;
}
inline int ceil8(int v) { return (v + 7) & ~7; }

// Features (already in dn->raw) -> network input -> forward -> composed colour in dn->color.
// fb_col / fb_rgba8: compose straight into a rendered frame's framebuffer (B = 1) instead of the
// colour buffer
int dn_run(mmhrt_ctx c, int B, int H, int W, float depth_scale, const float bg[4], const float* target, float* loss,
           bool train, float4* fb_col = nullptr, uint32_t* fb_rgba8 = nullptr, bool packed = false,
           const __half* aux8 = nullptr) {
    Denoiser& d = *c->dn;
    int Hp = ceil8(H), Wp = ceil8(W);
    if (int e = d.ensure_acts(c, B, Hp, Wp, train)) return e;
    size_t P = (size_t)B * H * W, Pp = (size_t)B * Hp * Wp;
    if (!fb_col)
        if (int e = mmhrt::ensure(c, d.color, P * 3, "denoiser colour")) return e;
    hipStream_t s = c->stream;
    float inv_depth = depth_scale > 0.0f ? 1.0f / depth_scale : 1.0f;
    if (!packed)  // a rendered frame arrives already packed (k_frame_input)
        hipLaunchKernelGGL(mmhrt::dn::k_pack_input, dim3((unsigned)((Pp + 255) / 256)), dim3(256), 0, s, d.raw.p, B,
                           H, W, Hp, Wp, inv_depth, d.x0);
    d.k32 = c->opt_dn_k32 != 0;
    d.conv_shape = (int)c->opt_dn_conv_shape;
    d.forward(s, B, Hp, Wp);
    float4 bgc = make_float4(bg ? bg[0] : 0.f, bg ? bg[1] : 0.f, bg ? bg[2] : 0.f, bg ? bg[3] : 1.f);
    if (loss) mmhrt::launch_fill_u32(s, reinterpret_cast<uint32_t*>(d.scalar.p), 0u, 1);
    // a rendered frame's albedo and hit distance come straight from its aux buffer (no raw copy)
    const uint16_t* feat = aux8 ? reinterpret_cast<const uint16_t*>(aux8) : d.raw.p;
    mmhrt::dn::launch_compose(s, d.out.p, feat, aux8 ? 8 : 12, aux8 ? 0 : 3, aux8 ? 6 : 9, B, H, W, Hp, Wp, bgc,
                              target, 1.0f / (float)(P * 3), fb_col ? nullptr : d.color.p, train ? d.gbuf.p : nullptr,
                              loss ? d.scalar.p : nullptr, fb_col, fb_rgba8);
    return 0;
}
}  // namespace

extern "C" {

int32_t mmhrt_denoiser_init(mmhrt_ctx c, uint64_t seed) {
    if (int e = dn_check(c, false)) return e;
    Denoiser& d = *c->dn;
    if (int e = d.ensure_params(c)) return e;
    std::vector<float> h(mmhrt::dn::param_count());
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    size_t o = 0;
    for (int l = 0; l < kNumLayers; ++l) {
        const auto& L = kLayers[l];
        int CI = L.c1 + L.c2;
        int fan_in = 9 * (l == 0 ? 12 : CI);
        float sd = sqrtf(2.0f / (float)fan_in) * (l == kNumLayers - 1 ? 0.1f : 1.0f);
        for (int t = 0; t < 9; ++t)
            for (int ci = 0; ci < CI; ++ci)
                for (int co = 0; co < L.co; ++co) {
                    bool live = !(l == 0 && ci >= 12) && !(l == kNumLayers - 1 && co > 0);
                    h[o++] = live ? sd * nd(rng) : 0.0f;
                }
        for (int co = 0; co < L.co; ++co) h[o++] = (l == kNumLayers - 1 && co == 0) ? 0.5f : 0.0f;
    }
    return mmhrt_denoiser_set_params(c, h.data(), h.size());
}

int32_t mmhrt_denoiser_param_count(mmhrt_ctx c, uint64_t* n) {
    if (!n) return MMHRT_E_INVALID;
    (void)c;
    *n = mmhrt::dn::param_count();
    return MMHRT_OK;
}

int32_t mmhrt_denoiser_get_params(mmhrt_ctx c, float* out, uint64_t n) {
    if (int e = dn_check(c, true)) return e;
    if (!out || n != mmhrt::dn::param_count()) return mmhrt::fail(c, MMHRT_E_INVALID, "denoiser_get_params: bad size");
    hipError_t e = hipMemcpyAsync(out, c->dn->params.p, n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? MMHRT_OK : mmhrt::hip_fail(c, e, "denoiser_get_params");
}

int32_t mmhrt_denoiser_set_params(mmhrt_ctx c, const float* in, uint64_t n) {
    if (int e = dn_check(c, false)) return e;
    if (!in || n != mmhrt::dn::param_count()) return mmhrt::fail(c, MMHRT_E_INVALID, "denoiser_set_params: bad size");
    Denoiser& d = *c->dn;
    if (int e = d.ensure_params(c)) return e;
    hipError_t e = hipMemcpyAsync(d.params.p, in, n * 4, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return mmhrt::hip_fail(c, e, "denoiser_set_params");
    mmhrt::dn::zero(c->stream, d.adam_m.p, (long)n);
    mmhrt::dn::zero(c->stream, d.adam_v.p, (long)n);
    d.relayout(c->stream);
    d.step = 0;
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return mmhrt::hip_fail(c, e, "denoiser_set_params");
    d.ready = true;
    return MMHRT_OK;
}

int32_t mmhrt_denoise(mmhrt_ctx c, float depth_scale) {
    if (int e = dn_check(c, true)) return e;
    if (c->fb_w == 0 || !c->fb_aux.p || !c->last_aux)
        return mmhrt::fail(c, MMHRT_E_STATE, "denoise: render a frame with MMHRT_FLAG_AUX first");
    Denoiser& d = *c->dn;
    int H = (int)c->fb_h, W = (int)c->fb_w;
    const int Hp = ceil8(H), Wp = ceil8(W);
    if (int e = d.ensure_acts(c, 1, Hp, Wp, false)) return e;
    const float inv_depth = depth_scale > 0.0f ? 1.0f / depth_scale : 1.0f;
    hipLaunchKernelGGL(mmhrt::dn::k_frame_input, dim3((unsigned)(((size_t)Hp * Wp + 255) / 256)), dim3(256), 0,
                       c->stream, c->fb_col.p, c->fb_aux.p, H, W, Hp, Wp, inv_depth, nullptr, d.x0);
    if (int e = dn_run(c, 1, H, W, depth_scale, c->last_bg, nullptr, nullptr, false, c->fb_col.p, c->fb_rgba8.p, true,
                       c->fb_aux.p))
        return e;
    hipError_t he = hipGetLastError();
    return he == hipSuccess ? MMHRT_OK : mmhrt::hip_fail(c, he, "denoise launch");
}

int32_t mmhrt_denoise_features(mmhrt_ctx c, const uint16_t* feats, uint32_t b, uint32_t h, uint32_t w,
                               float depth_scale, const float bg[4], float* out_color) {
    if (int e = dn_check(c, true)) return e;
    if (!feats || !out_color || b == 0 || h == 0 || w == 0)
        return mmhrt::fail(c, MMHRT_E_INVALID, "denoise_features: bad arguments");
    Denoiser& d = *c->dn;
    size_t P = (size_t)b * h * w;
    if (int e = mmhrt::ensure(c, d.raw, P * 12, "denoiser features")) return e;
    hipError_t he = hipMemcpyAsync(d.raw.p, feats, P * 24, hipMemcpyHostToDevice, c->stream);
    if (he != hipSuccess) return mmhrt::hip_fail(c, he, "denoise_features upload");
    if (int e = dn_run(c, (int)b, (int)h, (int)w, depth_scale, bg, nullptr, nullptr, false)) return e;
    he = hipMemcpyAsync(out_color, d.color.p, P * 12, hipMemcpyDeviceToHost, c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    return he == hipSuccess ? MMHRT_OK : mmhrt::hip_fail(c, he, "denoise_features");
}

int32_t mmhrt_denoiser_train_step(mmhrt_ctx c, const uint16_t* feats, const float* target, uint32_t b, uint32_t h,
                                  uint32_t w, float depth_scale, const float bg[4], float lr, int32_t apply,
                                  float* loss_out, float* grads_out) {
    if (int e = dn_check(c, true)) return e;
    if (!feats || !target || b == 0 || h == 0 || w == 0)
        return mmhrt::fail(c, MMHRT_E_INVALID, "train_step: bad arguments");
    Denoiser& d = *c->dn;
    size_t P = (size_t)b * h * w;
    if (int e = mmhrt::ensure(c, d.raw, P * 12, "denoiser features")) return e;
    if (int e = mmhrt::ensure(c, d.target, P * 3, "denoiser target")) return e;
    hipStream_t s = c->stream;
    hipError_t he = hipMemcpyAsync(d.raw.p, feats, P * 24, hipMemcpyHostToDevice, s);
    if (he == hipSuccess) he = hipMemcpyAsync(d.target.p, target, P * 12, hipMemcpyHostToDevice, s);
    if (he != hipSuccess) return mmhrt::hip_fail(c, he, "train_step upload");
    if (int e = dn_run(c, (int)b, (int)h, (int)w, depth_scale, bg, d.target.p, loss_out ? d.scalar.p : d.scalar.p, true))
        return e;
    size_t n = mmhrt::dn::param_count();
    mmhrt::dn::zero(s, d.grads.p, (long)n);
    d.backward(s, (int)b, ceil8((int)h), ceil8((int)w), c->cu_count);
    if (grads_out) {
        he = hipMemcpyAsync(grads_out, d.grads.p, n * 4, hipMemcpyDeviceToHost, s);
        if (he != hipSuccess) return mmhrt::hip_fail(c, he, "train_step grads");
    }
    if (apply) {
        d.step++;
        const float b1 = 0.9f, b2 = 0.999f;
        float c1 = 1.0f / (1.0f - powf(b1, (float)d.step)), c2 = 1.0f / (1.0f - powf(b2, (float)d.step));
        mmhrt::dn::adam(s, d.params.p, d.grads.p, d.adam_m.p, d.adam_v.p, (long)n, lr, b1, b2, 1e-8f, c1, c2);
        d.relayout(s);
    }
    if (loss_out) {
        he = hipMemcpyAsync(loss_out, d.scalar.p, 4, hipMemcpyDeviceToHost, s);
        if (he != hipSuccess) return mmhrt::hip_fail(c, he, "train_step loss");
    }
    he = hipStreamSynchronize(s);
    return he == hipSuccess ? MMHRT_OK : mmhrt::hip_fail(c, he, "train_step");
}

}  // extern "C"

namespace mmhrt {
namespace dn {
void Denoiser::release_all() {
    mmhrt::release(params); mmhrt::release(grads); mmhrt::release(adam_m); mmhrt::release(adam_v);
    mmhrt::release(scalar); mmhrt::release(wf); mmhrt::release(wb); mmhrt::release(wf32); mmhrt::release(wl); mmhrt::release(acts); mmhrt::release(gbuf);
    mmhrt::release(raw); mmhrt::release(out); mmhrt::release(target); mmhrt::release(color);
}
}  // namespace dn
}  // namespace mmhrt
