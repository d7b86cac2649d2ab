// denoise.hpp — U-Net denoiser state and kernel argument blocks.
#pragma once
#include "ctx.hpp"

namespace mmhrt {
namespace dn {

constexpr int kNumLayers = 15;

struct LayerSpec {
    const char* name;
    int c1;    // input channels at the layer's resolution
    int c2;    // input channels at half resolution (nearest-upsampled), 0 if none
    int co;    // output channels
    int relu;
};
extern const LayerSpec kLayers[kNumLayers];
size_t param_count();
void launch_compose(hipStream_t s, const float* out, const uint16_t* feat, int fstride, int fa, int ft, int B, int H,
                    int W, int Hp, int Wp, float4 bg, const float* target, float inv_count, float* color,
                    uint16_t* dout, float* loss, float4* fb_col, uint32_t* fb_rgba8);
void zero(hipStream_t s, float* p, long n);
void adam(hipStream_t s, float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
          float c1, float c2);

struct ConvArgs {
    const uint16_t* X1;
    const uint16_t* X2;
    int C1, C2, CO;
    int B, H, W;
    const uint16_t* Wt;    // [tap][ci/16][co/16][co%16][ci%16] (16x16x16 kernel)
    const uint16_t* Wt32;  // [k/32][co/16][lane][8], k = tap*CI + ci zero-padded (16x16x32 kernel)
    const uint16_t* Wl;    // [tap][ci/32][co/16][lane][8], ci zero-padded (LDS-staged kernel)
    const float* bias;
    uint16_t* Y;
    float* Yf;
    int relu;
    int YfC;  // channels written to Yf, which is also its pixel stride (0: all CO)
};
struct ConvBwdArgs {
    const uint16_t* dY;
    const uint16_t* Y;
    int relu;
    int CI, CO;
    int B, H, W;
    const uint16_t* Wt;
    uint16_t* dX;
};
struct ConvWgtArgs {
    const uint16_t* X1;
    const uint16_t* X2;
    int C1, C2, CO;
    int B, H, W;
    const uint16_t* dY;
    const uint16_t* Y;
    int relu;
    float* dW;
};

struct Denoiser {
    bool ready = false;
    int64_t step = 0;
    DevBuf<float> params, grads, adam_m, adam_v, scalar;
    DevBuf<uint16_t> wf, wb;   // bf16 weights in forward / backward fragment order
    DevBuf<uint16_t> wf32;     // bf16 weights in the 16x16x32 forward fragment order
    DevBuf<uint16_t> wl;       // bf16 weights in the LDS-staged forward fragment order
    bool k32 = true;           // forward with v_mfma_f32_16x16x32_bf16
    int conv_shape = 0;        // 0: automatic launch shape of the k32 kernel (see conv_fwd)
    DevBuf<uint16_t> acts;     // activations
    DevBuf<uint16_t> gbuf;     // backward temporaries
    DevBuf<uint16_t> raw;      // raw per-pixel features (12 halfs)
    DevBuf<float> out;         // network output, fp32 [B,Hp,Wp,16]
    DevBuf<float> target;      // training target colour
    DevBuf<float> color;       // composed colour [B,H,W,3]
    uint16_t* x0 = nullptr;
    uint16_t* a[kNumLayers] = {};
    uint16_t* pool[3] = {};

    int ensure_params(mmhrt_ctx_s* c);
    int ensure_acts(mmhrt_ctx_s* c, int B, int Hp, int Wp, bool train);
    void relayout(hipStream_t s);
    void forward(hipStream_t s, int B, int Hp, int Wp);
    void backward(hipStream_t s, int B, int Hp, int Wp, int cu);
    void release_all();
};

}  // namespace dn
}  // namespace mmhrt
