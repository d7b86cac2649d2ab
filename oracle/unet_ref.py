"""CPU reference of the denoiser network — TEST INFRASTRUCTURE ONLY (the checker of the BF16
MFMA kernels in megamol_amd/csrc/denoise.hip; the product never imports it).

Restates, with PyTorch on the CPU (fp32 or fp64), the network that include/mmhrt.h documents:
input packing of the 12-value feature records, the 15 convolutions of the 4-level U-Net
(3x3, padding 1, ReLU except the last), 2x2 average pooling, nearest 2x upsampling with
[skip, upsampled] concatenation, the albedo * v + residual composition and the mean squared error.
There is no reference counterpart (MegaMol has no denoiser); finite differences in float64 pin
the autograd gradients of this restatement.
"""
from __future__ import annotations

import numpy as np

LAYERS = [("enc0a", 16, 0, 32, 1), ("enc0b", 32, 0, 32, 1), ("enc1a", 32, 0, 48, 1), ("enc1b", 48, 0, 48, 1),
          ("enc2a", 48, 0, 64, 1), ("enc2b", 64, 0, 64, 1), ("bota", 64, 0, 96, 1), ("botb", 96, 0, 96, 1),
          ("dec2a", 64, 96, 64, 1), ("dec2b", 64, 0, 64, 1), ("dec1a", 48, 64, 48, 1), ("dec1b", 48, 0, 48, 1),
          ("dec0a", 32, 48, 32, 1), ("dec0b", 32, 0, 32, 1), ("out", 32, 0, 16, 0)]


def bf16_round(x, grad=True):
    """Round a float32 tensor to bfloat16 precision (nearest even), keeping float32. With grad,
    the gradient flowing back through this point is rounded the same way: the GPU stores both
    activations and activation gradients in bf16 between layers."""
    import torch

    class _Round(torch.autograd.Function):
        @staticmethod
        def forward(ctx, v):
            return v.to(torch.bfloat16).to(v.dtype)

        @staticmethod
        def backward(ctx, g):
            return g.to(torch.bfloat16).to(g.dtype) if grad else g

    return _Round.apply(x)


def bf16_round_grad_only(x):
    """Identity forward; bf16-rounded gradient backward (the network output stays fp32 on the GPU
    but its gradient is stored in bf16)."""
    import torch

    class _G(torch.autograd.Function):
        @staticmethod
        def forward(ctx, v):
            return v.clone()

        @staticmethod
        def backward(ctx, g):
            return g.to(torch.bfloat16).to(g.dtype)

    return _G.apply(x)


def pack(feats: np.ndarray, depth_scale: float, dtype="float32", round_bf16=False):
    """Feature records (B,H,W,12 f16) -> network input (B,16,Hp,Wp), Hp/Wp rounded up to 8."""
    import torch
    f = feats.astype(np.float32)
    B, H, W, _ = f.shape
    Hp, Wp = (H + 7) // 8 * 8, (W + 7) // 8 * 8
    x = np.zeros((B, Hp, Wp, 16), np.float32)
    t = f[..., 9]
    hit = t > 0
    inv = np.float32(1.0) / np.float32(depth_scale) if depth_scale > 0 else np.float32(1.0)
    x[:, :H, :W, 0] = f[..., 10]
    x[:, :H, :W, 1:4] = f[..., 3:6]
    x[:, :H, :W, 4:7] = f[..., 6:9]
    x[:, :H, :W, 7] = np.where(hit, t * inv, 0).astype(np.float32)
    x[:, :H, :W, 8:11] = f[..., 0:3]
    x[:, :H, :W, 11] = hit.astype(np.float32)
    tx = torch.from_numpy(x).permute(0, 3, 1, 2).contiguous()
    if round_bf16:
        tx = bf16_round(tx)
    return tx.to(getattr(torch, dtype))


def torch_params(params: np.ndarray, dtype="float32", requires_grad=False):
    """Canonical flat parameters -> list of (weight [co, ci, 3, 3], bias [co]) torch tensors."""
    import torch
    out, o = [], 0
    for _, c1, c2, co, _ in LAYERS:
        ci = c1 + c2
        w = params[o:o + 9 * ci * co].reshape(3, 3, ci, co)
        o += 9 * ci * co
        b = params[o:o + co]
        o += co
        tw = torch.tensor(np.ascontiguousarray(w.transpose(3, 2, 0, 1)), dtype=getattr(torch, dtype),
                          requires_grad=requires_grad)
        tb = torch.tensor(b, dtype=getattr(torch, dtype), requires_grad=requires_grad)
        out.append((tw, tb))
    return out


def flatten_grads(tp) -> np.ndarray:
    """Gradients of torch_params back to the canonical flat order W[t][ci][co], b[co]."""
    parts = []
    for tw, tb in tp:
        g = tw.grad.detach().double().numpy()  # [co, ci, 3, 3]
        parts.append(g.transpose(2, 3, 1, 0).reshape(-1))
        parts.append(tb.grad.detach().double().numpy())
    return np.concatenate(parts)


def forward(tp, x0, round_bf16=False):
    import torch
    import torch.nn.functional as F

    def conv(i, x, skip=None):
        if skip is not None:
            x = torch.cat([skip, F.interpolate(x, scale_factor=2, mode="nearest")], 1)
            if round_bf16:
                # the GPU keeps the gradient of the concatenated input in bf16 before splitting it
                # into the skip part and the 2x2 sums of the upsampled part
                x = bf16_round_grad_only(x)
        w, b = tp[i]
        if round_bf16:
            # the GPU multiplies bf16 copies of the fp32 master weights; their gradient is taken
            # with respect to the master weights (straight through the rounding)
            w = bf16_round(w, grad=False)
        y = F.conv2d(x, w, b, padding=1)
        if LAYERS[i][4]:
            y = F.relu(y)
        if round_bf16 and i < len(LAYERS) - 1:
            y = bf16_round(y)
        return y

    def pool(x):
        y = F.avg_pool2d(x, 2)
        return bf16_round(y) if round_bf16 else y

    a0 = conv(0, x0)
    a1 = conv(1, a0)
    a2 = conv(2, pool(a1))
    a3 = conv(3, a2)
    a4 = conv(4, pool(a3))
    a5 = conv(5, a4)
    a6 = conv(6, pool(a5))
    a7 = conv(7, a6)
    a8 = conv(8, a7, skip=a5)
    a9 = conv(9, a8)
    a10 = conv(10, a9, skip=a3)
    a11 = conv(11, a10)
    a12 = conv(12, a11, skip=a1)
    a13 = conv(13, a12)
    out = conv(14, a13)
    return bf16_round_grad_only(out) if round_bf16 else out


def compose(out, feats: np.ndarray, bg=(0, 0, 0, 1)):
    """out (B,16,Hp,Wp) -> colour (B,H,W,3): albedo * v + residual (channels 1-3) on hit pixels,
    background elsewhere."""
    import torch
    B, H, W, _ = feats.shape
    f = torch.from_numpy(feats.astype(np.float32)).to(out.dtype)
    v = out[:, 0, :H, :W]
    res = out[:, 1:4, :H, :W].permute(0, 2, 3, 1)
    hit = (f[..., 9] > 0)
    alb = f[..., 3:6]
    bgt = torch.tensor(bg[:3], dtype=out.dtype).expand(B, H, W, 3)
    return torch.where(hit[..., None], alb * v[..., None] + res, bgt)


def loss_and_grads(params, feats, target, depth_scale, dtype="float64", round_bf16=False):
    import torch
    tp = torch_params(params, dtype, requires_grad=True)
    x0 = pack(feats, depth_scale, dtype, round_bf16)
    col = compose(forward(tp, x0, round_bf16), feats)
    loss = ((col - torch.from_numpy(target).to(col.dtype)) ** 2).mean()
    loss.backward()
    return float(loss.detach()), flatten_grads(tp), col.detach().double().numpy()


def denoise(params, feats, depth_scale, dtype="float32", round_bf16=True):
    import torch
    with torch.no_grad():
        tp = torch_params(params, dtype)
        return compose(forward(tp, pack(feats, depth_scale, dtype, round_bf16), round_bf16), feats).double().numpy()
