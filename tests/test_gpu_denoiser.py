"""GPU tests of the BF16 MFMA denoiser against the CPU reference network (oracle/unet_ref.py).

Weights and activations are bf16 on the GPU, so the comparisons are against a float32 reference
that rounds weights, activations and activation gradients to bf16 at the same points: relative
L2 error of the output below 1e-2; per layer, weight gradients no further (x1.2) from the exact
float64 gradients than that reference is, and within 2e-2 (cosine above 0.9998) of it. Only the
fp32 accumulation order differs (the LDS-staged forward sums each 32-channel slice tap by tap,
the reference sums over tap-major k), which moves individual activations to the neighbouring
bf16 value; through the backward pass of the coarse levels that adds up to about 1% per layer.
"""
import numpy as np
import pytest

from helpers import case
from megamol_amd import denoiser as D

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from oracle import unet_ref as R  # noqa: E402
from test_denoiser_cpu import random_features  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def dnz(gpu_renderer):
    d = D.Denoiser(gpu_renderer, seed=7)
    d.p0 = d.get_params()  # every test perturbs these, so its parameters do not depend on test order
    return d


@pytest.mark.parametrize("k32", [0, 1, 14, 24, 34, 2, 4, 5, 6, 7])
@pytest.mark.parametrize("shape", [(1, 16, 16), (2, 24, 40), (1, 37, 70)])
def test_forward_matches_reference(dnz, shape, k32):
    """k32: 0 = 16x16x16 kernels, 1 = 16x16x32 with the automatic launch shape, 5 = the LDS-staged
    kernel (96-channel layers as two groups of 48 output channels), 7 = the LDS-staged
    kernel without channel groups, 6 = the LDS-DMA double-buffered kernel for single-slice layers
    (the LDS-staged one for the others), other values force the 16x16x32 launch shape (channel
    blocks * 10 + pixel blocks per wavefront)."""
    case(f"denoiser forward {shape} k32={k32}")
    dnz.r.set_option("dn_k32", 1 if k32 else 0)
    dnz.r.set_option("dn_conv_shape", k32 if k32 > 1 else 0)
    b, h, w = shape
    feats = random_features(b, h, w, seed=h)
    p = dnz.p0
    g = np.random.default_rng(1)
    p = (p + g.normal(size=p.shape).astype(np.float32) * 0.02).astype(np.float32)  # non-trivial biases too
    dnz.set_params(p)
    out = dnz.denoise(feats, depth_scale=2.0)
    ref = R.denoise(p, feats, 2.0, round_bf16=True)
    dnz.r.set_option("dn_k32", 1)
    dnz.r.set_option("dn_conv_shape", 0)
    assert out.shape == ref.shape
    print(f"forward rel. L2 error {rel(out, ref):.2e}")
    assert rel(out, ref) < 1e-2, rel(out, ref)


def test_gradients_match_reference(dnz):
    case("denoiser gradients")
    b, h, w = 2, 32, 32
    feats = random_features(b, h, w, seed=5)
    target = np.random.default_rng(2).random((b, h, w, 3)).astype(np.float32)
    p = dnz.p0
    p = (p + np.random.default_rng(3).normal(size=p.shape).astype(np.float32) * 0.02).astype(np.float32)
    dnz.set_params(p)
    loss, g = dnz.train_step(feats, target, 2.0, lr=0.0, apply=False, want_grads=True)
    rloss, rg, _ = R.loss_and_grads(p, feats, target, 2.0, "float32", round_bf16=True)
    _, eg, _ = R.loss_and_grads(p, feats, target, 2.0, "float64", round_bf16=False)
    assert abs(loss - rloss) <= 2e-2 * rloss
    o = 0
    rows = []
    for name, c1, c2, co, _ in D.LAYERS:
        n = 9 * (c1 + c2) * co + co
        gg, rr, ee = g[o:o + n], rg[o:o + n], eg[o:o + n]
        o += n
        if np.linalg.norm(rr) == 0:
            continue
        g64 = gg.astype(np.float64)
        cos = float(np.sum(g64 * rr) / (np.sqrt(np.sum(g64 * g64)) * np.sqrt(np.sum(rr * rr))))
        nb = 9 * (c1 + c2) * co  # bias gradients follow the weights
        rows.append((name, rel(gg, ee), rel(rr, ee), rel(gg, rr), cos, rel(gg[nb:], ee[nb:]), rel(rr[nb:], ee[nb:])))
    for r in rows:
        print("%-6s gpu-exact %.4f bf16ref-exact %.4f gpu-bf16ref %.4f cos %.5f | bias gpu %.4f ref %.4f" % r)
    for name, err_gpu, err_bf16, err_pair, cos, _, _ in rows:
        # bf16 weights and bf16 storage of activations and their gradients move the deepest
        # layers' weight gradients by about 3% from the exact float64 ones; the reference rounds
        # at the same points, so the GPU must be no further from exact than it, and close to it
        # up to accumulation-order flips of bf16 activations (about 1% at the coarse levels)
        assert err_gpu < max(1.2 * err_bf16, 1e-3), (name, err_gpu, err_bf16)
        assert err_pair < 2e-2 and cos > 0.9998, (name, err_pair, cos)
    assert np.array_equal(dnz.get_params(), p), "apply=0 must not update the parameters"


def test_frame_path_equals_feature_path(gpu_renderer):
    """mmhrt_denoise on a rendered frame (its features packed into the network input on the device
    in one pass, colours composed into the framebuffer) gives the same colours, bit for bit, as
    mmhrt_denoise_features on that frame's downloaded colour and features."""
    from helpers import small_camera
    from megamol_amd import _lib as L, params, scenes
    case("denoiser frame path vs feature path")
    dn = D.Denoiser(gpu_renderer, seed=3)
    xyzr, rgba = scenes.random_spheres(3000, seed=5, rmin=0.01, rmax=0.04)
    w, h = 77, 45  # not multiples of 8: the padded input border is exercised
    cam = small_camera(w, h)
    bg = (0.1, 0.2, 0.3, 1.0)
    p = params(w, h, spp=1, mode=L.MODE_AO, ao_distance=0.2, flags=L.FLAG_AUX, background=bg)
    gpu_renderer.set_spheres(xyzr, rgba)
    gpu_renderer.build()
    fr = gpu_renderer.render(cam, p)
    aux = gpu_renderer.download(w, h, outputs=(), aux=True).aux
    feats = np.zeros((h, w, 12), np.float16)
    feats[..., 0:3] = fr.color[..., :3].astype(np.float16)
    feats[..., 3:11] = aux
    ref = dn.denoise(feats, 1.5, bg=bg)[0]
    dn.denoise_frame(1.5)
    got = gpu_renderer.download(w, h, outputs=("color",)).color[..., :3]
    np.testing.assert_array_equal(got, ref)


def test_training_reduces_loss_and_set_get_roundtrip(dnz, tmp_path):
    case("denoiser training smoke")
    dnz.set_params(D.Denoiser(dnz.r, seed=11).get_params())
    feats = random_features(4, 32, 32, seed=9)
    # a learnable target: colour = albedo * visibility of a smooth field
    f = feats.astype(np.float32)
    yy, xx = np.mgrid[0:32, 0:32] / 32.0
    v = 0.5 + 0.4 * np.sin(6 * xx) * np.cos(5 * yy)
    target = np.where(f[..., 9:10] > 0, f[..., 3:6] * v[None, ..., None], 0).astype(np.float32)
    losses = [dnz.train_step(feats, target, 2.0, lr=2e-3) for _ in range(60)]
    assert losses[-1] < 0.5 * losses[0], (losses[0], losses[-1])
    path = tmp_path / "w.bin"
    dnz.save(str(path), {"loss": losses[-1]})
    p = dnz.get_params()
    dnz.set_params(np.zeros_like(p))
    meta = dnz.load(str(path))
    assert meta["loss"] == losses[-1]
    np.testing.assert_array_equal(dnz.get_params(), p)


def test_trained_network_gains_over_the_method_without_it(gpu_renderer):
    """The committed weights on held-out crops of the 50M-sphere FCC scene (a scene seed and
    cameras never used in training): the network against the raw 1-spp frame and against the
    non-learned alternative, a 5x5 box filter of the 1-spp visibility times the albedo, all
    compared with 256-spp frames of the same crops."""
    import os
    from megamol_amd import train_denoiser as T
    from megamol_amd.denoiser import psnr
    case("trained network vs box filter on held-out FCC crops")
    path = os.path.join(os.path.dirname(D.__file__), "weights", "denoiser_fcc.bin")
    dn = D.Denoiser(gpu_renderer, seed=None)
    dn.load(path)
    T.load_fcc(gpu_renderer, 232, seed=31)
    feats, targets = T.make_pairs(gpu_renderer, 12, 232, seed=31, crop=128, spp_target=256)
    den = dn.denoise(feats, 1.0)
    raw = feats[..., 0:3].astype(np.float32)
    box = np.stack([D.box_filter_baseline(f[..., 3:11], (0.0, 0.0, 0.0)) for f in feats])
    p_den, p_raw, p_box = psnr(den, targets), psnr(raw, targets), psnr(box, targets)
    print(f"held-out PSNR vs 256 spp: network {p_den:.2f} dB, box filter {p_box:.2f} dB, raw 1 spp {p_raw:.2f} dB")
    assert p_den > p_box + 2.0
    assert p_den > p_raw + 6.0
