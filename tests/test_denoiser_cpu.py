"""CPU tests of the denoiser's reference (the checker) and of the weights file format."""
import numpy as np
import pytest

from megamol_amd import denoiser as D

torch = pytest.importorskip("torch")
from oracle import unet_ref as R  # noqa: E402


def random_features(b, h, w, seed=0):
    g = np.random.default_rng(seed)
    f = np.zeros((b, h, w, 12), np.float16)
    hit = g.random((b, h, w)) < 0.8
    f[..., 0:3] = g.random((b, h, w, 3))
    f[..., 3:6] = g.random((b, h, w, 3))
    n = g.normal(size=(b, h, w, 3))
    f[..., 6:9] = n / np.linalg.norm(n, axis=-1, keepdims=True)
    f[..., 9] = np.where(hit, g.uniform(1, 3, (b, h, w)), 0)
    f[..., 10] = np.where(hit, g.integers(0, 2, (b, h, w)), 0)
    f[~hit, 3:9] = 0
    return f


def test_layer_table_matches_reference():
    assert D.LAYERS == R.LAYERS
    assert D.param_count() == sum(9 * (c1 + c2) * co + co for _, c1, c2, co, _ in R.LAYERS)
    # the network has about half a million parameters
    assert 400_000 < D.param_count() < 700_000


def test_reference_gradients_match_finite_differences():
    g = np.random.default_rng(3)
    p = (g.normal(size=D.param_count()) * 0.05).astype(np.float64)
    feats = random_features(1, 8, 8, seed=1)
    target = g.random((1, 8, 8, 3))
    loss, grads, _ = R.loss_and_grads(p.astype(np.float32).astype(np.float64), feats, target, 2.0, "float64")
    # central differences on a sample of parameters (including the first and last layer)
    idx = np.concatenate([g.choice(D.param_count(), 20, replace=False), [0, 5, D.param_count() - 16 - 1]])
    import torch as T

    def loss_at(q):
        tp = R.torch_params(q, "float64")
        col = R.compose(R.forward(tp, R.pack(feats, 2.0, "float64")), feats)
        return float(((col - T.from_numpy(target)) ** 2).mean())

    base = p.astype(np.float32).astype(np.float64)
    for i in idx:
        eps = 1e-6
        qp, qm = base.copy(), base.copy()
        qp[i] += eps
        qm[i] -= eps
        fd = (loss_at(qp) - loss_at(qm)) / (2 * eps)
        assert abs(fd - grads[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, grads[i])


def test_reference_conv_is_cross_correlation_with_tap_layout():
    # a single non-zero weight W[t][ci][co] must shift the input by the tap offset o_t = (ky-1, kx-1)
    import torch.nn.functional as F
    x = torch.zeros(1, 16, 5, 5)
    x[0, 3, 2, 2] = 1.0
    for t in range(9):
        p = np.zeros(D.param_count(), np.float32)
        w = p[:9 * 16 * 32].reshape(9, 16, 32)
        w[t, 3, 7] = 1.0
        tw, tb = R.torch_params(p)[0]
        y = F.conv2d(x, tw, tb, padding=1)[0, 7]
        ky, kx = divmod(t, 3)
        # y[p] = x[p + o_t]  ->  the impulse at (2,2) appears at (2 - (ky-1), 2 - (kx-1))
        assert y[2 - (ky - 1), 2 - (kx - 1)] == 1.0 and float(y.abs().sum()) == 1.0


def test_weights_file_roundtrip_and_validation(tmp_path):
    p = np.random.default_rng(0).normal(size=D.param_count()).astype(np.float32)
    f = tmp_path / "w.bin"
    D.save_weights(str(f), p, {"steps": 12})
    q, meta = D.load_weights(str(f))
    np.testing.assert_array_equal(p, q)
    assert meta == {"steps": 12}
    raw = bytearray(f.read_bytes())
    raw[0] = ord("X")
    (tmp_path / "bad.bin").write_bytes(bytes(raw))
    with pytest.raises(ValueError):
        D.load_weights(str(tmp_path / "bad.bin"))


def test_psnr_and_ssim_closed_forms():
    a = np.zeros((16, 16, 3))
    b = a + 0.1
    assert abs(D.psnr(a, b) - 20.0) < 1e-9  # mse 0.01 -> 20 dB
    assert D.ssim(a + 0.5, a + 0.5) == pytest.approx(1.0)


def test_box_filter_baseline_closed_forms():
    from megamol_amd.denoiser import box_filter_baseline
    aux = np.zeros((9, 9, 8), np.float16)
    aux[..., 0:3] = np.array([0.2, 0.4, 0.6], np.float16)
    aux[..., 6] = 1.0  # every pixel hit
    aux[..., 7] = 1.0
    aux[4, 4, 7] = 0.0  # one occluded sample in the middle
    out = box_filter_baseline(aux, (0.1, 0.1, 0.1), k=3)
    # the middle 3x3 window holds 8 visible samples of 9; windows that miss it keep v = 1
    np.testing.assert_allclose(out[4, 4], np.array([0.2, 0.4, 0.6]) * (8 / 9), rtol=2e-3)
    np.testing.assert_allclose(out[0, 0], np.array([0.2, 0.4, 0.6]), rtol=2e-3)
    # misses take the background and do not count in their neighbours' averages
    aux[0, :, 6] = 0.0
    out = box_filter_baseline(aux, (0.1, 0.1, 0.1), k=3)
    np.testing.assert_allclose(out[0, 3], [0.1, 0.1, 0.1])
    np.testing.assert_allclose(out[1, 0], np.array([0.2, 0.4, 0.6]), rtol=2e-3)
