"""The warm-start U-Net's HIP kernels (BF16 matrix-core convolutions, FP32
accumulation) against the float64 numpy reference oracle/unet_ref.py, whose
own gradients are pinned by finite differences in test_unet_ref.py.

BF16 activations carry about three significant digits, so forward outputs
and gradients are compared by relative L2 error and direction (cosine)
rather than bit for bit; the weights are first rounded to BF16 so that the
remaining differences come from activation rounding alone."""
import os
import sys

import numpy as np
import pytest

import artss_amd
import cases

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import unet_ref as U  # noqa: E402

pytestmark = pytest.mark.gpu


def _bf16(a):
    b = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def _flat(p):
    return np.concatenate([np.concatenate([p[n + ".W"].reshape(-1), p[n + ".b"]]) for n, *_ in U.LAYERS])


def _unflat(flat):
    p, off = {}, 0
    for n, ci, co, _ in U.LAYERS:
        p[n + ".W"] = flat[off:off + 27 * ci * co].reshape(27, ci, co)
        off += 27 * ci * co
        p[n + ".b"] = flat[off:off + co]
        off += co
    return p


@pytest.fixture(scope="module")
def setup():
    dom = artss_amd.Domain(cases.ns_taylor_green(n=16, levels=2, steps=1))
    rng = np.random.default_rng(5)
    p = U.init_params(seed=6)
    for n, *_ in U.LAYERS:
        p[n + ".W"] = _bf16(p[n + ".W"])
        p[n + ".b"] = 0.05 * rng.standard_normal(p[n + ".b"].shape)
    x = rng.standard_normal((16, 16, 16, 2))
    target = rng.standard_normal((16, 16, 16, 1))
    net = artss_amd.UNet(dom, seed=1)
    assert net.param_count() == _flat(p).size
    net.set_params(_flat(p))
    return dom, net, p, x, target


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_matches_reference(setup):
    dom, net, p, x, _ = setup
    xb = artss_amd.Buffer(dom, x.astype(np.float32))
    yb = artss_amd.Buffer(dom, np.zeros((16, 16, 16, 1), np.float32))
    net.forward(xb, yb)
    y = yb.download().astype(np.float64)
    y_ref, _ = U.forward(p, x)
    err = _rel(y, y_ref)
    print("forward relative L2 error", err)
    assert err < 3e-2


def test_forward_matches_reference_wide(setup):
    """288 x 32 x 16 voxels: a wavefront row of the forward convolution
    covers 128, 64 and 32 voxels at 16, 32 and 64 output channels, so every
    level runs several workgroups along x with a partial last tile"""
    dom, net, p, _, _ = setup
    x = np.random.default_rng(11).standard_normal((16, 32, 288, 2))
    xb = artss_amd.Buffer(dom, x.astype(np.float32))
    yb = artss_amd.Buffer(dom, np.zeros((16, 32, 288, 1), np.float32))
    net.forward(xb, yb)
    y = yb.download().astype(np.float64)
    y_ref, _ = U.forward(p, x)
    err = _rel(y, y_ref)
    print("wide forward relative L2 error", err)
    assert err < 3e-2


def test_gradients_match_reference(setup):
    dom, net, p, x, target = setup
    xb = artss_amd.Buffer(dom, x.astype(np.float32))
    tb = artss_amd.Buffer(dom, target.astype(np.float32))
    loss, g = net.gradients(xb, tb)
    loss_ref, g_ref = U.loss_and_grad(p, x, target)
    print("loss", loss, "reference", loss_ref)
    assert abs(loss - loss_ref) <= 3e-2 * loss_ref
    gg = _unflat(g.astype(np.float64))
    for n, *_ in U.LAYERS:
        for key in (n + ".W", n + ".b"):
            a, b = gg[key].reshape(-1), g_ref[key].reshape(-1)
            cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
            err = _rel(a, b)
            print(key, "rel err %.3g cos %.5f" % (err, cos))
            assert cos > 0.995 and err < 0.1, key


def test_training_reduces_loss(setup):
    dom, net, p, x, _ = setup
    # learn a fixed linear smoothing of the first input channel
    smooth = x[..., :1].copy()
    for ax in range(3):
        smooth = 0.25 * np.roll(smooth, 1, ax) + 0.5 * smooth + 0.25 * np.roll(smooth, -1, ax)
    xb = artss_amd.Buffer(dom, x.astype(np.float32))
    tb = artss_amd.Buffer(dom, smooth.astype(np.float32))
    net.set_params(_flat(U.init_params(seed=7)))
    first, _ = net.gradients(xb, tb, want_grads=False)
    loss = first
    for _ in range(150):
        net.adam_step(2e-3)
        loss, _ = net.gradients(xb, tb, want_grads=False)
    print("training loss", first, "->", loss)
    assert loss < 0.2 * first
