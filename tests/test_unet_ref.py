"""The numpy reference of the warm-start U-Net (oracle/unet_ref.py) pinned by
mathematics: a brute-force convolution at single voxels, the adjoint
identities of down/up-sampling, and central finite differences of the loss
against its analytic gradient for entries of every parameter tensor."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import unet_ref as U  # noqa: E402


def test_conv_matches_brute_force():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 5, 6, 3))
    W = rng.standard_normal((27, 3, 2))
    b = rng.standard_normal(2)
    y = U.conv(x, W, b)
    for (k, j, i) in [(0, 0, 0), (3, 4, 5), (1, 2, 3), (2, 0, 5)]:
        want = b.copy()
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    kk, jj, ii = k + dz, j + dy, i + dx
                    if 0 <= kk < 4 and 0 <= jj < 5 and 0 <= ii < 6:
                        want += x[kk, jj, ii] @ W[9 * (dz + 1) + 3 * (dy + 1) + (dx + 1)]
        assert np.allclose(y[k, j, i], want, rtol=1e-13, atol=1e-13)


def test_sampling_adjoints():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((4, 6, 8, 3))
    c = rng.standard_normal((2, 3, 4, 3))
    assert np.isclose((U.down(a) * c).sum(), (a * U.down_backward(c)).sum())
    assert np.isclose((U.up(c) * a).sum(), (c * U.up_backward(a)).sum())


@pytest.mark.parametrize("name", [n for n, *_ in U.LAYERS])
def test_gradient_matches_finite_differences(name):
    rng = np.random.default_rng(3)
    p = U.init_params(seed=4)
    for k in p:
        if k.endswith(".b"):
            p[k] = 0.05 * rng.standard_normal(p[k].shape)
    x = rng.standard_normal((8, 8, 8, 2))
    target = rng.standard_normal((8, 8, 8, 1))
    _, g = U.loss_and_grad(p, x, target)
    for key in (name + ".W", name + ".b"):
        flat = p[key].reshape(-1)
        for idx in rng.choice(flat.size, size=min(3, flat.size), replace=False):
            h = 1e-6
            old = flat[idx]
            flat[idx] = old + h
            lp, _ = U.loss_and_grad(p, x, target)
            flat[idx] = old - h
            lm, _ = U.loss_and_grad(p, x, target)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            an = g[key].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(fd)), (key, idx, fd, an)
