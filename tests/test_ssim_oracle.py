"""Pins of the oracle's SSIM restatement (pytorch_msssim's algorithm, float64)
against what the mathematics fixes: SSIM(x, x) = 1, the closed form on
constant images, the separable valid filter against scipy.ndimage, a
brute-force 2-D window sum (also with a side shorter than the window), the
window's closed form, and torch autograd against central differences.
CPU only."""
import numpy as np
import pytest
import torch

from oracle import oracle as O


def test_window_closed_form():
    g = O.gaussian_window_1d(11, 1.5).double().numpy()
    c = np.arange(11) - 5
    ref = np.exp(-c ** 2 / 4.5)
    np.testing.assert_allclose(g, ref / ref.sum(), rtol=1e-6)
    np.testing.assert_array_equal(g, g[::-1])


def test_identical_images_and_constant_closed_form():
    x = torch.rand(2, 3, 23, 17, dtype=torch.float64, generator=torch.Generator().manual_seed(0))
    assert abs(float(O.ssim_fp64(x, x)) - 1.0) < 1e-12
    # on constant images every variance vanishes and S = (2ab + C1) / (a^2 +
    # b^2 + C1), for a window normalised exactly (the float32 window of
    # pytorch_msssim sums to 1 only to an ulp, which moves S by ~1e-5 here)
    a, b = 0.3, 0.7
    X = torch.full((1, 2, 15, 15), a, dtype=torch.float64)
    Y = torch.full((1, 2, 15, 15), b, dtype=torch.float64)
    g = O.gaussian_window_1d().double()
    c1 = 0.01 ** 2
    np.testing.assert_allclose(float(O.ssim_fp64(X, Y, win=g / g.sum())), (2 * a * b + c1) / (a * a + b * b + c1),
                               rtol=1e-12)


def test_filter_matches_scipy_correlate():
    from scipy import ndimage
    rng = np.random.default_rng(1)
    x = rng.uniform(size=(1, 1, 30, 41))
    g = O.gaussian_window_1d().double().numpy()
    got = O.gaussian_filter_valid(torch.from_numpy(x), O.gaussian_window_1d()).numpy()[0, 0]
    ref = ndimage.correlate1d(ndimage.correlate1d(x[0, 0], g, axis=0, mode="constant"), g, axis=1, mode="constant")
    np.testing.assert_allclose(got, ref[5:-5, 5:-5], rtol=1e-12, atol=1e-14)


def _brute_ssim(x, y, kh, kw, c1=1e-4, c2=9e-4):
    """Direct 2-D window sums per window (one channel)."""
    g = O.gaussian_window_1d().double().numpy()
    gy = g if kh == 11 else np.ones(1)
    gx = g if kw == 11 else np.ones(1)
    H, W = x.shape
    out = []
    for i in range(H - kh + 1):
        for j in range(W - kw + 1):
            w = np.outer(gy, gx)
            px, py = x[i:i + kh, j:j + kw], y[i:i + kh, j:j + kw]
            mx, my = (w * px).sum(), (w * py).sum()
            sxx = (w * px * px).sum() - mx * mx
            syy = (w * py * py).sum() - my * my
            sxy = (w * px * py).sum() - mx * my
            out.append((2 * mx * my + c1) / (mx * mx + my * my + c1) * (2 * sxy + c2) / (sxx + syy + c2))
    return np.mean(out)


@pytest.mark.parametrize("shape", [(13, 12), (7, 16), (14, 5)])
def test_brute_force_windows(shape):
    rng = np.random.default_rng(2)
    x, y = rng.uniform(size=shape), rng.uniform(size=shape)
    kh = 11 if shape[0] >= 11 else 1
    kw = 11 if shape[1] >= 11 else 1
    got = float(O.ssim_fp64(torch.from_numpy(x)[None, None], torch.from_numpy(y)[None, None]))
    np.testing.assert_allclose(got, _brute_ssim(x, y, kh, kw), rtol=1e-12)


def test_loss_gradient_against_central_differences():
    rng = np.random.default_rng(3)
    c = rng.uniform(size=(1, 12, 13, 3))
    t = rng.uniform(size=(1, 12, 13, 3))
    c[0, 4, 4, 1] = t[0, 4, 4, 1] + 0.3   # keep the probed entries off the L1 kink
    _, grad = O.l1_ssim_loss_fp64(c, t)
    h = 1e-6
    for idx in [(0, 0, 0, 0), (0, 4, 4, 1), (0, 11, 12, 2), (0, 6, 2, 0)]:
        cp, cm = c.copy(), c.copy()
        cp[idx] += h
        cm[idx] -= h
        fd = (O.l1_ssim_loss_fp64(cp, t)[0] - O.l1_ssim_loss_fp64(cm, t)[0]) / (2 * h)
        np.testing.assert_allclose(grad[idx], fd, rtol=1e-5, atol=1e-10)
