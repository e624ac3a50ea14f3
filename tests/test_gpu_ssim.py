"""SSIM and the reference's L1 + SSIM loss (examples/02_synthetic.py:376-380)
on the HIP kernels, against the oracle's float64 restatement of
pytorch_msssim.ssim with gradients from torch autograd.

Tolerances (fp32 kernels against fp64): SSIM values within 2e-6 absolute
(variances are differences of fp32 moments, ~1e-7 absolute, against C2 =
9e-4); gradients within 1e-3 of the largest gradient magnitude."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

VAL_TOL = 2e-6
GRAD_TOL = 1e-3


def _images(shape, seed, anti=False):
    g = torch.Generator().manual_seed(seed)
    x = torch.rand(shape, generator=g, dtype=torch.float64)
    y = torch.rand(shape, generator=g, dtype=torch.float64)
    if anti:
        y = (1 - x) * 0.8 + 0.2 * y
    return x, y


@pytest.mark.parametrize("shape", [(2, 3, 64, 80), (1, 3, 11, 11), (1, 3, 7, 30), (3, 1, 100, 37), (1, 4, 150, 203)])
def test_ssim_values_and_gradients(dev, shape):
    from diffsoup_amd import losses
    print(f"ssim shape {shape}", flush=True)
    x, y = _images(shape, 1)
    yg = y.float().to(dev).requires_grad_(True)
    s = losses.ssim(x.float().to(dev), yg, data_range=1.0)
    s.backward()
    y64 = y.clone().requires_grad_(True)
    ref = O.ssim_fp64(x, y64, data_range=1.0)
    ref.backward()
    assert abs(float(s.detach()) - float(ref.detach())) <= VAL_TOL
    r = y64.grad.numpy()
    np.testing.assert_allclose(yg.grad.cpu().numpy(), r, rtol=0, atol=GRAD_TOL * np.abs(r).max())


def test_ssim_per_image_nonnegative_and_first_image_gradient(dev):
    from diffsoup_amd import losses
    x, y = _images((3, 3, 40, 50), 2, anti=True)
    x[1] = y[1] * 0.9 + 0.05   # one image pair is similar, the other two are anti-correlated
    xg = x.float().to(dev).requires_grad_(True)
    yg = y.float().to(dev).requires_grad_(True)
    s = losses.ssim(xg, yg, data_range=1.0, size_average=False, nonnegative_ssim=True)
    (s * torch.tensor([1.0, 2.0, 3.0], device=dev)).sum().backward()
    x64, y64 = x.clone().requires_grad_(True), y.clone().requires_grad_(True)
    ref = O.ssim_fp64(x64, y64, data_range=1.0, size_average=False, nonnegative_ssim=True)
    (ref * torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64)).sum().backward()
    np.testing.assert_allclose(s.detach().cpu().numpy(), ref.detach().numpy(), rtol=0, atol=VAL_TOL)
    assert (ref.detach().numpy() == 0).sum() >= 1   # the relu clamps at least one image
    for got, r in ((xg.grad, x64.grad), (yg.grad, y64.grad)):
        r = r.numpy()
        np.testing.assert_allclose(got.cpu().numpy(), r, rtol=0, atol=GRAD_TOL * np.abs(r).max())


def test_ssim_reads_permuted_channels_last_images(dev):
    """The reference's call form: ssim(gt.permute(0,3,1,2), color.permute(0,3,1,2))."""
    from diffsoup_amd import losses
    x, y = _images((2, 48, 70, 3), 3)
    a = y.float().to(dev).requires_grad_(True)
    s1 = losses.ssim(x.float().to(dev).permute(0, 3, 1, 2), a.permute(0, 3, 1, 2), data_range=1.0)
    s1.backward()
    b = y.float().to(dev).permute(0, 3, 1, 2).contiguous().requires_grad_(True)
    s2 = losses.ssim(x.float().to(dev).permute(0, 3, 1, 2).contiguous(), b, data_range=1.0)
    s2.backward()
    assert float(s1.detach()) == float(s2.detach())
    np.testing.assert_array_equal(a.grad.cpu().numpy(), b.grad.permute(0, 2, 3, 1).cpu().numpy())


def test_identical_images(dev):
    from diffsoup_amd import losses
    x, _ = _images((1, 3, 33, 45), 4)
    y = x.float().to(dev).requires_grad_(True)
    s = losses.ssim(x.float().to(dev), y, data_range=1.0)
    s.backward()
    assert abs(float(s.detach()) - 1.0) < 1e-6
    assert float(y.grad.abs().max()) < 1e-4   # SSIM is stationary at Y = X


def test_l1_ssim_loss(dev):
    from diffsoup_amd import losses
    x, y = _images((2, 256, 256, 3), 5)
    y[0, :10] = x[0, :10]      # exact ties: the L1 term contributes sign(0) = 0 there
    c = y.float().to(dev).requires_grad_(True)
    loss = losses.l1_ssim_loss(c, x.float().to(dev))
    loss.backward()
    ref_loss, ref_grad = O.l1_ssim_loss_fp64(y.float().double().numpy(), x.float().double().numpy())
    assert abs(float(loss) - ref_loss) <= 1e-6
    np.testing.assert_allclose(c.grad.cpu().numpy(), ref_grad, rtol=0, atol=GRAD_TOL * np.abs(ref_grad).max())


def test_batches_beyond_one_launch(dev):
    """More images than one launch takes (64): the batch runs in chunks and
    matches the oracle, per image and through the loss."""
    from diffsoup_amd import losses
    x, y = _images((70, 3, 12, 14), 6)
    yg = y.float().to(dev).requires_grad_(True)
    s = losses.ssim(x.float().to(dev), yg, data_range=1.0, size_average=False)
    (s * torch.linspace(0.5, 1.5, 70, device=dev)).sum().backward()
    y64 = y.clone().requires_grad_(True)
    ref = O.ssim_fp64(x, y64, data_range=1.0, size_average=False)
    (ref * torch.linspace(0.5, 1.5, 70, dtype=torch.float64)).sum().backward()
    np.testing.assert_allclose(s.detach().cpu().numpy(), ref.detach().numpy(), rtol=0, atol=VAL_TOL)
    r = y64.grad.numpy()
    np.testing.assert_allclose(yg.grad.cpu().numpy(), r, rtol=0, atol=GRAD_TOL * np.abs(r).max())
    xc, yc = _images((66, 9, 13, 3), 7)
    c = yc.float().to(dev).requires_grad_(True)
    loss = losses.l1_ssim_loss(c, xc.float().to(dev))
    loss.backward()
    ref_loss, ref_grad = O.l1_ssim_loss_fp64(yc.float().double().numpy(), xc.float().double().numpy())
    assert abs(float(loss.detach()) - ref_loss) <= 1e-6
    np.testing.assert_allclose(c.grad.cpu().numpy(), ref_grad, rtol=0, atol=GRAD_TOL * np.abs(ref_grad).max())
