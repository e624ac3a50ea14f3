"""The fused optimiser step against torch.optim.Adam and the reference's
VectorAdam formula, and the HIP-graph-captured training iteration
(forward + backward + optimiser, seed read from device memory) against the
same iteration run eagerly."""
import numpy as np
import pytest
import torch

from diffsoup_amd import scenes

pytestmark = pytest.mark.gpu


def test_fused_adam_matches_torch_and_vector_adam(dev):
    from diffsoup_amd.optimize import FusedAdam, VectorAdam
    g = torch.Generator().manual_seed(0)
    shapes = [(3001, 7), (13,), (1000, 3)]   # crosses the 4096-element block boundary; odd sizes
    ps = [torch.randn(*s, generator=g) for s in shapes]
    ours = [p.clone().to(dev) for p in ps]
    grads = [torch.zeros_like(p) for p in ours]
    opt = FusedAdam([dict(param=ours[0], grad=grads[0], lr=5e-2),
                     dict(param=ours[1], grad=grads[1], lr=1e-2, betas=(0.8, 0.99)),
                     dict(param=ours[2], grad=grads[2], lr=1e-3, vector_dim=3)])
    ref = [p.clone().to(dev).requires_grad_(True) for p in ps]
    ta = torch.optim.Adam([{"params": [ref[0]], "lr": 5e-2}, {"params": [ref[1]], "lr": 1e-2, "betas": (0.8, 0.99)}],
                          foreach=False)
    va = VectorAdam([ref[2]], lr=1e-3)
    for it in range(4):
        gs = [torch.randn(*s, generator=g).to(dev) * (10.0 ** (it - 2)) for s in shapes]
        before = [r.detach().clone() for r in ref]
        for gr, x in zip(grads, gs):
            gr.copy_(x)
        for r, x in zip(ref, gs):
            r.grad = x.clone()
        opt.step()
        ta.step()
        va.step()
        torch.cuda.synchronize()
        for o, r, b in zip(ours, ref, before):
            du, dr = (o - b).cpu().numpy(), (r.detach() - b).cpu().numpy()
            np.testing.assert_allclose(du, dr, rtol=1e-4, atol=1e-4 * np.abs(dr).max())
    assert int(opt.step_count.item()) == 4

    # out of place: same values in param_out, param untouched
    p_in = [o.clone() for o in ours]
    p_out = [torch.empty_like(o) for o in ours]
    opt2 = FusedAdam([dict(param=p_in[0], grad=grads[0], lr=5e-2, param_out=p_out[0]),
                      dict(param=p_in[1], grad=grads[1], lr=1e-2, betas=(0.8, 0.99), param_out=p_out[1]),
                      dict(param=p_in[2], grad=grads[2], lr=1e-3, vector_dim=3, param_out=p_out[2])])
    opt3 = FusedAdam([dict(param=ours[0], grad=grads[0], lr=5e-2),
                      dict(param=ours[1], grad=grads[1], lr=1e-2, betas=(0.8, 0.99)),
                      dict(param=ours[2], grad=grads[2], lr=1e-3, vector_dim=3)])
    snapshot = [o.clone() for o in ours]
    opt2.step()
    opt3.step()
    torch.cuda.synchronize()
    for a, b, c, d in zip(p_out, ours, p_in, snapshot):
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
        np.testing.assert_array_equal(c.cpu().numpy(), d.cpu().numpy())

    # the bf16 copy written beside the update is torch's round-to-nearest-even
    p16 = [torch.empty(o.shape, dtype=torch.bfloat16, device=dev) for o in ours]
    opt4 = FusedAdam([dict(param=ours[0], grad=grads[0], lr=5e-2, param_bf16_out=p16[0]),
                      dict(param=ours[1], grad=grads[1], lr=1e-2, param_bf16_out=p16[1]),
                      dict(param=ours[2], grad=grads[2], lr=1e-3, vector_dim=3, param_bf16_out=p16[2])])
    opt4.step()
    torch.cuda.synchronize()
    for a, b in zip(p16, ours):
        np.testing.assert_array_equal(a.view(torch.int16).cpu().numpy(), b.to(torch.bfloat16).view(torch.int16).cpu().numpy())


@pytest.mark.parametrize("n", [(1 << 20) + 3, (20 << 20) + 5, (130 << 20) + 7])
def test_fused_adam_launch_shapes(dev, n):
    """The update at the three launch shapes (up to 16M units, up to 128M,
    above: one, two and four float4s per thread), against torch.optim.Adam
    over two steps; the odd sizes take the scalar tail."""
    from diffsoup_amd.optimize import FusedAdam
    g = torch.Generator(device=dev).manual_seed(1)
    p0 = torch.randn(n, generator=g, device=dev)
    ours, grad = p0.clone(), torch.empty_like(p0)
    opt = FusedAdam([dict(param=ours, grad=grad, lr=5e-2)])
    ref = p0.clone().requires_grad_(True)
    ta = torch.optim.Adam([ref], lr=5e-2, foreach=False)
    for it in range(2):
        x = torch.randn(n, generator=g, device=dev)
        grad.copy_(x)
        ref.grad = x
        before = ref.detach().clone()
        opt.step()
        ta.step()
        torch.cuda.synchronize()
        du, dr = ours - before, ref.detach() - before
        assert float((du - dr).abs().max()) <= 1e-4 * float(dr.abs().max()), f"step {it}"
        del before, du, dr
    assert int(opt.step_count.item()) == 2


def _snap(p):
    return {k: getattr(p, k).detach().clone() for k in ("verts", "alpha_logits", "feat", "w1", "b3")}


def test_captured_step_matches_eager(dev):
    from diffsoup_amd import fused
    s = scenes.sphere_torus_soup(6, n_surface=3000, n_float=800, views=2, res=(144, 176))
    mvp = torch.from_numpy(s.mvp).to(dev)
    target = torch.from_numpy(s.target).to(dev)
    mk = lambda: fused.FusedStep(s.resolutions, s.T, s.verts.shape[0], s.alpha_level, s.levels, 8, 32, device=dev)
    s0 = 4242

    # eager: step + optimiser, seeds s0, s0 + 1, s0 + 2
    pe = fused.SoupParams.from_soup(s, dev)
    pe0 = fused.SoupParams.from_soup(s, dev)
    se = mk()
    oe = fused.soup_optimizer(se, pe)
    eager_rast, eager_params, eager_loss = [], [], []
    for i in range(3):
        se.run(pe, mvp, target, seed=s0 + i)
        eager_rast.append(se.rast.clone())
        eager_loss.append(se.loss.clone())
        oe.step()
        eager_params.append(_snap(pe))
    torch.cuda.synchronize()

    # graph: the same three iterations as replays, with eager work between them
    pg = fused.SoupParams.from_soup(s, dev)
    sg = mk()
    og = fused.soup_optimizer(sg, pg)
    cap = fused.CapturedStep(sg, pg, mvp, target, og, seed=s0)
    check = mk()
    for i in range(3):
        before = fused.SoupParams(*[t.clone() for t in (pg.verts, pg.tri, pg.alpha_logits, pg.feat, pg.w1, pg.b1,
                                                         pg.w2, pg.b2, pg.w3, pg.b3)])
        loss = cap.replay()
        torch.cuda.synchronize()
        # the replay used seed s0 + i on the parameters it started from
        check.run(before, mvp, target, seed=s0 + i)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(sg.rast.cpu().numpy().view(np.uint32), check.rast.cpu().numpy().view(np.uint32))
        if i == 0:
            np.testing.assert_array_equal(sg.rast.cpu().numpy().view(np.uint32),
                                          eager_rast[0].cpu().numpy().view(np.uint32))
        assert abs(float(loss.item()) - float(eager_loss[i].item())) <= 1e-4 * float(eager_loss[i].item())
        # parameters follow the eager run.  Atomics reorder fp32 sums, and a
        # gradient that cancels to ~0 in one order and not the other moves
        # its element differently under Adam's normalisation, so a handful
        # of elements may differ; a missing, doubled or mis-scaled update
        # moves nearly all of them.
        for k, v in _snap(pg).items():
            ref = eager_params[i][k]
            moved = (ref - eager_params[i - 1][k]).abs() if i else (ref - _snap(pe0)[k]).abs()
            off = (v - ref).abs() > 1e-5 * max(1.0, float(ref.abs().max()))
            assert float(off.float().mean()) < 1e-3, f"replay {i} {k}: {int(off.sum())} of {off.numel()} differ"
            assert float(moved.max()) > 0, f"eager {k} did not move"
        _ = torch.randn(4096, device=dev).sum().item()   # eager work between replays
    assert int(og.step_count.item()) == 3


def test_captured_l1_ssim_step_matches_eager(dev):
    """The L1 + SSIM step (colour pass, SSIM kernels, adjoint pass) inside a
    captured graph: each replay, with eager work between replays, gives the
    raster, loss and gradients of an eager step on the same parameters and
    seed."""
    from diffsoup_amd import fused
    s = scenes.sphere_torus_soup(7, n_surface=3000, n_float=800, views=2, res=(144, 176))
    mvp = torch.from_numpy(s.mvp).to(dev)
    target = torch.from_numpy(s.target).to(dev)
    mk = lambda: fused.FusedStep(s.resolutions, s.T, s.verts.shape[0], s.alpha_level, s.levels, 8, 32, device=dev,
                                 loss="l1_ssim")
    pg = fused.SoupParams.from_soup(s, dev)
    sg = mk()
    cap = fused.CapturedStep(sg, pg, mvp, target, None, seed=77)
    check = mk()
    for i in range(3):
        loss = cap.replay()
        torch.cuda.synchronize()
        check.run(pg, mvp, target, seed=77 + i)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(sg.rast.cpu().numpy().view(np.uint32), check.rast.cpu().numpy().view(np.uint32))
        assert abs(float(loss.item()) - float(check.loss.item())) <= 1e-6 * float(check.loss.item())
        for name in ("grad_color", "grad_feat", "grad_mlp", "grad_verts", "grad_alpha_logits"):
            a, b = getattr(sg, name).cpu().numpy(), getattr(check, name).cpu().numpy()
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-5 * max(1e-12, np.abs(b).max()), err_msg=name)
        _ = torch.randn(4096, device=dev).sum().item()   # eager work between replays
