"""torch.autograd wrapper of the batched DRR (robartsvtk_amd.autograd), GPU.

The wrapper's gradients come from rvk_drr_plan_backward; these tests check
the plumbing end to end against independent references: the fused
registration step's pose gradient, and FP64 central differences of the CPU
oracle for a loss written in torch.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from robartsvtk_amd import abi, configs

import helpers as H
from test_gpu_batch import _fd_pose_gradient, _plan_with_derivatives, _registration_scene, baked64, ncc64

pytestmark = pytest.mark.gpu


def _renderer(s, volume=None):
    from robartsvtk_amd.autograd import DrrRenderer
    img = s.extra["img"]
    return DrrRenderer(s.volume if volume is None else volume, s.spacing, s.extra["cameras"],
                       [(img, img)] * len(s.extra["cameras"]), s.extra["center"], configs.march_params(s.march),
                       configs.drr_params(s.drr))


def _ncc_loss(torch, images, targets, offsets):
    total = 0.0
    for v in range(len(offsets) - 1):
        a = images[offsets[v]:offsets[v + 1]].double()
        b = targets[offsets[v]:offsets[v + 1]].double()
        a = a - a.mean()
        b = b - b.mean()
        total = total + 1.0 - (a * b).sum() / torch.sqrt((a * a).sum() * (b * b).sum())
    return total


def test_autograd_pose_gradient(gpu_ctx):
    import torch
    from robartsvtk_amd.autograd import drr_render
    s, tg = _registration_scene(n_views=2)
    r = _renderer(s)
    poses = torch.tensor(np.asarray(s.extra["poses"]), dtype=torch.float64, requires_grad=True)
    images = drr_render(r, poses)
    tgt = torch.from_numpy(np.concatenate([t.ravel() for t in tg])).cuda()
    loss = _ncc_loss(torch, images, tgt, r.plan.offsets)
    loss.backward()
    g = poses.grad.numpy()
    # the fused registration step of an independent plan
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    plan = _plan_with_derivatives(gpu_ctx, s, vol)
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    lv = torch.empty(plan.n_views, dtype=torch.float32, device="cuda")
    gr = torch.empty(plan.n_views * 6, dtype=torch.float32, device="cuda")
    plan.register(out, tgt, lv, gr)
    torch.cuda.synchronize()
    assert torch.equal(out, images.detach())
    gr = gr.cpu().numpy().reshape(-1, 6)
    assert abs(float(loss) - float(lv.sum())) <= 1e-5
    np.testing.assert_allclose(g, gr, rtol=2e-4, atol=1e-6 * np.abs(gr).max())
    # and FP64 central differences (h = 1e-6, see test_gpu_batch) for view 0
    fd6 = _fd_pose_gradient(s, tg, 0, 1e-6)
    e = np.linalg.norm(g[0] - fd6) / np.linalg.norm(fd6)
    print(f"autograd pose grad view 0 {g[0]}, fd(1e-6) {fd6}, rel err {e:.2e}")
    assert e <= 1e-3
    # a second forward at new poses reuses the plan (views updated in place)
    poses2 = (poses.detach() + 0.01).requires_grad_(True)
    images2 = drr_render(r, poses2)
    _ncc_loss(torch, images2, tgt, r.plan.offsets).backward()
    assert not torch.equal(images2, images.detach()) and torch.isfinite(poses2.grad).all()
    plan.close()
    r.close()


def test_autograd_volume_gradient_vs_fp64_fd(gpu_ctx):
    import torch
    from robartsvtk_amd.autograd import drr_render
    s, _ = _registration_scene(n_views=2, scale=0.1)
    mu = torch.from_numpy(s.volume).cuda().requires_grad_(True)
    r = _renderer(s, volume=mu)
    poses = torch.tensor(np.asarray(s.extra["poses"]), dtype=torch.float64)
    images = drr_render(r, poses, mu)
    rng = np.random.default_rng(3)
    w = torch.from_numpy(rng.standard_normal(r.plan.total).astype(np.float32)).cuda()
    loss = (images.double() * w.double()).sum()
    loss.backward()
    g = mu.grad.cpu().numpy().astype(np.float64)
    img = s.extra["img"]
    wv = w.cpu().numpy().astype(np.float64)
    vds = [{"m": np.zeros(16, np.float32), "W": img, "H": img,
            "m64": baked64(s.extra["cameras"][v], s.extra["poses"][v], s.extra["center"], img, img, s.spacing)}
           for v in range(r.n_views)]

    def total(vol):
        return sum(float(orc.drr(vol, s.spacing, vds[v], s.march, s.drr, "f64").ravel() @
                         wv[r.plan.offsets[v]:r.plan.offsets[v + 1]]) for v in range(r.n_views))

    body = s.volume > 0.015  # mu > 0 in the whole trilinear neighbourhood: no kinks of the gate
    d = np.where(body, rng.standard_normal(s.volume.shape), 0.0).astype(np.float32)
    eps = 1e-3
    fd = (total((s.volume + eps * d).astype(np.float32)) - total((s.volume - eps * d).astype(np.float32))) / (2 * eps)
    an = float((g * d).sum())
    print(f"autograd volume directional derivative: analytic {an:.6e}, fd {fd:.6e}")
    assert abs(an - fd) <= 1e-3 * abs(fd)
    r.close()


def test_autograd_two_forwards_one_backward(gpu_ctx):
    """Two renders in one loss share the renderer's plan: each backward must
    differentiate the poses and volume of its own forward, not the last."""
    import torch
    from robartsvtk_amd.autograd import drr_render
    s, tg = _registration_scene(n_views=2, scale=0.1)
    mu0 = torch.from_numpy(s.volume).cuda()
    r = _renderer(s)
    tgt = torch.from_numpy(np.concatenate([t.ravel() for t in tg])).cuda()
    p1 = np.asarray(s.extra["poses"], np.float64)
    p2 = p1 + np.array([0.01, -0.02, 0.015, 1.0, -0.5, 0.7])
    mus = [mu0, (mu0 * 1.1).contiguous()]
    rng = np.random.default_rng(17)
    w = torch.from_numpy(rng.standard_normal(r.plan.total).astype(np.float32)).cuda().double()

    def one(p, mu, scale):
        x = torch.tensor(p, dtype=torch.float64, requires_grad=True)
        m = mu.clone().requires_grad_(True)
        img = drr_render(r, x, m)
        loss = scale * (_ncc_loss(torch, img, tgt, r.plan.offsets) + (img.double() * w).sum())
        return x, m, loss

    # separately: forward then backward, one render at a time
    sep = []
    for p, mu, sc in ((p1, mus[0], 1.0), (p2, mus[1], 2.0)):
        x, m, loss = one(p, mu, sc)
        loss.backward()
        sep.append((x.grad.clone(), m.grad.clone()))
    # together: both forwards first, then one backward through both
    x1, m1, l1 = one(p1, mus[0], 1.0)
    x2, m2, l2 = one(p2, mus[1], 2.0)
    (l1 + l2).backward()
    torch.cuda.synchronize()
    for (gx, gm), x, m in zip(sep, (x1, x2), (m1, m2)):
        torch.testing.assert_close(x.grad, gx, rtol=1e-6, atol=0.0)
        torch.testing.assert_close(m.grad, gm, rtol=1e-6, atol=0.0)
    assert not torch.allclose(x1.grad, x2.grad)
    r.close()
