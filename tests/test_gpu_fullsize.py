"""Parity at the benchmarked sizes (GPU): the 2D caster on 512^3 u16 at
1920x1080, the 64-pose batched DRR on 512x512x400 int16, and the 16-view
registration step on 256^3 float32 -- the exact workloads bench.py times.

Tolerances as in the reduced-size tests: RGBA8 within 1 LSB on >= 99.9 % of
pixels; DRR <= 2e-6 relative against the FP32 oracle in the same march mode
and <= 1e-4 against the FP64 oracle replaying the FP32 branch decisions; NCC
losses <= 1e-4; pose gradients <= 1e-3 relative against FP64 central
differences at h = 1e-6.  The unconstrained FP64 oracle (its own branch
decisions) is reported beside them.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from robartsvtk_amd import abi, configs

import helpers as H
from test_gpu_batch import _fd_pose_gradient, _plan_with_derivatives, _registration_scene, baked64, ncc64

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_c3_full_size(gpu_ctx):
    s = configs.scene_c3(scale=1.0)
    gmax = s.tf2d["gradmag_max"]
    vh = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    vh.compute_gradmag(0, gmax)
    g = H.gpu_2d(gpu_ctx, vh, s.tf2d, s.views[0], s.march, use_gradmag=True)
    vh.close()
    _say("c3: GPU render done; oracle FP32 / FP64")
    gm = H.gradmag_u8(s.volume, s.spacing, gmax)
    tf = dict(s.tf2d, gradmag=gm, gradmag_scale=float(np.float32(gmax) / np.float32(255.0)))
    vd = configs.view_dict(s.views[0])
    o32, _, _ = orc.cast2d(s.volume, s.spacing, tf, vd, s.march, "f32")
    o64, _, _ = orc.cast2d(s.volume, s.spacing, tf, vd, dict(s.march, f32_control=1), "f64")
    f32, m32 = H.lsb_stats(g, o32)
    f64, m64 = H.lsb_stats(g, o64)
    o64u, _, _ = orc.cast2d(s.volume, s.spacing, tf, vd, s.march, "f64")
    fu, mu = H.lsb_stats(g, o64u)
    exact = float(np.mean(np.all(g == o32, axis=-1)))
    _say(f"c3 1920x1080: within 1 LSB vs f32 {f32:.6f} (max {m32}, bit-exact {exact:.6f}); vs f64 with f32 "
         f"decisions {f64:.6f} (max {m64}); vs unconstrained f64 {fu:.6f} (max {mu})")
    assert f32 >= 0.999 and f64 >= 0.999
    assert g[..., 3].max() > 100


def test_c4_full_size_views_and_pose_gradient(gpu_ctx):
    import torch
    s = configs.scene_c4(scale=1.0, with_derivatives=True)
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    mp, dp = configs.march_params(s.march), configs.drr_params(s.drr)
    plan = abi.DrrPlan(gpu_ctx, vol, s.views, mp, dp, dview=s.extra["dview"])
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    plan.autotune(out, reps=1)
    plan.render(out)
    torch.cuda.synchronize()
    sizes = s.extra["sizes"]
    firsts = {}
    for v, n in enumerate(sizes):
        firsts.setdefault(n, v)
    assert sorted(firsts) == [256, 512, 768, 1024]
    for n, v in sorted(firsts.items()):
        view = s.views[v]
        g = plan.split(out)[v].cpu().numpy()
        single = torch.empty((view.height, view.width), dtype=torch.float32, device="cuda")
        abi.render_drr(gpu_ctx, vol, view, mp, dp, single)
        torch.cuda.synchronize()
        assert np.array_equal(single.cpu().numpy(), g), v
        vd = configs.view_dict(view)
        o32 = orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f32")
        o64 = orc.drr(s.volume, s.spacing, vd, dict(s.march, f32_control=1), s.drr, "f64")
        o64u = orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f64")
        e32, e64, eu = (float(H.rel_err(g, o).max()) for o in (o32, o64, o64u))
        _say(f"c4 view {v} ({n}^2): max rel err vs f32 {e32:.2e}, vs f64 (f32 decisions) {e64:.2e}, "
             f"vs unconstrained f64 {eu:.2e}")
        assert e32 <= 2e-6 and e64 <= 1e-4
    # pose gradient of the registration step for the two first 256^2 views,
    # whose targets are FP64 DRRs at poses moved by a few degrees / mm
    test_views = [v for v, n in enumerate(sizes) if n == 256][:2]
    rng = np.random.default_rng(44)
    tgt = (out + 0.01 * torch.randn(out.shape, device="cuda", generator=torch.Generator("cuda").manual_seed(4)))
    tg = {}
    c = s.extra["center"]
    for v in test_views:
        gt = s.extra["poses"][v] + np.concatenate([np.radians(rng.uniform(-3, 3, 3)), rng.uniform(-5, 5, 3)])
        vd = {"m": np.zeros(16, np.float32), "W": 256, "H": 256,
              "m64": baked64(s.extra["cameras"][v], gt, c, 256, 256, s.spacing)}
        t = orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f64")
        tg[v] = (t + rng.normal(0.0, 0.01, t.shape)).astype(np.float32)
        tgt[plan.offsets[v]:plan.offsets[v + 1]] = torch.from_numpy(tg[v].ravel()).cuda()
    loss = torch.empty(plan.n_views, dtype=torch.float32, device="cuda")
    grad = torch.empty(plan.n_views * 6, dtype=torch.float32, device="cuda")
    plan.capture(out, tgt, loss, grad)
    plan.register(out, tgt, loss, grad)
    torch.cuda.synchronize()
    loss, grad = loss.cpu().numpy(), grad.cpu().numpy().reshape(-1, 6)
    tgl = [tg.get(v) for v in range(plan.n_views)]
    for v in test_views:
        vd = {"m": np.zeros(16, np.float32), "W": 256, "H": 256,
              "m64": baked64(s.extra["cameras"][v], s.extra["poses"][v], c, 256, 256, s.spacing)}
        l64 = 1.0 - ncc64(orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f64"), tg[v])
        fd6 = _fd_pose_gradient(s, tgl, v, 1e-6)
        fd3 = _fd_pose_gradient(s, tgl, v, 1e-3)
        e6 = np.linalg.norm(grad[v] - fd6) / np.linalg.norm(fd6)
        e3r = np.linalg.norm(grad[v][:3] - fd3[:3]) / np.linalg.norm(fd3[:3])
        e3t = np.linalg.norm(grad[v][3:] - fd3[3:]) / np.linalg.norm(fd3[3:])
        _say(f"c4 view {v}: loss {loss[v]:.6f} (fp64 {l64:.6f}); pose grad rel err vs fd(1e-6) {e6:.2e}; "
             f"vs fd(1e-3): rot {e3r:.2e}, trans {e3t:.2e}")
        assert abs(loss[v] - l64) <= 1e-4 * max(1.0, abs(l64))
        assert e6 <= 1e-3
    plan.close()
    vol.close()


def _c5_full():
    return _registration_scene(n_views=16, scale=1.0)


def plan_views(s):
    """The float32 views the registration plan renders (rvk_posed_view)."""
    c, img = s.extra["center"], s.extra["img"]
    return [abi.posed_view(cam, s.extra["poses"][v], c, img, img, spacing=s.spacing)
            for v, cam in enumerate(s.extra["cameras"])]


def test_c5_full_size_loss_and_pose_gradient(gpu_ctx):
    import torch
    s, tg = _c5_full()
    _say("c5: targets ready")
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    plan = _plan_with_derivatives(gpu_ctx, s, vol)
    tgt = torch.from_numpy(np.concatenate([t.ravel() for t in tg])).cuda()
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    loss = torch.empty(plan.n_views, dtype=torch.float32, device="cuda")
    grad = torch.empty(plan.n_views * 6, dtype=torch.float32, device="cuda")
    plan.autotune(out, reps=1)
    plan.capture(out, tgt, loss, grad)
    plan.register(out, tgt, loss, grad)
    torch.cuda.synchronize()
    img = s.extra["img"]
    imgs = out.cpu().numpy()
    loss, grad = loss.cpu().numpy(), grad.cpu().numpy().reshape(-1, 6)
    worst, worst_u = 0.0, 0.0
    pv = plan_views(s)
    for v in range(plan.n_views):
        gi = imgs[plan.offsets[v]:plan.offsets[v + 1]]
        # the FP64 march on the same (float32) view matrix, replaying the FP32 decisions
        oc = orc.drr(s.volume, s.spacing, configs.view_dict(pv[v]), dict(s.march, f32_control=1), s.drr,
                     "f64")
        worst = max(worst, float(H.rel_err(gi, oc.ravel()).max()))
        # and the unconstrained FP64 march on the pose matrices formed in double
        vd = {"m": np.zeros(16, np.float32), "W": img, "H": img,
              "m64": baked64(s.extra["cameras"][v], s.extra["poses"][v], s.extra["center"], img, img, s.spacing)}
        I = orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f64")
        worst_u = max(worst_u, float(H.rel_err(gi, I.ravel()).max()))
        l64 = 1.0 - ncc64(I, tg[v])
        assert abs(loss[v] - l64) <= 1e-4 * max(1.0, abs(l64)), (v, loss[v], l64)
    _say(f"c5 16 views: max rel DRR err vs f64 (f32 decisions) {worst:.2e}; vs unconstrained f64 on double-precision "
         f"pose matrices {worst_u:.2e}")
    assert worst <= 1e-4
    for v in (0, 9):
        fd6 = _fd_pose_gradient(s, tg, v, 1e-6)
        fd3 = _fd_pose_gradient(s, tg, v, 1e-3)
        e6 = np.linalg.norm(grad[v] - fd6) / np.linalg.norm(fd6)
        e3r = np.linalg.norm(grad[v][:3] - fd3[:3]) / np.linalg.norm(fd3[:3])
        e3t = np.linalg.norm(grad[v][3:] - fd3[3:]) / np.linalg.norm(fd3[3:])
        _say(f"c5 view {v}: loss {loss[v]:.6f}; pose grad rel err vs fd(1e-6) {e6:.2e}; vs fd(1e-3): "
             f"rot {e3r:.2e}, trans {e3t:.2e}")
        assert e6 <= 1e-3
        assert e3t <= 1e-3
    plan.close()
    vol.close()


def test_c5_full_size_volume_gradient(gpu_ctx):
    import torch
    s, tg = _c5_full()
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    plan = _plan_with_derivatives(gpu_ctx, s, vol)
    tgt = torch.from_numpy(np.concatenate([t.ravel() for t in tg])).cuda()
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    loss = torch.empty(plan.n_views, dtype=torch.float32, device="cuda")
    dldi = torch.empty_like(out)
    plan.render(out)
    plan.ncc(out, tgt, loss, dldi)
    gvol = torch.zeros(s.volume.size, dtype=torch.float32, device="cuda")
    plan.backward(out, dldi, grad_volume=gvol)
    torch.cuda.synchronize()
    _say("c5: volume gradient done; FP64 directional difference")
    g = gvol.cpu().numpy().reshape(s.volume.shape)
    img = s.extra["img"]
    vds = [{"m": np.zeros(16, np.float32), "W": img, "H": img,
            "m64": baked64(s.extra["cameras"][v], s.extra["poses"][v], s.extra["center"], img, img, s.spacing)}
           for v in range(plan.n_views)]

    def total_loss(mu):
        return sum(1.0 - ncc64(orc.drr(mu, s.spacing, vds[v], s.march, s.drr, "f64"), tg[v])
                   for v in range(plan.n_views))

    # a random direction over the body (mu > 0 in every voxel's trilinear neighbourhood)
    rng = np.random.default_rng(19)
    body = s.volume > 0.015
    r = np.where(body, rng.standard_normal(s.volume.shape), 0.0).astype(np.float32)
    eps = 1e-3
    fd = (total_loss((s.volume + eps * r).astype(np.float32)) -
          total_loss((s.volume - eps * r).astype(np.float32))) / (2 * eps)
    an = float((g.astype(np.float64) * r).sum())
    _say(f"c5 full-size volume gradient, directional derivative over the body: analytic {an:.6e}, fd {fd:.6e}")
    assert abs(an - fd) <= 1e-3 * abs(fd)
    plan.close()
    vol.close()
