"""The registration loop on the device (GPU): the reference's uncentred NCC
beside the zero-mean one, and the optimizer step that keeps poses, views and
the whole iteration (render, NCC, pose gradient, update) in one replayed
hipGraph, replacing the derivative-free Powell loop.

Loss and gradients are checked against FP64 closed forms and FP64 central
differences of the CPU oracle; the loop against the ground-truth poses it
is meant to recover.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from robartsvtk_amd import abi, configs

import helpers as H
from test_gpu_batch import _plan_with_derivatives, _registration_scene, baked64

pytestmark = pytest.mark.gpu


def ncc_uncentred64(img, tgt):
    a = img.astype(np.float64).ravel()
    b = tgt.astype(np.float64).ravel()
    return float(a @ b / (np.sqrt(a @ a) * np.sqrt(b @ b)))


def _tensors(torch, plan, tg):
    tgt = torch.from_numpy(np.concatenate([t.ravel() for t in tg])).cuda()
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    loss = torch.empty(plan.n_views, dtype=torch.float32, device="cuda")
    grad = torch.empty(plan.n_views * 6, dtype=torch.float32, device="cuda")
    return tgt, out, loss, grad


def test_uncentred_ncc_loss_image_and_pose_gradient(gpu_ctx):
    """vtkImageNormalizedCrossCorrelation::GetResult's form: loss, d loss / d
    image against the FP64 closed form, and the fused step's pose gradient
    against FP64 central differences of the uncentred loss (h = 1e-6)."""
    import torch
    s, tg = _registration_scene(n_views=2)
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    plan = _plan_with_derivatives(gpu_ctx, s, vol)
    plan.set_ncc(abi.NCC_UNCENTERED)
    tgt, out, loss, grad = _tensors(torch, plan, tg)
    plan.register(out, tgt, loss, grad)
    loss2, dldi = torch.empty_like(loss), torch.empty_like(out)
    plan.ncc(out, tgt, loss2, dldi)
    torch.cuda.synchronize()
    assert torch.equal(loss, loss2)
    I = out.cpu().numpy().astype(np.float64)
    img = s.extra["img"]
    for v in range(plan.n_views):
        a = I[plan.offsets[v]:plan.offsets[v + 1]]
        b = tg[v].ravel().astype(np.float64)
        saa, sbb = a @ a, b @ b
        ncc = (a @ b) / (np.sqrt(saa) * np.sqrt(sbb))
        assert abs(float(loss[v]) - (1.0 - ncc)) <= 1e-6
        g = -(b / (np.sqrt(saa) * np.sqrt(sbb)) - ncc * a / saa)
        np.testing.assert_allclose(dldi.cpu().numpy()[plan.offsets[v]:plan.offsets[v + 1]], g, rtol=1e-4,
                                   atol=1e-6 * np.abs(g).max())
        # the uncentred loss is not the centred one on these DRRs
        ac, bc = a - a.mean(), b - b.mean()
        assert abs((ac @ bc) / np.sqrt((ac @ ac) * (bc @ bc)) - ncc) > 1e-4
    grad = grad.cpu().numpy().reshape(-1, 6)
    cam, c = s.extra["cameras"], s.extra["center"]
    h = 1e-6
    for v in range(plan.n_views):
        fd = np.zeros(6)
        for i in range(6):
            vals = []
            for sgn in (1.0, -1.0):
                pose = np.array(s.extra["poses"][v], float)
                pose[i] += sgn * h
                vd = {"m": np.zeros(16, np.float32), "W": img, "H": img,
                      "m64": baked64(cam[v], pose, c, img, img, s.spacing)}
                vals.append(1.0 - ncc_uncentred64(orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f64"), tg[v]))
            fd[i] = (vals[0] - vals[1]) / (2 * h)
        e = np.linalg.norm(grad[v] - fd) / np.linalg.norm(fd)
        print(f"uncentred NCC view {v}: grad {grad[v]}, fd(1e-6) {fd}, rel err {e:.2e}")
        assert e <= 1e-3
    plan.close()
    vol.close()


def _adam(shared, lr_rot=0.01, lr_trans=1.0):
    return abi.PoseOptimizer(kind=abi.OPT_ADAM, shared_pose=int(shared), lr_rotation=lr_rot, lr_translation=lr_trans,
                             beta1=0.9, beta2=0.999, eps=1e-8)


def _pose_err(p, gt):
    d = np.asarray(p) - np.asarray(gt)
    return np.degrees(np.abs(d[..., :3])).max(axis=-1), np.abs(d[..., 3:]).max(axis=-1)


def test_device_poses_render_as_host_posed_views(gpu_ctx):
    """set_poses computes the views on the device with the host's geometry
    (rvk_posed_view): renders and fused steps agree with a plan built from the
    host's views, and a first step reports the loss and gradient of the
    poses before its update."""
    import torch
    s, tg = _registration_scene(n_views=3)
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    host = _plan_with_derivatives(gpu_ctx, s, vol)
    dev = _plan_with_derivatives(gpu_ctx, s, vol)
    dev.set_poses(s.extra["cameras"], s.extra["center"], np.zeros((3, 6)), _adam(False))
    dev.set_poses(s.extra["cameras"], s.extra["center"], s.extra["poses"], _adam(False))
    np.testing.assert_array_equal(dev.poses(), np.asarray(s.extra["poses"]))
    tgt, out_h, loss_h, grad_h = _tensors(torch, host, tg)
    _, out_d, loss_d, grad_d = _tensors(torch, dev, tg)
    host.register(out_h, tgt, loss_h, grad_h)
    dev.step(out_d, tgt, loss_d, grad_d)
    torch.cuda.synchronize()
    torch.testing.assert_close(out_d, out_h, rtol=1e-5, atol=1e-7)
    torch.testing.assert_close(loss_d, loss_h, rtol=1e-5, atol=1e-7)
    torch.testing.assert_close(grad_d, grad_h, rtol=1e-3, atol=1e-6 * float(grad_h.abs().max()))
    # the step moved every pose against its gradient (Adam's first step: lr * sign)
    p1 = dev.poses()
    g = grad_d.cpu().numpy().reshape(-1, 6)
    d = p1 - np.asarray(s.extra["poses"])
    lr = np.array([0.01] * 3 + [1.0] * 3)
    nz = np.abs(g) > 0
    np.testing.assert_allclose(d[nz], (-lr * np.sign(g))[nz], rtol=1e-4)
    host.close()
    dev.close()
    vol.close()


@pytest.mark.parametrize("shared", [False, True])
def test_device_registration_loop_recovers_poses(gpu_ctx, shared):
    """Replaying the captured step (render + NCC + pose gradient + Adam
    update + new views) N times from poses perturbed by up to 5 deg / 10 mm
    moves them to the ground truth: targets are FP64 oracle DRRs at the
    ground truth plus noise (sigma 0.01).  shared: one pose for four views."""
    import torch
    n = 4
    s = configs.scene_c5(scale=0.25, n_views=n)
    c, img = s.extra["center"], s.extra["img"]
    cams = s.extra["cameras"]
    rng = np.random.default_rng(31)
    gt = s.extra["gt_poses"][:1].repeat(n, 0) if shared else s.extra["gt_poses"]
    start = s.extra["poses"][:1].repeat(n, 0) if shared else s.extra["poses"]
    tg = []
    for v in range(n):
        vd = {"m": np.zeros(16, np.float32), "W": img, "H": img, "m64": baked64(cams[v], gt[v], c, img, img, s.spacing)}
        t = orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f64")
        tg.append((t + rng.normal(0.0, 0.01, t.shape)).astype(np.float32))
    views, dms = [], []
    for v in range(n):
        view, dm = abi.posed_view(cams[v], start[v], c, img, img, spacing=s.spacing, with_derivatives=True)
        views.append(view)
        dms.append(dm)
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    plan = abi.DrrPlan(gpu_ctx, vol, views, configs.march_params(s.march), configs.drr_params(s.drr),
                       dview=np.stack(dms))
    plan.set_poses(cams, c, start[:1] if shared else start, _adam(shared, 0.005, 0.5))
    tgt, out, loss, grad = _tensors(torch, plan, tg)
    plan.capture_step(out, tgt, loss, grad)
    e_rot0, e_tr0 = _pose_err(start[:1] if shared else start, gt[:1] if shared else gt)
    losses = []
    steps = 150
    for it in range(steps):
        plan.step(out, tgt, loss, grad)
        if it in (0, steps - 1):
            torch.cuda.synchronize()
            losses.append(float(loss.sum()))
    final = plan.poses()
    e_rot, e_tr = _pose_err(final, gt[:1] if shared else gt)
    print(f"shared={shared}: loss {losses[0]:.5f} -> {losses[1]:.5f}; max rot err {e_rot0.max():.3f} -> "
          f"{e_rot.max():.3f} deg, max trans err {e_tr0.max():.3f} -> {e_tr.max():.3f} mm")
    assert losses[1] < 0.5 * losses[0]
    assert np.all(e_rot < 0.5 * e_rot0 + 0.2) and np.all(e_tr < 0.5 * e_tr0 + 0.5)
    assert e_rot.max() < 1.0 and e_tr.max() < 2.0
    plan.close()
    vol.close()


def test_device_steps_identical_across_relayout_and_work_order(gpu_ctx):
    """Three replayed registration steps give the same images, losses,
    gradients and poses when, between the capture and the replays, the
    volume is re-encoded in another cell layout (the stale graph is not
    replayed: the steps are launched directly) or the forward takes the
    tiles in the region order."""
    import torch
    s, tg = _registration_scene(n_views=3)
    runs = []
    for case in ("plain", "relayout", "regions"):
        vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
        plan = _plan_with_derivatives(gpu_ctx, s, vol)
        plan.set_poses(s.extra["cameras"], s.extra["center"], s.extra["poses"], _adam(False))
        tgt, out, loss, grad = _tensors(torch, plan, tg)
        if case == "regions":
            plan.set_work_order(abi.ORDER_REGIONS)
        plan.capture_step(out, tgt, loss, grad)
        if case == "relayout":
            vol.relayout(abi.CELLS_BRICK4 if vol.layout() != abi.CELLS_BRICK4 else abi.CELLS_PLAIN)
        for _ in range(3):
            plan.step(out, tgt, loss, grad)
        torch.cuda.synchronize()
        runs.append((out.cpu(), loss.cpu(), grad.cpu(), plan.poses()))
        plan.close()
        vol.close()
    for other in runs[1:]:
        for a, b in zip(runs[0][:3], other[:3]):
            assert torch.equal(a, b)
        np.testing.assert_array_equal(runs[0][3], other[3])
