"""Batched DRR, hipGraph replay, NCC, and the DRR adjoint (GPU).

Gradients are checked against FP64 central finite differences of the CPU
oracle (h = 1e-3 rad / mm for poses), with the pose matrices formed in double
precision (no float32 rounding inside the difference quotient).
"""
import numpy as np
import pytest

from oracle import oracle as orc
from robartsvtk_amd import abi, configs

import helpers as H

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def baked64(cam, pose, center, W, H, spacing):
    u, _ = abi.pose_matrix(pose, center)
    m, _ = abi.compute_matrices(cam, W / H, spacing=spacing, user_matrix=u)
    return np.array([[m[r, 0] * -2.0, m[r, 1] * 2.0, m[r, 2], m[r, 3] + m[r, 0] - m[r, 1]] for r in range(4)])


def ncc64(img, tgt):
    a = img.astype(np.float64).ravel() - img.mean()
    b = tgt.astype(np.float64).ravel() - tgt.mean()
    return float(a @ b / np.sqrt((a @ a) * (b @ b)))


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("skip", [0, 1])
@pytest.mark.parametrize("schedule", [abi.SCHED_GRID, abi.SCHED_PERSISTENT])
def test_batch_matches_single_view_renders(gpu_ctx, schedule, skip):
    torch = _torch()
    s = configs.scene_c4(scale=0.125, n_poses=24)
    s.march = dict(s.march, skip_rule=skip)
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    mp, dp = configs.march_params(s.march), configs.drr_params(s.drr)
    plan = abi.DrrPlan(gpu_ctx, vol, s.views, mp, dp)
    plan.set_schedule(schedule, 2)
    out = torch.full((plan.total,), float("nan"), dtype=torch.float32, device="cuda")
    plan.render(out)
    torch.cuda.synchronize()
    for v, view in enumerate(s.views):
        single = torch.empty((view.height, view.width), dtype=torch.float32, device="cuda")
        abi.render_drr(gpu_ctx, vol, view, mp, dp, single)
        torch.cuda.synchronize()
        assert torch.equal(plan.split(out)[v], single), v
    # and against the oracle for two views of different size
    for v in (0, 5):
        o = orc.drr(s.volume, s.spacing, configs.view_dict(s.views[v]), s.march, s.drr, "f32")
        assert H.rel_err(plan.split(out)[v].cpu().numpy(), o).max() <= 2e-6
    plan.close()


def _with_keyhole(view, n):
    key = [(1, 0, 0.01, -0.35 * n), (-1, 0, 0.01, 0.65 * n), (0, 1, 0.01, -0.3 * n), (0, -1, 0.01, 0.7 * n),
           (0.01, 0, 1, -0.3 * n), (0.01, 0, -1, 0.6 * n)]
    view.n_keyhole_planes = 6
    view.keyhole_planes[:] = [float(x) for p in key for x in p]
    return view


@pytest.mark.parametrize("order", [abi.ORDER_VIEWS, abi.ORDER_REGIONS])
def test_batch_views_with_keyhole_planes(gpu_ctx, order):
    """A plan with keyhole planes on some views renders every view as the
    single-view caster does (which specialises views without planes), also
    after update_views adds or removes the planes of a captured plan (and,
    in the region work order, recomputes the order)."""
    torch = _torch()
    s = configs.scene_c4(scale=0.125, n_poses=6)
    n = s.dims[0]
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    mp, dp = configs.march_params(s.march), configs.drr_params(s.drr)
    plain = list(s.views)
    keyed = [_with_keyhole(abi.View.from_buffer_copy(v), n) if i % 2 else v for i, v in enumerate(plain)]

    def check(plan, views, out):
        for v, view in enumerate(views):
            single = torch.empty((view.height, view.width), dtype=torch.float32, device="cuda")
            abi.render_drr(gpu_ctx, vol, view, mp, dp, single)
            torch.cuda.synchronize()
            assert torch.equal(plan.split(out)[v], single), v

    plan = abi.DrrPlan(gpu_ctx, vol, plain, mp, dp)
    plan.set_work_order(order)
    out = torch.zeros(plan.total, dtype=torch.float32, device="cuda")
    plan.capture(out)
    plan.render(out)
    torch.cuda.synchronize()
    check(plan, plain, out)
    base = out.clone()
    plan.update_views(keyed)
    out.zero_()
    plan.render(out)
    torch.cuda.synchronize()
    check(plan, keyed, out)
    assert not torch.equal(out, base)  # the keyholes changed some pixels
    plan.update_views(plain)
    out.zero_()
    plan.render(out)
    torch.cuda.synchronize()
    assert torch.equal(out, base)
    plan.close()
    vol.close()


def test_autotune_and_graph_replays(gpu_ctx):
    torch = _torch()
    s = configs.scene_c4(scale=0.125, n_poses=16)
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    mp, dp = configs.march_params(s.march), configs.drr_params(s.drr)
    plan = abi.DrrPlan(gpu_ctx, vol, s.views, mp, dp)
    out = torch.zeros(plan.total, dtype=torch.float32, device="cuda")
    ms = plan.autotune(out, reps=2)
    assert ms > 0 and plan.schedule()[0] in (abi.SCHED_GRID, abi.SCHED_PERSISTENT)
    plan.render(out)
    torch.cuda.synchronize()
    eager = out.clone()
    plan.capture(out)
    for rep in range(4):
        out.fill_(-1.0)                      # eager work between replays
        torch.cuda.synchronize()
        plan.render(out)                     # graph replay
        torch.cuda.synchronize()
        assert torch.equal(out, eager), rep
    # a persistent-schedule graph needs its queue reset on every replay
    plan.set_schedule(abi.SCHED_PERSISTENT, 4)
    plan.capture(out)
    for rep in range(3):
        out.zero_()
        plan.render(out)
        torch.cuda.synchronize()
        assert torch.equal(out, eager), rep
    # re-uploading the frame with values spanning more than 32767 changes its
    # cell encoding: the captured graph is not replayed for it, the render is
    # launched directly and matches a fresh plan
    wide = np.clip(s.volume.astype(np.int32) * 20, -32768, 32767).astype(np.int16)
    assert int(wide.max()) - int(wide.min()) > 32767
    vol.set_frame(0, wide)
    plan.render(out)
    fresh = abi.DrrPlan(gpu_ctx, vol, s.views, mp, dp)
    ref = torch.zeros_like(out)
    fresh.render(ref)
    torch.cuda.synchronize()
    assert torch.equal(out, ref) and not torch.equal(out, eager)
    fresh.close()
    plan.close()


def test_tile_shapes_and_layout_autotune(gpu_ctx):
    """Every wave tile and work order renders the same images; the layout autotuner
    re-encodes the plan's volume, after which a graph captured before is not
    replayed (its cell pointers are stale) and renders stay identical."""
    torch = _torch()
    s = configs.scene_c4(scale=0.125, n_poses=12)
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    mp, dp = configs.march_params(s.march), configs.drr_params(s.drr)
    plan = abi.DrrPlan(gpu_ctx, vol, s.views, mp, dp)
    ref = torch.zeros(plan.total, dtype=torch.float32, device="cuda")
    plan.render(ref)
    torch.cuda.synchronize()
    out = torch.empty_like(ref)
    for order in (abi.ORDER_REGIONS, abi.ORDER_VIEWS):
        for shape in (abi.TILE_16X4, abi.TILE_4X16, abi.TILE_8X8):
            for sched in (abi.SCHED_GRID, abi.SCHED_PERSISTENT):
                plan.set_work_order(order)
                plan.set_tile_shape(shape)
                plan.set_schedule(sched, 4)
                assert plan.tile_shape() == shape and plan.work_order() == order
                out.fill_(float("nan"))
                plan.render(out)
                torch.cuda.synchronize()
                assert torch.equal(out, ref), (order, shape, sched)
    with pytest.raises(abi.RvkError):
        plan.set_tile_shape(3)
    with pytest.raises(abi.RvkError):
        plan.set_work_order(2)
    plan.capture(out)
    other = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    with pytest.raises(abi.RvkError):
        plan.autotune_layout(other, out, reps=1)   # not the plan's volume
    other.close()
    ms, layout = plan.autotune_layout(vol, out, reps=2)
    assert ms > 0 and vol.layout() == layout
    assert plan.tile_shape() in (abi.TILE_8X8, abi.TILE_16X4, abi.TILE_4X16)
    for rep in range(2):
        out.fill_(-1.0)
        plan.render(out)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
    for layout in (abi.CELLS_PAIR8, abi.CELLS_BRICK4, abi.CELLS_PLAIN):
        plan.capture(out)
        vol.relayout(layout)
        out.fill_(-1.0)
        plan.render(out)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), layout
    plan.close()
    vol.close()


# ---------------------------------------------------------------------------
def _registration_scene(n_views=3, scale=0.125):
    s = configs.scene_c5(scale=scale, n_views=n_views)
    c = s.extra["center"]
    img = s.extra["img"]
    # targets: FP64 oracle DRRs at the ground-truth poses plus seeded noise
    rng = np.random.default_rng(5)
    tg = []
    for v in range(n_views):
        vd = {"m": np.zeros(16, np.float32), "W": img, "H": img,
              "m64": baked64(s.extra["cameras"][v], s.extra["gt_poses"][v], c, img, img, s.spacing)}
        t = orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f64")
        tg.append((t + rng.normal(0.0, s.extra["noise_sigma"], t.shape)).astype(np.float32))
    return s, tg


def _view_size(s, v):
    if "sizes" in s.extra:
        return s.extra["sizes"][v], s.extra["sizes"][v]
    return s.extra["img"], s.extra["img"]


def _fd_pose_gradient(s, tg, v, h=1e-3):
    """FP64 central differences of 1 - NCC(DRR(pose), target) at view v's
    pose, with the pose matrices formed in double precision."""
    cam, c = s.extra["cameras"][v], s.extra["center"]
    W, Hh = _view_size(s, v)
    g = np.zeros(6)
    for i in range(6):
        vals = []
        for sgn in (1.0, -1.0):
            pose = np.array(s.extra["poses"][v], float)
            pose[i] += sgn * h
            vd = {"m": np.zeros(16, np.float32), "W": W, "H": Hh,
                  "m64": baked64(cam, pose, c, W, Hh, s.spacing)}
            I = orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f64")
            vals.append(1.0 - ncc64(I, tg[v]))
        g[i] = (vals[0] - vals[1]) / (2 * h)
    return g


def _plan_with_derivatives(ctx, s, vol):
    c, img = s.extra["center"], s.extra["img"]
    views, dms = [], []
    for v, cam in enumerate(s.extra["cameras"]):
        view, dm = abi.posed_view(cam, s.extra["poses"][v], c, img, img, spacing=s.spacing, with_derivatives=True)
        views.append(view)
        dms.append(dm)
    return abi.DrrPlan(ctx, vol, views, configs.march_params(s.march), configs.drr_params(s.drr),
                       dview=np.stack(dms))


def test_register_loss_and_pose_gradient_vs_fp64_fd(gpu_ctx):
    torch = _torch()
    s, tg = _registration_scene(n_views=3)
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    plan = _plan_with_derivatives(gpu_ctx, s, vol)
    tgt = torch.from_numpy(np.concatenate([t.ravel() for t in tg])).cuda()
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    loss = torch.empty(plan.n_views, dtype=torch.float32, device="cuda")
    grad = torch.empty(plan.n_views * 6, dtype=torch.float32, device="cuda")
    plan.register(out, tgt, loss, grad)
    torch.cuda.synchronize()
    loss, grad = loss.cpu().numpy(), grad.cpu().numpy().reshape(-1, 6)
    img = s.extra["img"]
    for v in range(plan.n_views):
        vd = {"m": np.zeros(16, np.float32), "W": img, "H": img,
              "m64": baked64(s.extra["cameras"][v], s.extra["poses"][v], s.extra["center"], img, img, s.spacing)}
        I = orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f64")
        l64 = 1.0 - ncc64(I, tg[v])
        assert abs(loss[v] - l64) <= 1e-4 * max(1.0, abs(l64)), (v, loss[v], l64)
        # h = 1e-3 rad / mm moves samples by ~1% of a voxel under rotation, so
        # the difference quotient straddles kinks of the (C0) trilinear
        # interpolant; the quotient converges to the analytic gradient as h -> 0
        fd3 = _fd_pose_gradient(s, tg, v, 1e-3)
        fd6 = _fd_pose_gradient(s, tg, v, 1e-6)
        e3r = np.linalg.norm(grad[v][:3] - fd3[:3]) / np.linalg.norm(fd3[:3])
        e3t = np.linalg.norm(grad[v][3:] - fd3[3:]) / np.linalg.norm(fd3[3:])
        e6 = np.linalg.norm(grad[v] - fd6) / np.linalg.norm(fd6)
        print(f"view {v}: loss {loss[v]:.6f} (fp64 {l64:.6f}); grad {grad[v]}; fd(1e-6) {fd6}; "
              f"rel err vs fd(1e-3): rot {e3r:.2e} trans {e3t:.2e}; vs fd(1e-6) {e6:.2e}")
        assert e6 <= 1e-3, (v, e6)
        assert e3t <= 1e-3, (v, e3t)
        # rotations at h = 1e-3 are reported only: the secant error is the
        # quotient's own (it shrinks linearly with h, see fd(1e-6) above)
    plan.close()


def test_backward_with_explicit_image_gradient_matches_register(gpu_ctx):
    torch = _torch()
    s, tg = _registration_scene(n_views=2)
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    plan = _plan_with_derivatives(gpu_ctx, s, vol)
    tgt = torch.from_numpy(np.concatenate([t.ravel() for t in tg])).cuda()
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    loss = torch.empty(plan.n_views, dtype=torch.float32, device="cuda")
    grad = torch.empty(plan.n_views * 6, dtype=torch.float32, device="cuda")
    plan.register(out, tgt, loss, grad)
    loss2 = torch.empty_like(loss)
    dldi = torch.empty_like(out)
    plan.ncc(out, tgt, loss2, dldi)
    grad2 = torch.empty_like(grad)
    plan.backward(out, dldi, grad_pose=grad2)
    torch.cuda.synchronize()
    assert torch.equal(loss, loss2)
    torch.testing.assert_close(grad2, grad, rtol=1e-5, atol=1e-7)
    # dloss/dI against the closed form in FP64
    I = out.cpu().numpy().astype(np.float64)
    for v in range(plan.n_views):
        a = I[plan.offsets[v]:plan.offsets[v + 1]]
        b = tg[v].ravel().astype(np.float64)
        ac, bc = a - a.mean(), b - b.mean()
        saa, sbb = ac @ ac, bc @ bc
        ncc = (ac @ bc) / np.sqrt(saa * sbb)
        g = -(bc / np.sqrt(saa * sbb) - ncc * ac / saa)
        np.testing.assert_allclose(dldi.cpu().numpy()[plan.offsets[v]:plan.offsets[v + 1]], g, rtol=1e-4,
                                   atol=1e-6 * np.abs(g).max())
    plan.close()


def test_registration_graph_replay(gpu_ctx):
    torch = _torch()
    s, tg = _registration_scene(n_views=2)
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    plan = _plan_with_derivatives(gpu_ctx, s, vol)
    tgt = torch.from_numpy(np.concatenate([t.ravel() for t in tg])).cuda()
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    loss = torch.empty(plan.n_views, dtype=torch.float32, device="cuda")
    grad = torch.empty(plan.n_views * 6, dtype=torch.float32, device="cuda")
    plan.register(out, tgt, loss, grad)
    torch.cuda.synchronize()
    ref = (out.clone(), loss.clone(), grad.clone())
    plan.capture(out, tgt, loss, grad)
    for rep in range(3):
        for t in (out, loss, grad):
            t.fill_(123.0)
        torch.cuda.synchronize()
        plan.register(out, tgt, loss, grad)
        torch.cuda.synchronize()
        assert torch.equal(out, ref[0]) and torch.equal(loss, ref[1]) and torch.equal(grad, ref[2]), rep
    plan.close()


def test_volume_gradient_vs_fp64_fd(gpu_ctx):
    torch = _torch()
    s, tg = _registration_scene(n_views=2, scale=0.1)
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
    g = gvol.cpu().numpy().reshape(s.volume.shape)
    img = s.extra["img"]
    vds = [{"m": np.zeros(16, np.float32), "W": img, "H": img,
            "m64": baked64(s.extra["cameras"][v], s.extra["poses"][v], s.extra["center"], img, img, s.spacing)}
           for v in range(plan.n_views)]

    def total_loss(mu):
        return sum(1.0 - ncc64(orc.drr(mu, s.spacing, vds[v], s.march, s.drr, "f64"), tg[v])
                   for v in range(plan.n_views))

    # a random direction over the body voxels (mu > 0 in the whole trilinear neighbourhood)
    rng = np.random.default_rng(9)
    body = s.volume > 0.015
    r = np.where(body, rng.standard_normal(s.volume.shape), 0.0).astype(np.float32)
    eps = 1e-3
    fd = (total_loss((s.volume + eps * r).astype(np.float32)) -
          total_loss((s.volume - eps * r).astype(np.float32))) / (2 * eps)
    an = float((g.astype(np.float64) * r).sum())
    print(f"directional derivative: analytic {an:.6e}, fd {fd:.6e}")
    assert abs(an - fd) <= 1e-3 * abs(fd)
    plan.close()


def test_volume_gradient_repeatable_zero_and_non_finite(gpu_ctx):
    """The volume gradient sums in fixed point, so two calls give the same
    bits whatever order the workgroups' waves add in; a zero d loss / d image
    gives an exactly zero gradient, and a NaN on one ray (with samples) gives
    NaN, not a silently finite gradient."""
    torch = _torch()
    s, tg = _registration_scene(n_views=2, scale=0.1)
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    plan = _plan_with_derivatives(gpu_ctx, s, vol)
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    plan.render(out)
    rng = np.random.default_rng(77)
    dldi = torch.from_numpy(rng.standard_normal(plan.total).astype(np.float32)).cuda()
    g1 = torch.full((s.volume.size,), 3.0, dtype=torch.float32, device="cuda")
    g2 = torch.full_like(g1, -5.0)
    plan.backward(out, dldi, grad_volume=g1)
    plan.backward(out, dldi, grad_volume=g2)
    torch.cuda.synchronize()
    assert torch.equal(g1, g2)
    assert bool(torch.isfinite(g1).all()) and int((g1 != 0).sum()) > 0
    plan.backward(out, torch.zeros_like(dldi), grad_volume=g2)
    torch.cuda.synchronize()
    assert int((g2 != 0).sum()) == 0
    # a pixel at the image centre sees the body
    bad = dldi.clone()
    img = s.extra["img"]
    bad[(img // 2) * img + img // 2] = float("nan")
    plan.backward(out, bad, grad_volume=g2)
    torch.cuda.synchronize()
    assert bool(torch.isnan(g2).any())
    plan.close()
    vol.close()


def test_volume_gradient_many_views_additive_and_homogeneous(gpu_ctx):
    """70 views (more than one LDS chunk of view boxes) on a small positive
    f32 volume whose edge 8-voxel bricks are partial: the gradient of the
    70-view plan is the sum of those of its two 35-view halves (to the
    fixed-point rounding), and doubling d loss / d image doubles the
    gradient bit for bit (the fixed-point scale moves by one power of two)."""
    torch = _torch()
    dims = (21, 19, 18)
    rng = np.random.default_rng(31)
    mu = (0.01 + 0.02 * rng.random(dims[::-1])).astype(np.float32)
    spacing = (1.0, 1.0, 1.0)
    march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 1, "skip_rule": 0, "precise": 1}
    drr = {"intercept": 0.0, "slope": 1.0, "offset": 1.0, "output": abi.DRR_F32_EXP_NEG_L, "ref_termination": 0}
    c = configs.volume_center(dims, spacing)
    views = [abi.compute_view(configs.drr_camera(c, 2.0 * np.pi * i / 70, 16, sod=60.0, sid=100.0, fov=40.0,
                                                 half_extent=20.0), 16, 16, spacing=spacing) for i in range(70)]
    vol = abi.Volume.from_array(gpu_ctx, mu, spacing)

    # the same per-pixel weights for the whole plan and its halves
    w_all = np.random.default_rng(5).standard_normal(70 * 256).astype(np.float32)

    def grad_w(vs, w):
        plan = abi.DrrPlan(gpu_ctx, vol, vs, configs.march_params(march), configs.drr_params(drr))
        out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
        plan.render(out)
        g = torch.zeros(mu.size, dtype=torch.float32, device="cuda")
        plan.backward(out, torch.from_numpy(w).cuda(), grad_volume=g)
        torch.cuda.synchronize()
        plan.close()
        return g.cpu().numpy().astype(np.float64)

    ga = grad_w(views, w_all)
    gb = grad_w(views[:35], w_all[:35 * 256]) + grad_w(views[35:], w_all[35 * 256:])
    scale = np.abs(ga).max()
    assert scale > 0 and np.count_nonzero(ga) > 0.5 * ga.size
    assert np.abs(ga - gb).max() <= 1e-5 * scale, np.abs(ga - gb).max() / scale
    assert np.array_equal(grad_w(views, 2.0 * w_all), 2.0 * ga)
    vol.close()


def test_volume_gradient_of_int16_ct_volume_equals_its_float_copy(gpu_ctx):
    """The volume gradient for the reference's own CT type: a reduced c4
    int16 HU volume (binary16-pair cells, intercept 0.02, slope 2e-5) and a
    float32 copy of it render the same images and give the same gradient
    with respect to the stored voxel values, bit for bit: their trilinear
    values agree to the bit and the sums are fixed-point integers."""
    torch = _torch()
    s = configs.scene_c4(scale=0.125, n_poses=4)
    assert s.volume.dtype == np.int16
    rng = np.random.default_rng(12)
    res = []
    dldi = None
    for arr in (s.volume, s.volume.astype(np.float32)):
        vol = abi.Volume.from_array(gpu_ctx, arr, s.spacing)
        plan = abi.DrrPlan(gpu_ctx, vol, s.views, configs.march_params(s.march), configs.drr_params(s.drr))
        out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
        plan.render(out)
        if dldi is None:
            dldi = torch.from_numpy(rng.standard_normal(plan.total).astype(np.float32)).cuda()
        g = torch.full((arr.size,), 9.0, dtype=torch.float32, device="cuda")
        plan.backward(out, dldi, grad_volume=g)
        torch.cuda.synchronize()
        res.append((out.cpu(), g.cpu()))
        plan.close()
        vol.close()
    assert torch.equal(res[0][0], res[1][0])
    assert torch.equal(res[0][1], res[1][1])
    assert int((res[0][1] != 0).sum()) > 0 and bool(torch.isfinite(res[0][1]).all())


def test_register_and_volume_gradient_identical_across_cell_layouts(gpu_ctx):
    torch = _torch()
    s, tg = _registration_scene(n_views=2, scale=0.1)
    tgt = torch.from_numpy(np.concatenate([t.ravel() for t in tg])).cuda()
    res = []
    for layout in (abi.CELLS_PLAIN, abi.CELLS_BRICK4, abi.CELLS_PAIR8):
        vol = abi.Volume.from_array(gpu_ctx, s.volume, s.spacing, layout=layout)
        plan = _plan_with_derivatives(gpu_ctx, s, vol)
        out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
        loss = torch.empty(plan.n_views, dtype=torch.float32, device="cuda")
        grad = torch.empty(plan.n_views * 6, dtype=torch.float32, device="cuda")
        plan.register(out, tgt, loss, grad)
        dldi = torch.empty_like(out)
        plan.ncc(out, tgt, loss, dldi)
        gvol = torch.zeros(s.volume.size, dtype=torch.float32, device="cuda")
        plan.backward(out, dldi, grad_volume=gvol)
        torch.cuda.synchronize()
        res.append((out.cpu(), loss.cpu(), grad.cpu(), gvol.cpu()))
        plan.close()
        vol.close()
    for other in res[1:]:
        for a, b in zip(res[0], other):
            assert torch.equal(a, b)


# ---------------------------------------------------------------------------
def test_skip_rule_pose_gradient_vs_fp64_fd(gpu_ctx):
    """The adjoint of the skip-rule march (the batched kernels' SKIP
    instances): the skip rule changes which samples are taken only where the
    march leaves and re-enters contributing values, so between such changes
    the accumulated samples are fixed and the pose gradient is that of their
    sum.  Two poses of a reduced c4 scene (int16 HU volume), against FP64
    central differences of the skip-rule march at h = 1e-6."""
    torch = _torch()
    s = configs.scene_c4(scale=0.125, n_poses=2, with_derivatives=True)
    s.march = dict(s.march, skip_rule=1)
    c = s.extra["center"]
    rng = np.random.default_rng(41)
    gt = s.extra["poses"] + np.concatenate([np.radians(rng.uniform(-2, 2, (2, 3))), rng.uniform(-4, 4, (2, 3))], 1)
    tg = []
    for v in range(2):
        W = s.extra["sizes"][v]
        vd = {"m": np.zeros(16, np.float32), "W": W, "H": W,
              "m64": baked64(s.extra["cameras"][v], gt[v], c, W, W, s.spacing)}
        t = orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f64")
        tg.append((t + rng.normal(0.0, 0.01, t.shape)).astype(np.float32))
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    plan = abi.DrrPlan(gpu_ctx, vol, s.views, configs.march_params(s.march), configs.drr_params(s.drr),
                       dview=s.extra["dview"])
    tgt = torch.from_numpy(np.concatenate([t.ravel() for t in tg])).cuda()
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    loss = torch.empty(2, dtype=torch.float32, device="cuda")
    grad = torch.empty(12, dtype=torch.float32, device="cuda")
    plan.register(out, tgt, loss, grad)
    # and the explicit backward of the skip-rule march agrees with the fused step
    dldi = torch.empty_like(out)
    plan.ncc(out, tgt, torch.empty_like(loss), dldi)
    grad2 = torch.empty_like(grad)
    plan.backward(out, dldi, grad_pose=grad2)
    torch.cuda.synchronize()
    torch.testing.assert_close(grad2, grad, rtol=1e-5, atol=1e-7)
    grad = grad.cpu().numpy().reshape(2, 6)
    for v in range(2):
        fd6 = _fd_pose_gradient(s, tg, v, 1e-6)
        e6 = np.linalg.norm(grad[v] - fd6) / np.linalg.norm(fd6)
        print(f"skip-rule view {v}: grad {grad[v]}, fd(1e-6) {fd6}, rel err {e6:.2e}")
        assert e6 <= 1e-3, (v, e6)
    plan.close()
    vol.close()


def _ortho_view(W, H, x0, a, z0, b, y_from, length):
    """Parallel rays along +y: pixel (x, y) starts at (x0 + a vx, y_from,
    z0 + b vy) with vx = 1 - x/W, vy = y/H, and runs `length` voxels."""
    v = abi.View()
    v.view_to_voxels[:] = [a, 0, 0, x0, 0, 0, length, y_from, 0, b, 0, z0, 0, 0, 0, 1]
    v.width, v.height = W, H
    return v


def test_volume_gradient_per_voxel_edges_and_keyhole(gpu_ctx):
    """d loss / d volume of a linear loss, voxel by voxel, against FP64
    central differences (the FP64 march replays the FP32 step counts), on a
    small strictly positive volume: so no sample switches between
    contributing and not, and the loss is smooth in every voxel.  Views:
    (1) parallel rays through x in [0.1, 0.45] voxels, i.e. cell coordinates
    below 1, where the clamped edge cell holds voxel 0 twice; (2) a cone-beam
    view with keyhole planes, whose window's samples are not accumulated."""
    torch = _torch()
    dims = (11, 10, 9)
    rng = np.random.default_rng(23)
    mu = (0.02 + 0.03 * rng.random(dims[::-1])).astype(np.float32)
    spacing = (1.0, 1.2, 0.9)
    march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 1, "skip_rule": 0, "precise": 1}
    drr = {"intercept": 0.0, "slope": 1.0, "offset": 1.0, "output": abi.DRR_F32_EXP_NEG_L, "ref_termination": 0}
    views = [_ortho_view(8, 6, 0.05, 0.4, 2.3, 5.0, -2.0, dims[1] + 4.0)]
    c = configs.volume_center(dims, spacing)
    cam = configs.drr_camera(c, 0.7, 12, sod=60.0, sid=100.0, fov=40.0, half_extent=14.0)
    key = abi.compute_view(cam, 12, 12, spacing=spacing)
    key = _with_keyhole(key, dims[0])
    views.append(key)
    vol = abi.Volume.from_array(gpu_ctx, mu, spacing)
    plan = abi.DrrPlan(gpu_ctx, vol, views, configs.march_params(march), configs.drr_params(drr))
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    plan.render(out)
    w = torch.from_numpy(rng.standard_normal(plan.total).astype(np.float32)).cuda()
    gvol = torch.zeros(mu.size, dtype=torch.float32, device="cuda")
    plan.backward(out, w, grad_volume=gvol)
    torch.cuda.synchronize()
    g = gvol.cpu().numpy().reshape(mu.shape).astype(np.float64)
    wv = w.cpu().numpy().astype(np.float64)
    vds = [configs.view_dict(v) for v in views]
    m64 = dict(march, f32_control=1)
    # the keyhole window does exclude samples of view 2
    full = orc.drr(mu, spacing, dict(vds[1], keyhole=[]), march, drr, "f32")
    assert not np.array_equal(full, orc.drr(mu, spacing, vds[1], march, drr, "f32"))

    def loss(vol):
        return sum(float(orc.drr(vol, spacing, vds[v], m64, drr, "f64", nthreads=1).ravel() @
                         wv[plan.offsets[v]:plan.offsets[v + 1]]) for v in range(len(views)))

    h = 1e-3
    fd = np.zeros_like(g)
    for idx in np.ndindex(*mu.shape):
        up, dn = mu.copy(), mu.copy()
        up[idx] += h
        dn[idx] -= h
        fd[idx] = (loss(up) - loss(dn)) / (2 * h)
    err = np.abs(g - fd)
    scale = np.abs(fd).max()
    print(f"per-voxel volume gradient: max |g - fd| {err.max():.3e} of max |fd| {scale:.3e}; "
          f"x = 0 slab max |fd| {np.abs(fd[:, :, 0]).max():.3e}")
    assert np.abs(fd[:, :, 0]).max() > 0.1 * scale  # the edge rays weigh on voxel x = 0
    np.testing.assert_allclose(g, fd, rtol=1e-3, atol=1e-4 * scale)
    plan.close()
    vol.close()


def test_volume_gradient_per_voxel_single_brick(gpu_ctx):
    """A 7x6x6 volume is one 8^3 brick of the volume adjoint, so the clamped
    edge voxels of both faces of every axis fold into that one brick:
    d loss / d volume of a linear loss, voxel by voxel, against FP64 central
    differences (the FP64 march replays the FP32 step counts), for two
    cone-beam views and parallel views along y through the cells next to
    the low-x and the high-x faces."""
    torch = _torch()
    dims = (7, 6, 6)
    rng = np.random.default_rng(29)
    mu = (0.02 + 0.03 * rng.random(dims[::-1])).astype(np.float32)
    spacing = (1.0, 1.2, 0.9)
    march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 1, "skip_rule": 0, "precise": 1}
    drr = {"intercept": 0.0, "slope": 1.0, "offset": 1.0, "output": abi.DRR_F32_EXP_NEG_L, "ref_termination": 0}
    c = configs.volume_center(dims, spacing)
    views = [abi.compute_view(configs.drr_camera(c, a, 9, sod=60.0, sid=100.0, fov=10.0, half_extent=8.0), 9, 7,
                              spacing=spacing) for a in (0.3, 1.9)]
    views.append(_ortho_view(6, 4, 0.1, 4.6, 1.2, 2.6, -2.0, dims[1] + 4.0))
    views.append(_ortho_view(6, 4, 5.6, 0.9, 1.2, 2.6, -2.0, dims[1] + 4.0))
    vol = abi.Volume.from_array(gpu_ctx, mu, spacing)
    plan = abi.DrrPlan(gpu_ctx, vol, views, configs.march_params(march), configs.drr_params(drr))
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    plan.render(out)
    w = torch.from_numpy(rng.standard_normal(plan.total).astype(np.float32)).cuda()
    gvol = torch.zeros(mu.size, dtype=torch.float32, device="cuda")
    plan.backward(out, w, grad_volume=gvol)
    torch.cuda.synchronize()
    g = gvol.cpu().numpy().reshape(mu.shape).astype(np.float64)
    wv = w.cpu().numpy().astype(np.float64)
    vds = [configs.view_dict(v) for v in views]
    m64 = dict(march, f32_control=1)

    def loss(vol):
        return sum(float(orc.drr(vol, spacing, vds[v], m64, drr, "f64", nthreads=1).ravel() @
                         wv[plan.offsets[v]:plan.offsets[v + 1]]) for v in range(len(views)))

    h = 1e-3
    fd = np.zeros_like(g)
    for idx in np.ndindex(*mu.shape):
        up, dn = mu.copy(), mu.copy()
        up[idx] += h
        dn[idx] -= h
        fd[idx] = (loss(up) - loss(dn)) / (2 * h)
    scale = np.abs(fd).max()
    print(f"single-brick volume gradient: max |g - fd| {np.abs(g - fd).max():.3e} of max |fd| {scale:.3e}")
    assert scale > 0 and np.count_nonzero(fd) > 0.5 * fd.size
    # both x faces carry gradient
    assert np.abs(fd[:, :, 0]).max() > 0.05 * scale and np.abs(fd[:, :, -1]).max() > 0.05 * scale
    np.testing.assert_allclose(g, fd, rtol=1e-3, atol=1e-4 * scale)
    plan.close()
    vol.close()


def test_pose_backward_needs_dview(gpu_ctx):
    torch = _torch()
    s = configs.scene_c4(scale=0.125, n_poses=2)
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    plan = abi.DrrPlan(gpu_ctx, vol, s.views, configs.march_params(s.march), configs.drr_params(s.drr))
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    plan.render(out)
    with pytest.raises(abi.RvkError) as e:
        plan.backward(out, torch.ones_like(out), grad_pose=torch.empty(12, dtype=torch.float32, device="cuda"))
    assert e.value.status == 5
    plan.close()
    vol.close()


def test_volume_backward_rejects_marches_it_does_not_differentiate(gpu_ctx):
    """The volume gradient is that of the precise plain march: a plan with the
    skip rule, with the reference's termination or with the reference FP32
    march is refused with RVK_ERR_UNSUPPORTED (6), before anything runs."""
    torch = _torch()
    s = configs.scene_c4(scale=0.0625, n_poses=2)
    vol = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    g = torch.full((s.volume.size,), 4.0, dtype=torch.float32, device="cuda")
    for march, drr in ((dict(s.march, skip_rule=1), s.drr), (dict(s.march, precise=0), s.drr),
                       (s.march, dict(s.drr, ref_termination=1))):
        plan = abi.DrrPlan(gpu_ctx, vol, s.views, configs.march_params(march), configs.drr_params(drr))
        out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
        plan.render(out)
        with pytest.raises(abi.RvkError) as e:
            plan.backward(out, torch.ones_like(out), grad_volume=g)
        assert e.value.status == 6
        plan.close()
    torch.cuda.synchronize()
    assert bool((g == 4.0).all())
    vol.close()


def test_pose_gradient_at_h_1e3_on_a_smooth_phantom(gpu_ctx):
    """Against FP64 central differences at h = 1e-3 rad / mm -- the step the
    error budget names -- the pose gradient agrees within 1e-3 for rotations
    and translations when the volume is smooth at the voxel scale (a wide
    Gaussian, zero at the volume's faces).  On the CT phantoms the h = 1e-3
    quotient straddles trilinear kinks (per-voxel noise, bone edges) and
    carries an O(h) secant error of its own; the h -> 0 tests above pin the
    adjoint there."""
    torch = _torch()
    n, sig = 96, 10.0
    sp = (1.0, 1.0, 1.0)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    c0 = (n - 1) / 2.0
    r2 = (i - c0 - 6) ** 2 + (j - c0 + 4) ** 2 * 1.3 + (k - c0) ** 2 * 0.8
    mu = (0.03 * np.exp(-r2 / (2 * sig * sig))).astype(np.float32)
    c = configs.volume_center((n, n, n), sp)
    img = 96
    march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 1, "skip_rule": 0, "precise": 1}
    drr = {"intercept": 0.0, "slope": 1.0, "offset": 1.0, "output": abi.DRR_F32_EXP_NEG_L, "ref_termination": 0}
    cams = [configs.drr_camera(c, a, img, sod=3.0 * n, sid=5.0 * n, fov=1.6 * n, half_extent=0.9 * n)
            for a in (0.4, 1.9)]
    gt = [np.array([0.02, -0.015, 0.01, 2.0, -1.0, 1.5]), np.array([-0.01, 0.02, 0.015, -1.0, 2.0, -0.5])]
    poses = [g + d for g, d in zip(gt, (np.array([0.03, 0.02, -0.025, -3.0, 2.0, 2.5]),
                                        np.array([-0.02, 0.03, 0.02, 2.0, -3.0, 1.5])))]

    def render64(v, p):
        vd = {"m": np.zeros(16, np.float32), "W": img, "H": img, "m64": baked64(cams[v], p, c, img, img, sp)}
        return orc.drr(mu, sp, vd, march, drr, "f64")

    tg = [render64(v, gt[v]).astype(np.float32) for v in range(2)]
    views, dms = [], []
    for v in range(2):
        view, dm = abi.posed_view(cams[v], poses[v], c, img, img, spacing=sp, with_derivatives=True)
        views.append(view)
        dms.append(dm)
    vol = abi.Volume.from_array(gpu_ctx, mu, sp)
    plan = abi.DrrPlan(gpu_ctx, vol, views, configs.march_params(march), configs.drr_params(drr), dview=np.stack(dms))
    tgt = torch.from_numpy(np.concatenate([t.ravel() for t in tg])).cuda()
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    loss = torch.empty(2, dtype=torch.float32, device="cuda")
    grad = torch.empty(12, dtype=torch.float32, device="cuda")
    plan.register(out, tgt, loss, grad)
    torch.cuda.synchronize()
    grad = grad.cpu().numpy().reshape(2, 6)
    h = 1e-3
    for v in range(2):
        fd = np.zeros(6)
        for a in range(6):
            p1, p2 = poses[v].copy(), poses[v].copy()
            p1[a] += h
            p2[a] -= h
            fd[a] = ((1.0 - ncc64(render64(v, p1), tg[v])) - (1.0 - ncc64(render64(v, p2), tg[v]))) / (2 * h)
        er = np.linalg.norm(grad[v][:3] - fd[:3]) / np.linalg.norm(fd[:3])
        et = np.linalg.norm(grad[v][3:] - fd[3:]) / np.linalg.norm(fd[3:])
        print(f"smooth phantom view {v}: grad {grad[v]}, fd(1e-3) {fd}, rel err rot {er:.2e}, trans {et:.2e}")
        assert er <= 1e-3 and et <= 1e-3, (v, er, et)
    plan.close()
    vol.close()
