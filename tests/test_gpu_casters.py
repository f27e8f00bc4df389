"""Parity of the gfx950 casters (called through the C-ABI) with the CPU oracle.

Tolerances: DRR float outputs <= 1e-4 relative error against the FP64 oracle
(replaying the FP32 branch decisions) and, against the FP32 oracle in the
same march mode, <= 2e-6 (both follow the same operation order; only
expf/logf differ); RGBA8 outputs within 1 LSB on >= 99.9 % of pixels.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from robartsvtk_amd import abi, configs

import helpers as H

pytestmark = pytest.mark.gpu

DRR_REL_TOL = 1e-4
SAME_MODE_REL_TOL = 2e-6


def _drr_check(ctx, vol, spacing, view, march, drr, tol_f64=DRR_REL_TOL, tol_f32=SAME_MODE_REL_TOL):
    vh = H.gpu_volume(ctx, vol, spacing)
    g = H.gpu_drr(ctx, vh, view, march, drr)
    vd = configs.view_dict(view)
    o32 = orc.drr(vol, spacing, vd, march, drr, "f32")
    o64 = orc.drr(vol, spacing, vd, dict(march, f32_control=1), drr, "f64")
    vh.close()
    if drr["output"] == abi.DRR_RGBA8_REF:
        frac32, mx32 = H.lsb_stats(g, o32)
        frac64, mx64 = H.lsb_stats(g, o64)
        print(f"rgba8 within 1 LSB: vs f32 {frac32:.6f} (max {mx32}), vs f64 {frac64:.6f} (max {mx64})")
        assert frac32 >= 0.999 and frac64 >= 0.999
        return g
    assert np.all(np.isfinite(g))
    if drr["output"] == abi.DRR_F32_L:
        # intensity exp(-L) has relative error |dL|: compare L absolutely
        e32 = np.abs(g - o32).max()
        e64 = np.abs(g - o64).max()
    else:
        e32 = H.rel_err(g, o32).max()
        e64 = H.rel_err(g, o64).max()
    print(f"max rel err vs f32 oracle {e32:.3e}, vs f64 oracle {e64:.3e}")
    assert e32 <= tol_f32, e32
    assert e64 <= tol_f64, e64
    return g


# ---------------------------------------------------------------------------
# DRR
@pytest.mark.parametrize("dtype", ["u8", "i8", "u16", "i16", "u32", "i32", "f32"])
def test_drr_all_voxel_types(gpu_ctx, dtype):
    s = configs.scene_c2(scale=0.125)
    hu = s.volume.astype(np.float64)
    drr = dict(s.drr)
    # store the phantom in each native type, with slope/intercept giving the same mu
    if dtype == "u8":
        vol = np.clip(np.rint((hu + 1000.0) / 10.0), 0, 255).astype(np.uint8)
        drr.update(slope=2e-4, intercept=0.0)
    elif dtype == "i8":
        vol = np.clip(np.rint(hu / 20.0), -128, 127).astype(np.int8)
        drr.update(slope=4e-4, intercept=0.02)
    elif dtype == "u16":
        vol = (hu + 1024.0).astype(np.uint16)
        drr.update(intercept=0.02 - 1024 * 2e-5)
    elif dtype == "i16":
        vol = s.volume
    elif dtype == "u32":
        vol = (hu + 1024.0).astype(np.uint32)
        drr.update(intercept=0.02 - 1024 * 2e-5)
    elif dtype == "i32":
        vol = hu.astype(np.int32)
    else:
        vol = configs.mu_from_hu(s.volume)
        drr.update(slope=1.0, intercept=0.0)
    _drr_check(gpu_ctx, vol, s.spacing, s.views[0], s.march, drr)


@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 2, 5), (4, 3, 2), (5, 4, 3), (7, 6, 6)])
def test_drr_smallest_volumes_ragged_views(gpu_ctx, dims):
    """The smallest volumes the library accepts (2 or 3 voxels on an axis:
    the reference's crop box [1, dim - 2] is then empty or flat, so every
    ray is rejected) and a 7x6x6 volume whose rays do cross it, f32 and
    int16, on 9x7 views (no tile is whole), skip rule on and off, precise
    march and the reference's: single-view renders against the oracle, and
    a two-view batched plan bit for bit against them."""
    import torch
    rng = np.random.default_rng(sum(dims))
    spacing = (1.0, 1.2, 0.9)
    c = configs.volume_center(dims, spacing)
    views = [abi.compute_view(configs.drr_camera(c, a, 9, sod=60.0, sid=100.0, fov=10.0, half_extent=8.0), 9, 7,
                              spacing=spacing) for a in (0.3, 1.9)]
    for vol, drr in ((rng.uniform(0.0, 0.05, dims[::-1]).astype(np.float32),
                      {"intercept": 0.0, "slope": 1.0, "offset": 1.0, "output": abi.DRR_F32_EXP_NEG_L,
                       "ref_termination": 0}),
                     (rng.integers(-1000, 1500, dims[::-1]).astype(np.int16),
                      {"intercept": 0.02, "slope": 2e-5, "offset": 1.0, "output": abi.DRR_F32_EXP_NEG_L,
                       "ref_termination": 0})):
        for skip in (0, 1):
            for precise in (0, 1):
                march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 1, "skip_rule": skip, "precise": precise}
                singles = [_drr_check(gpu_ctx, vol, spacing, v, march, drr) for v in views]
                if min(dims) <= 3:
                    assert all(np.all(s == 1.0) for s in singles)
                else:
                    assert any(np.any(s < 1.0) for s in singles)
                vh = H.gpu_volume(gpu_ctx, vol, spacing)
                plan = abi.DrrPlan(gpu_ctx, vh, views, configs.march_params(march), configs.drr_params(drr))
                out = torch.full((plan.total,), float("nan"), dtype=torch.float32, device="cuda")
                plan.render(out)
                torch.cuda.synchronize()
                got = out.cpu().numpy()
                for i, s1 in enumerate(singles):
                    assert np.array_equal(got[plan.offsets[i]:plan.offsets[i + 1]], s1.ravel()), (skip, precise, i)
                plan.close()
                vh.close()


@pytest.mark.parametrize("dtype", ["u16", "i16"])
def test_drr_wide_range_16bit_frames(gpu_ctx, dtype):
    """16-bit frames whose values span 32768 or more keep plain (a, b) cells
    (their x-pair differences do not fit 16 bits); a value range just below
    keeps difference cells.  Both render as the oracle does."""
    s = configs.scene_c2(scale=0.125)
    hu = s.volume.astype(np.float64)
    drr = dict(s.drr)
    for stretch in (16.0, 7.9):  # HU span ~2500 x stretch: above and below 32767
        if dtype == "u16":
            vol = np.clip(np.rint((hu + 1000.0) * stretch), 0, 65535).astype(np.uint16)
            drr.update(slope=2e-5 / stretch, intercept=0.0)
        else:
            vol = np.clip(np.rint(hu * stretch), -32768, 32767).astype(np.int16)
            drr.update(slope=2e-5 / stretch, intercept=0.02)
        span = int(vol.max()) - int(vol.min())
        assert (span > 32767) == (stretch > 10)
        _drr_check(gpu_ctx, vol, s.spacing, s.views[0], s.march, drr)


@pytest.mark.parametrize("dtype", ["i16", "u16"])
def test_binary16_base_frames(gpu_ctx, dtype):
    """16-bit frames whose values all lie in [-2048, 2048] keep binary16
    bases in their cells, and binary16 x-pair differences too when every
    difference is exact in binary16 (|d| <= 2048, or even up to 4096); they
    render exactly as the same values in float32 cells do, up to the end
    points of that interval and one past them (a frame reaching 2049 keeps
    int16 bases).  Cases: "pair" (values even: differences up to 4096, all
    exact), "base" (i16 only: one x-pair (-2048, 2047) planted in the middle,
    difference 4095, which binary16 would round), "int16" (top 2049)."""
    s = configs.scene_c2(scale=0.125)
    hu = s.volume.astype(np.float64)
    t = (hu - hu.min()) / (hu.max() - hu.min())  # 0 (air) .. 1 (densest bone)
    for case in ("pair", "base", "int16") if dtype == "i16" else ("pair", "int16"):
        top = 2049 if case == "int16" else 2048
        if dtype == "i16":  # stretched onto [-2048, top]
            v = np.rint(-2048 + t * (top + 2048))
            drr = dict(s.drr, slope=0.05 / (top + 2048), intercept=0.05 * 2048 / (top + 2048))
        else:
            v = np.rint(t * top)
            drr = dict(s.drr, slope=0.05 / top, intercept=0.0)
        if case != "int16":
            v = 2 * np.floor(v / 2)
        if case == "base":
            z, y, x = (n // 2 for n in v.shape)
            v[z, y, x : x + 2] = (-2048, 2047)
        vol = v.astype(np.int16 if dtype == "i16" else np.uint16)
        assert int(vol.max()) == top and int(vol.min()) == (-2048 if dtype == "i16" else 0)
        d = np.abs(np.diff(vol.astype(np.int32), axis=2))
        if case != "int16":
            assert ((d > 2048) & (d % 2 == 1)).any() == (case == "base")
        assert d.max() > 2048 or dtype == "u16"
        g = _drr_check(gpu_ctx, vol, s.spacing, s.views[0], s.march, drr)
        vf = H.gpu_volume(gpu_ctx, vol.astype(np.float32), s.spacing)
        gf = H.gpu_drr(gpu_ctx, vf, s.views[0], s.march, drr)
        vf.close()
        assert np.array_equal(g, gf), case
    # a compositing caster on binary16-base cells: c1's u8 values as uint16
    c1 = configs.scene_c1(scale=0.25)
    march = dict(c1.march, term_T=1.0 / 64)
    v8 = H.gpu_volume(gpu_ctx, c1.volume, c1.spacing)
    v16 = H.gpu_volume(gpu_ctx, c1.volume.astype(np.uint16), c1.spacing)
    a = H.gpu_1d(gpu_ctx, v8, c1.tf1d, c1.views[0], march)
    b = H.gpu_1d(gpu_ctx, v16, c1.tf1d, c1.views[0], march)
    v8.close()
    v16.close()
    assert np.array_equal(a, b) and a[..., 3].max() > 200
    # the 2D caster with the gradient-magnitude volume computed from them:
    # c3's values halved (0..2047) against the same values in float32 cells
    c3 = configs.scene_c3(scale=0.25)
    half = c3.volume // 2
    tf = dict(c3.tf2d, params=dict(c3.tf2d["params"], intensity_multiplier=float(np.float32(2.0 / 4095.0))))
    outs = []
    for vol in (half.astype(np.uint16), half.astype(np.float32)):
        vh = H.gpu_volume(gpu_ctx, vol, c3.spacing)
        vh.compute_gradmag(0, c3.tf2d["gradmag_max"])
        outs.append(H.gpu_2d(gpu_ctx, vh, tf, c3.views[0], c3.march, use_gradmag=True))
        vh.close()
    assert np.array_equal(outs[0], outs[1]) and outs[0][..., 3].max() > 100


@pytest.mark.parametrize("skip", [0, 1])
@pytest.mark.parametrize("precise", [0, 1])
@pytest.mark.parametrize("jitter", [0, 1])
def test_drr_march_modes(gpu_ctx, skip, precise, jitter):
    s = configs.scene_c2(scale=0.25)
    march = dict(s.march, skip_rule=skip, precise=precise, jitter=jitter)
    # the reference (non-precise) FP32 march drifts; its FP64 gap is only reported
    tol64 = DRR_REL_TOL if precise else 1.0
    _drr_check(gpu_ctx, s.volume, s.spacing, s.views[0], march, s.drr, tol_f64=tol64)


def _sparse_volume(dtype, sign, dims=(61, 45, 53), seed=11):
    """A background whose samples contribute nothing (alpha >= 0, exactly 0
    for f32) holding isolated contributing voxels and one small dense block,
    with its slope/intercept."""
    rng = np.random.default_rng(seed)
    shape = dims[::-1]
    hot = rng.random(shape) < 0.002
    hot[20:24, 10:13, 30:36] = True
    val = rng.uniform(0.3, 1.0, shape)
    if dtype == "f32":
        vol = np.where(hot, sign * val, 0.0).astype(np.float32)
        drr = dict(slope=float(sign), intercept=0.0)
    elif dtype == "f32c":  # the background contributes one constant alpha; rays may terminate in it
        vol = np.where(hot, sign * val, sign * 0.1).astype(np.float32)
        drr = dict(slope=float(sign), intercept=0.0, ref_termination=1)
    elif dtype == "i16":  # mu = 2e-5 (v + 1024) >= 0 from v = -1024 up
        vol = np.where(hot, -1024 + sign * np.rint(3000 * val), -1024).astype(np.int16)
        drr = dict(slope=sign * 2e-5, intercept=sign * 2e-5 * 1024)
    else:  # u8, mu = sign (v - 128) / 500
        vol = np.where(hot, 128 + sign * np.rint(120 * val), 128).astype(np.uint8)
        drr = dict(slope=sign / 500.0, intercept=-sign * 128 / 500.0)
    return vol, drr


@pytest.mark.parametrize("dtype,sign", [("f32", 1), ("f32", -1), ("f32c", 1), ("i16", 1), ("i16", -1), ("u8", -1)])
@pytest.mark.parametrize("layout", [abi.CELLS_PLAIN, abi.CELLS_BRICK4, abi.CELLS_PAIR8])
def test_drr_sparse_volumes(gpu_ctx, dtype, sign, layout):
    """Mostly empty volumes: the batched precise march passes over runs of
    samples that cannot contribute (or that contribute one constant alpha)
    and must give what the single-view march, which takes every sample,
    gives: with either skip rule, slope sign (the upper or the lower bound of
    a region's values decides), ragged dims, and perspective and
    axis-parallel views (rays with zero direction components)."""
    dims, spacing = (61, 45, 53), (1.0, 1.2, 0.9)
    vol, d = _sparse_volume(dtype, sign, dims)
    drr = {"offset": 1.0, "output": abi.DRR_F32_L, "ref_termination": 0, "tint": (0, 0, 0, 0)}
    drr.update(d)
    c = configs.volume_center(dims, spacing)
    views = []
    for k, ang in enumerate((0.0, 0.7, 2.3)):
        cam = configs.drr_camera(c, ang, 0, sod=300.0, half_extent=60.0)
        views.append(abi.compute_view(cam, 67 + 8 * k, 49, spacing=spacing))
    along_x = abi.camera(c + np.array([300.0, 0, 0]), c, (0, 0, 1), 30.0, (200.0, 400.0), parallel=True,
                         parallel_scale=30.0)
    views.append(abi.compute_view(along_x, 65, 47, spacing=spacing))
    # a keyhole: contributing runs must stop at its window
    kv = abi.compute_view(configs.drr_camera(c, 0.7, 0, sod=300.0, half_extent=60.0), 71, 45, spacing=spacing)
    n = dims[0]
    key = [(1, 0, 0.01, -0.35 * n), (-1, 0, 0.01, 0.65 * n), (0, 1, 0.01, -0.3 * n), (0, -1, 0.01, 0.7 * n),
           (0.01, 0, 1, -0.3 * n), (0.01, 0, -1, 0.6 * n)]
    kv.n_keyhole_planes = 6
    kv.keyhole_planes[:] = [float(x) for p in key for x in p]
    views.append(kv)
    import torch
    vh = abi.Volume.from_array(gpu_ctx, vol, spacing, layout=layout)
    try:
        for skip in (0, 1):
            march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 1, "skip_rule": skip, "precise": 1}
            plan = abi.DrrPlan(gpu_ctx, vh, views, configs.march_params(march), configs.drr_params(drr))
            packed = torch.full((plan.total,), float("nan"), dtype=torch.float32, device="cuda")
            plan.render(packed)
            torch.cuda.synchronize()
            batched = [b.cpu().numpy() for b in plan.split(packed)]
            plan.close()
            for i, view in enumerate(views):
                g = H.gpu_drr(gpu_ctx, vh, view, march, drr)
                o32 = orc.drr(vol, spacing, configs.view_dict(view), march, drr, "f32")
                assert np.abs(o32).max() > 0.1  # the rays meet contributing voxels
                err = np.abs(g - o32).max()
                print(f"view {i} skip {skip}: max |L - L_f32| {err:.3e}")
                assert err <= 1e-5 * max(1.0, np.abs(o32).max()), (i, skip, err)
                assert np.array_equal(batched[i], g), (i, skip, np.abs(batched[i] - g).max())
    finally:
        vh.close()


@pytest.mark.parametrize("output", [abi.DRR_RGBA8_REF, abi.DRR_F32_EXP_NEG_L, abi.DRR_F32_L])
@pytest.mark.parametrize("ref_term", [0, 1])
def test_drr_outputs_and_termination(gpu_ctx, output, ref_term):
    s = configs.scene_c2(scale=0.25)
    drr = dict(s.drr, output=output, ref_termination=ref_term, tint=(200, 120, 40, 90), offset=0.9)
    _drr_check(gpu_ctx, s.volume, s.spacing, s.views[0], s.march, drr)


def test_drr_reference_defaults(gpu_ctx):
    """Reference constants: CT 0.022/0.000022/1.0, step = MinSpacing, RGBA8."""
    s = configs.scene_c2(scale=0.25)
    march = {"step_scale": 1.0, "term_T": 0.0, "jitter": 1, "skip_rule": 1, "precise": 0}
    drr = {"intercept": 0.022, "slope": 0.000022, "offset": 1.0, "output": abi.DRR_RGBA8_REF,
           "ref_termination": 1, "tint": (0, 0, 0, 0)}
    _drr_check(gpu_ctx, s.volume, s.spacing, s.views[0], march, drr)


def test_drr_c2_full_size(gpu_ctx):
    """The benchmarked DRR workload at full size (256^3 i16 -> 1024^2 f32)."""
    s = configs.scene_c2(scale=1.0)
    g = _drr_check(gpu_ctx, s.volume, s.spacing, s.views[0], s.march, s.drr)
    assert g.min() > 0.0 and g.max() <= 1.0


# ---------------------------------------------------------------------------
# geometry edge cases
def _small_drr_scene():
    s = configs.scene_c2(scale=0.125)
    return s


@pytest.mark.parametrize("wh", [(1, 1), (37, 23), (16, 16), (130, 7)])
def test_drr_ragged_sizes(gpu_ctx, wh):
    s = _small_drr_scene()
    cam = s.extra["camera"]
    view = abi.compute_view(cam, wh[0], wh[1], spacing=s.spacing)
    _drr_check(gpu_ctx, s.volume, s.spacing, view, s.march, s.drr)


def test_drr_rays_missing_volume(gpu_ctx):
    s = _small_drr_scene()
    c = configs.volume_center(s.dims, s.spacing)
    cam = abi.camera(c + np.array([600.0, 0, 0]), c + np.array([1200.0, 0, 0]), (0, 0, 1), 20.0, (10.0, 400.0))
    view = abi.compute_view(cam, 40, 24, spacing=s.spacing)
    g = _drr_check(gpu_ctx, s.volume, s.spacing, view, s.march, s.drr)
    assert np.all(g == 1.0)  # exp(-0)


def test_drr_flipped_clip_and_keyhole(gpu_ctx):
    s = configs.scene_c2(scale=0.25)
    view = abi.compute_view(s.extra["camera"], 96, 80, spacing=s.spacing, flipped=True)
    n = s.dims[0]
    planes = [(1, 0, 0, -0.2 * n), (-1, 0, 0, 0.9 * n), (0, 1, 0, -0.1 * n), (0, -1, 0, 0.8 * n),
              (0, 0, 1, -0.15 * n), (0, 0, -1, 0.95 * n)]
    view.n_clip_planes = 6
    view.clip_planes[:] = [float(x) for p in planes for x in p]
    key = [(0, 0, 1, -0.5 * n), (0, 0, -1, 0.7 * n), (1, 0, 0, -0.3 * n), (-1, 0, 0, 0.6 * n),
           (0, 1, 0, -0.3 * n), (0, -1, 0, 0.7 * n)]
    view.n_keyhole_planes = 6
    view.keyhole_planes[:] = [float(x) for p in key for x in p]
    for skip in (0, 1):
        _drr_check(gpu_ctx, s.volume, s.spacing, view, dict(s.march, skip_rule=skip), s.drr)


def test_drr_zbuffer(gpu_ctx):
    import torch
    s = _small_drr_scene()
    view = abi.compute_view(s.extra["camera"], 64, 48, spacing=s.spacing)
    zb = np.linspace(0.3, 1.0, 50 * 30, dtype=np.float32).reshape(30, 50)
    zb_dev = torch.from_numpy(zb).cuda()
    view.zbuffer = zb_dev.data_ptr()
    view.zb_width, view.zb_height = 50, 30
    vh = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    g = H.gpu_drr(gpu_ctx, vh, view, s.march, s.drr)
    vd = dict(configs.view_dict(view), zbuffer=zb)
    o32 = orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f32")
    assert H.rel_err(g, o32).max() <= SAME_MODE_REL_TOL


# ---------------------------------------------------------------------------
# 1D transfer function
def _check_rgba(g, o32, o64, label):
    f32, m32 = H.lsb_stats(g, o32)
    f64, m64 = H.lsb_stats(g, o64)
    exact = float(np.mean(np.all(g == o32, axis=-1)))
    print(f"{label}: within 1 LSB vs f32 {f32:.6f} (max {m32}, bit-exact {exact:.6f}); vs f64 {f64:.6f} (max {m64})")
    assert f32 >= 0.999 and f64 >= 0.999


@pytest.mark.parametrize("scale", [0.25, 1.0])
def test_1d_c1(gpu_ctx, scale):
    s = configs.scene_c1(scale=scale)
    vh = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    g = H.gpu_1d(gpu_ctx, vh, s.tf1d, s.views[0], s.march)
    vd = configs.view_dict(s.views[0])
    o32, _, _ = orc.cast1d(s.volume, s.spacing, s.tf1d, vd, s.march, "f32")
    o64, _, _ = orc.cast1d(s.volume, s.spacing, s.tf1d, vd, dict(s.march, f32_control=1), "f64")
    _check_rgba(g, o32, o64, f"c1 scale {scale}")
    assert g[..., 3].max() > 200  # something opaque was hit


@pytest.mark.parametrize("skip", [0, 1])
@pytest.mark.parametrize("precise", [0, 1])
def test_1d_shaded_gradient_opacity(gpu_ctx, skip, precise):
    s = configs.scene_c1(scale=0.25)
    tf = dict(s.tf1d)
    n = tf["alpha"].shape[0]
    tf["galpha"] = np.linspace(0.2, 1.0, n).astype(np.float32)
    tf["params"] = dict(tf["params"], gradient_low=0.0, gradient_multiplier=1.0 / 40.0, ambient=0.3, diffuse=0.6,
                        specular=0.4, specular_power=8.0)
    march = dict(s.march, skip_rule=skip, precise=precise, term_T=1.0 / 64)
    vh = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    g = H.gpu_1d(gpu_ctx, vh, tf, s.views[0], march)
    vd = configs.view_dict(s.views[0])
    o32, _, _ = orc.cast1d(s.volume, s.spacing, tf, vd, march, "f32")
    o64, _, _ = orc.cast1d(s.volume, s.spacing, tf, vd, dict(march, f32_control=1), "f64")
    _check_rgba(g, o32, o64, "1d shaded")


def test_1d_aux_and_post_shade(gpu_ctx):
    import torch
    s = configs.scene_c1(scale=0.25)
    vh = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    g, (depth, mind, maxd) = H.gpu_1d(gpu_ctx, vh, s.tf1d, s.views[0], s.march, with_aux=True)
    vd = configs.view_dict(s.views[0])
    o32, _, odepth = orc.cast1d(s.volume, s.spacing, s.tf1d, vd, s.march, "f32")
    rays = orc.form_rays(s.dims, s.spacing, vd, s.march, "f32")
    assert np.array_equal(depth, odepth)
    assert np.array_equal(mind, rays["min_depth"]) and np.array_equal(maxd, rays["max_depth"])
    shade = {"celr": 0.7, "cela": 0.002, "celc": 1.0 / (0.005 - 0.002), "disr": 0.25, "disa": 0.45,
             "disc": 1.0 / (0.6 - 0.45)}
    rgba = torch.from_numpy(np.ascontiguousarray(g)).cuda()
    dd, dmin, dmax = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (depth, mind, maxd))
    sp = abi.ShadeParams(*[shade[k] for k in ("celr", "cela", "celc", "disr", "disa", "disc")])
    abi.check(abi.lib().rvk_post_shade(gpu_ctx.handle, g.shape[1], g.shape[0], sp, dd.data_ptr(), dmin.data_ptr(),
                                       dmax.data_ptr(), rgba.data_ptr()))
    torch.cuda.synchronize()
    od, orgba = orc.post_shade(shade, depth, mind, maxd, g)
    assert np.array_equal(rgba.cpu().numpy(), orgba)
    assert np.array_equal(dd.cpu().numpy(), od)


# ---------------------------------------------------------------------------
# 2D transfer function
def _scene_c3_small():
    return configs.scene_c3(scale=0.25)


def test_2d_c3_gradmag(gpu_ctx):
    s = _scene_c3_small()
    vh = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    gmax = s.tf2d["gradmag_max"]
    vh.compute_gradmag(0, gmax)
    g = H.gpu_2d(gpu_ctx, vh, s.tf2d, s.views[0], s.march, use_gradmag=True)
    gm = H.gradmag_u8(s.volume, s.spacing, gmax)
    tf = dict(s.tf2d, gradmag=gm, gradmag_scale=float(np.float32(gmax) / np.float32(255.0)))
    vd = configs.view_dict(s.views[0])
    o32, _, _ = orc.cast2d(s.volume, s.spacing, tf, vd, s.march, "f32")
    o64, _, _ = orc.cast2d(s.volume, s.spacing, tf, vd, dict(s.march, f32_control=1), "f64")
    _check_rgba(g, o32, o64, "c3 gradmag")
    assert g[..., 3].max() > 100


@pytest.mark.parametrize("skip", [0, 1])
def test_2d_on_the_fly_gradient_shaded(gpu_ctx, skip):
    s = _scene_c3_small()
    tf = dict(s.tf2d)
    N = tf["tables"]["alpha"].shape[0]
    r = np.random.default_rng(7)
    tables = dict(tf["tables"])
    tables["ambient"] = (0.2 + 0.3 * r.random((N, N))).astype(np.float32)
    tables["diffuse"] = (0.5 * r.random((N, N))).astype(np.float32)
    tables["specular"] = (0.4 * r.random((N, N))).astype(np.float32)
    tables["specular_power"] = (1.0 + 10.0 * r.random((N, N))).astype(np.float32)
    tf["tables"] = tables
    march = dict(s.march, skip_rule=skip, term_T=1.0 / 32)
    vh = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    g = H.gpu_2d(gpu_ctx, vh, tf, s.views[0], march, use_gradmag=False)
    vd = configs.view_dict(s.views[0])
    o32, _, _ = orc.cast2d(s.volume, s.spacing, tf, vd, march, "f32")
    o64, _, _ = orc.cast2d(s.volume, s.spacing, tf, vd, dict(march, f32_control=1), "f64")
    _check_rgba(g, o32, o64, "2d shaded")


def test_2d_keyhole_dual_tf(gpu_ctx):
    s = _scene_c3_small()
    tf = dict(s.tf2d)
    tf["keyhole_tables"] = configs.tf2d_bumps(99, tf["tables"]["alpha"].shape[0])
    view = abi.compute_view(s.extra["camera"], 120, 90, spacing=s.spacing)
    n = s.dims[0]
    key = [(1, 0, 0, -0.3 * n), (-1, 0, 0, 0.7 * n), (0, 1, 0, -0.3 * n), (0, -1, 0, 0.7 * n),
           (0, 0, 1, -0.2 * n), (0, 0, -1, 0.8 * n)]
    view.n_keyhole_planes = 6
    view.keyhole_planes[:] = [float(x) for p in key for x in p]
    vh = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    g, (depth, _, _) = H.gpu_2d(gpu_ctx, vh, tf, view, s.march, use_gradmag=False, with_aux=True)
    vd = configs.view_dict(view)
    o32, _, od = orc.cast2d(s.volume, s.spacing, tf, vd, s.march, "f32")
    o64, _, _ = orc.cast2d(s.volume, s.spacing, tf, vd, dict(s.march, f32_control=1), "f64")
    _check_rgba(g, o32, o64, "2d dual")


# ---------------------------------------------------------------------------
# error behaviour through the ABI
def test_errors(gpu_ctx):
    import torch
    v = abi.Volume(gpu_ctx, (8, 8, 8), (1, 1, 1), "u8", n_frames=2)
    cam = abi.camera((3.5, 3.5, 40.0), (3.5, 3.5, 3.5), (0, 1, 0), 20.0, (10.0, 80.0))
    view = abi.compute_view(cam, 4, 4)
    out = torch.zeros(16, dtype=torch.float32, device="cuda")
    m = abi.MarchParams(1.0, 0.0, 1, 1, 1)
    d = configs.drr_params({"intercept": 0.0, "slope": 2e-5, "offset": 1.0, "output": 1, "ref_termination": 0})
    with pytest.raises(abi.RvkError) as e:
        abi.render_drr(gpu_ctx, v, view, m, d, out, frame=0)
    assert e.value.status == 5  # not ready: frame has no data
    with pytest.raises(abi.RvkError) as e:
        abi.render_drr(gpu_ctx, v, view, m, d, out, frame=7)
    assert e.value.status == 1
    v.set_frame(1, np.zeros((8, 8, 8), np.uint8))
    abi.render_drr(gpu_ctx, v, view, m, d, out, frame=1)
    torch.cuda.synchronize()
    assert torch.all(out == 1.0)
    with pytest.raises(abi.RvkError):
        abi.Volume(gpu_ctx, (8, 8, 1), (1, 1, 1), "u8")
    with pytest.raises(abi.RvkError):
        abi.render_drr(gpu_ctx, v, view, abi.MarchParams(0.0, 0.0, 1, 1, 1), d, out, frame=1)


# ---------------------------------------------------------------------------
# the compositing benchmark's sample count: the steps-taken side output of a
# plain march (skip rule off) summed over pixels
@pytest.mark.parametrize("which", ["c1", "c3"])
def test_plain_march_sample_count(gpu_ctx, which):
    s = configs.scene_c1(scale=0.25) if which == "c1" else _scene_c3_small()
    march = dict(s.march, skip_rule=0)
    vh = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    vd = configs.view_dict(s.views[0])
    if which == "c1":
        _, (depth, _, _) = H.gpu_1d(gpu_ctx, vh, s.tf1d, s.views[0], march, with_aux=True)
        n64 = orc.count_samples(s.volume, s.spacing, vd, march, tf1d=s.tf1d)
    else:
        gmax = s.tf2d["gradmag_max"]
        vh.compute_gradmag(0, gmax)
        _, (depth, _, _) = H.gpu_2d(gpu_ctx, vh, s.tf2d, s.views[0], march, use_gradmag=True, with_aux=True)
        gm = H.gradmag_u8(s.volume, s.spacing, gmax)
        tf = dict(s.tf2d, gradmag=gm, gradmag_scale=float(np.float32(gmax) / np.float32(255.0)))
        n64 = orc.count_samples(s.volume, s.spacing, vd, march, tf2d=tf)
    vh.close()
    n_gpu = int(depth.astype(np.float64).sum())
    print(f"{which}: plain-march samples GPU {n_gpu}, FP64 oracle {n64}")
    assert n64 > 0 and abs(n_gpu - n64) <= 1e-3 * n64


def test_drr_sample_count_against_fp64_oracle(gpu_ctx):
    """The DRR benchmark's sample count, rvk_count_ray_steps (geometry only:
    the sum over pixels of floor(numSteps - offset)), against the FP64
    oracle's plain-march count on the c2 geometry at half size, and the
    samples the skip-rule march evaluates (FP32 oracle) below it."""
    s = configs.scene_c2(scale=0.5)
    vh = H.gpu_volume(gpu_ctx, s.volume, s.spacing)
    view = s.views[0]
    n_gpu = abi.count_ray_steps(gpu_ctx, vh, view, configs.march_params(s.march))
    vh.close()
    vd = configs.view_dict(view)
    n64 = orc.count_samples(s.volume, s.spacing, vd, dict(s.march, skip_rule=0))
    n_eval = orc.count_drr_evaluated(s.volume, s.spacing, vd, s.march, s.drr)
    print(f"c2/2: plain-march samples GPU {n_gpu}, FP64 oracle {n64}, skip-rule evaluated {n_eval}")
    assert n64 > 0 and abs(n_gpu - n64) <= 1e-5 * n64
    assert 0 < n_eval <= n64


# ---------------------------------------------------------------------------
# cell order: plain x-fastest and 4x4x4 bricks give identical renders
def test_cell_layouts_render_identically(gpu_ctx):
    import torch
    outs = {}
    for layout in (abi.CELLS_PLAIN, abi.CELLS_BRICK4, abi.CELLS_PAIR8):
        s2 = configs.scene_c2(scale=0.25)
        v2 = abi.Volume.from_array(gpu_ctx, s2.volume, s2.spacing, layout=layout)
        assert v2.layout() == layout
        d = H.gpu_drr(gpu_ctx, v2, s2.views[0], s2.march, s2.drr)
        s1 = configs.scene_c1(scale=0.25)
        v1 = abi.Volume.from_array(gpu_ctx, s1.volume, s1.spacing, layout=layout)
        c1 = H.gpu_1d(gpu_ctx, v1, s1.tf1d, s1.views[0], dict(s1.march, term_T=1.0 / 64))
        s3 = _scene_c3_small()
        v3 = abi.Volume.from_array(gpu_ctx, s3.volume, s3.spacing, layout=layout)
        v3.compute_gradmag(0, s3.tf2d["gradmag_max"])
        c3 = H.gpu_2d(gpu_ctx, v3, s3.tf2d, s3.views[0], s3.march, use_gradmag=True)
        c3b = H.gpu_2d(gpu_ctx, v3, s3.tf2d, s3.views[0], s3.march, use_gradmag=False)
        outs[layout] = (d, c1, c3, c3b)
        for v in (v1, v2, v3):
            v.close()
    for layout in (abi.CELLS_BRICK4, abi.CELLS_PAIR8):
        for a, b in zip(outs[abi.CELLS_PLAIN], outs[layout]):
            assert np.array_equal(a, b), layout
    assert outs[abi.CELLS_PLAIN][2][..., 3].max() > 100


def test_relayout_renders_identically(gpu_ctx):
    """rvk_volume_relayout re-encodes loaded frames (u8, int16 in x-pair
    binary16, u16 with its gradient-magnitude volume, float32 differences)
    in every other layout; renders stay bit-identical to a fresh upload."""
    s2 = configs.scene_c2(scale=0.25)
    s1 = configs.scene_c1(scale=0.25)
    s3 = _scene_c3_small()
    rng = np.random.default_rng(11)
    f32 = rng.normal(0.0, 300.0, s2.volume.shape).astype(np.float32)

    def render(v2, v1, v3, vf):
        return (H.gpu_drr(gpu_ctx, v2, s2.views[0], s2.march, s2.drr),
                H.gpu_1d(gpu_ctx, v1, s1.tf1d, s1.views[0], dict(s1.march, term_T=1.0 / 64)),
                H.gpu_2d(gpu_ctx, v3, s3.tf2d, s3.views[0], s3.march, use_gradmag=True),
                H.gpu_drr(gpu_ctx, vf, s2.views[0], s2.march, s2.drr))

    def upload(layout):
        v2 = abi.Volume.from_array(gpu_ctx, s2.volume, s2.spacing, layout=layout)
        v1 = abi.Volume.from_array(gpu_ctx, s1.volume, s1.spacing, layout=layout)
        v3 = abi.Volume.from_array(gpu_ctx, s3.volume, s3.spacing, layout=layout)
        v3.compute_gradmag(0, s3.tf2d["gradmag_max"])
        vf = abi.Volume.from_array(gpu_ctx, f32, s2.spacing, layout=layout)
        return v2, v1, v3, vf

    vols = upload(abi.CELLS_PLAIN)
    ref = render(*vols)
    for layout in (abi.CELLS_PAIR8, abi.CELLS_BRICK4, abi.CELLS_PLAIN, abi.CELLS_PAIR8):
        for v in vols:
            v.relayout(layout)
            assert v.layout() == layout
        for a, b in zip(ref, render(*vols)):
            assert np.array_equal(a, b), layout
    for v in vols:
        v.close()
    assert ref[2][..., 3].max() > 100
    # a layout beyond the volume's addressing limits is refused and the
    # volume keeps its frames and layout
    thin = abi.Volume(gpu_ctx, (2, 4096, 4096), (1.0, 1.0, 1.0), np.uint8)
    thin.set_frame(0, np.zeros((4096, 4096, 2), np.uint8))
    with pytest.raises(abi.RvkError):
        thin.relayout(abi.CELLS_PLAIN)
    assert thin.layout() == abi.CELLS_BRICK4
    thin.relayout(abi.CELLS_PAIR8)
    assert thin.layout() == abi.CELLS_PAIR8
    thin.close()


def test_cell_layout_rules(gpu_ctx):
    small = abi.Volume(gpu_ctx, (64, 64, 64), (1.0, 1.0, 1.0), np.int16)
    assert small.layout() == abi.CELLS_PLAIN          # 35 MB of cells
    big = abi.Volume(gpu_ctx, (512, 512, 512), (1.0, 1.0, 1.0), np.uint16)
    assert big.layout() == abi.CELLS_BRICK4           # 2.2 GB of cells
    big.set_layout(abi.CELLS_PLAIN)                   # allowed before data
    assert big.layout() == abi.CELLS_PLAIN
    with pytest.raises(abi.RvkError):
        big.set_layout(7)
    small.set_frame(0, np.zeros((64, 64, 64), np.int16))
    with pytest.raises(abi.RvkError):
        small.set_layout(abi.CELLS_BRICK4)           # fixed once a frame has data
    # x-fastest cells address (nz + 2)(ny + 2) cell rows with 24-bit
    # multiplies: a thin volume with more rows defaults to bricks
    thin = abi.Volume(gpu_ctx, (2, 4096, 4096), (1.0, 1.0, 1.0), np.uint8)
    assert thin.layout() == abi.CELLS_BRICK4          # 0.5 GB of cells, 4098^2 rows
    with pytest.raises(abi.RvkError):
        thin.set_layout(abi.CELLS_PLAIN)
    thin.close()
    small.close()
    big.close()
