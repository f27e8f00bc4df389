"""More pins of the CPU oracle (CPU only): the paths the uniform-medium closed
forms leave open.

Every expected value here comes from a closed form of the scene, not from the
oracle's code: fields whose trilinear reconstruction and central differences
are known exactly (a quadratic along z has the exact derivative under
piecewise-linear interpolation, a linear ramp a constant one), rays built
from a hand-made affine view matrix (so each sample position is known), a
textbook segment-against-half-spaces clip, and hand-computed shading factors.
Each test names the mistake it would catch.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from robartsvtk_amd import configs


def affine_view(W, H, S0, U, V, D, flipped=0, **extra):
    """A view whose pixel (x, y) casts the segment S(x, y) + z D, z in [0, 1],
    with S = S0 + vx U + vy V, vx = 1 - x/W (x/W flipped), vy = y/H, in
    voxel coordinates: the baked matrix's columns are U, V, D, S0 and its
    bottom row (0, 0, 0, 1)."""
    m = np.array([[U[0], V[0], D[0], S0[0]],
                  [U[1], V[1], D[1], S0[1]],
                  [U[2], V[2], D[2], S0[2]],
                  [0.0, 0.0, 0.0, 1.0]], np.float64)
    assert np.array_equal(m.astype(np.float32).astype(np.float64), m)  # exact in float32
    d = {"m": m.astype(np.float32).ravel(), "W": W, "H": H, "flipped": flipped, "clip": [], "keyhole": []}
    d.update(extra)
    return d


def pixel_start(view, x, y):
    m = np.asarray(view["m"], np.float64).reshape(4, 4)
    vx = x / view["W"] if view["flipped"] else 1.0 - x / view["W"]
    vy = y / view["H"]
    return m[:3, 0] * vx + m[:3, 1] * vy + m[:3, 3], m[:3, 2]


# ---------------------------------------------------------------------------
# 1D caster: gradient AFTER the step, gradient opacity, Phong-Lambert factor
# (V/CUDA_vtkCuda1DVolumeMapper_renderAlgo.cuh:100-131)
C2 = 0.5  # volume = C2 * k^2 along z (k the voxel index)


def _quadratic_z(dims):
    k = np.arange(dims[2], dtype=np.float64)
    vol = np.broadcast_to((C2 * k * k)[:, None, None], dims[::-1])
    return np.ascontiguousarray(vol, np.float32)


def _grad_z(p_z):
    # piecewise-linear interpolation f of C2 k^2 at voxel coordinate w = p - 0.5
    # has f(w + 1) - f(w - 1) = 4 C2 w exactly; times 0.5 / spacing (spacing 1)
    return 2.0 * C2 * (p_z - 0.5)


def _tf1d_ramp_galpha(n=256):
    # galpha[i] = i / n (exact in float32): galpha(u) = (n u - 0.5) / n in range
    return {"alpha": np.full(n, 0.25, np.float32), "galpha": (np.arange(n) / n).astype(np.float32),
            "r": np.full(n, 0.5, np.float32), "g": np.full(n, 0.25, np.float32), "b": np.full(n, 0.75, np.float32),
            "params": {"intensity_low": 0.0, "intensity_multiplier": 1.0 / 300.0, "gradient_low": 0.0,
                       "gradient_multiplier": 1.0 / 16.0, "ambient": 0.25, "diffuse": 0.5, "specular": 0.125,
                       "specular_power": 2.0}}


def _expect_1d(start, inc, n, post_step=True):
    """RGBA (floats) of the plain march: samples at start + t inc, t < n,
    gradient (0, 0, g(p)) at the next sample (post_step) or at the sample."""
    tf = _tf1d_ramp_galpha()
    P = tf["params"]
    col = np.array([0.5, 0.25, 0.75])
    T = 1.0
    rgb = np.zeros(3)
    ray_len = float(np.linalg.norm(inc))
    for t in range(n):
        pz = start[2] + (t + (1 if post_step else 0)) * inc[2]
        gm = abs(_grad_z(pz))
        ga = (256.0 * (P["gradient_multiplier"] * gm) - 0.5) / 256.0
        alpha = 0.25 * ga
        pl = min(1.0, abs(_grad_z(pz) * inc[2]) / (gm * ray_len))
        shade = np.clip((P["ambient"] + P["diffuse"] * pl) * col + P["specular"] * pl ** P["specular_power"], 0, 1)
        rgb += T * alpha * shade
        T *= 1.0 - alpha
    return np.array([*np.clip(rgb, 0, 1), 1.0 - T])


@pytest.mark.parametrize("direction", ["along_gradient", "across_gradient", "oblique"])
def test_1d_gradient_after_step_gradient_opacity_and_shading(direction):
    """Catches: the gradient taken before the step, a missing 0.5/spacing,
    gradient opacity looked up at the wrong |grad|, the Phong-Lambert factor
    (1 along the gradient, 0 across it, cos 45 deg oblique) and spec * PL^pow."""
    dims = (16, 12, 24)
    vol = _quadratic_z(dims)
    W, H = 8, 6
    if direction == "along_gradient":
        view = affine_view(W, H, (3.25, 2.75, 3.0), (8, 0, 0), (0, 6, 0), (0, 0, 10))
    elif direction == "across_gradient":
        view = affine_view(W, H, (2.0, 2.75, 6.25), (0, 0, 6), (0, 6, 0), (10, 0, 0))
    else:
        view = affine_view(W, H, (2.0, 2.75, 3.0), (2, 0, 0), (0, 6, 0), (8, 0, 8))
    march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 0, "skip_rule": 0, "precise": 1}
    _, f, depth = orc.cast1d(vol, (1.0, 1.0, 1.0), _tf1d_ramp_galpha(), view, march, "f64")
    rays = orc.form_rays(dims, (1.0, 1.0, 1.0), view, march, "f64")
    worst_mutant = 0.0
    for y in range(H):
        for x in range(W):
            st, inc = rays["start"][y, x], rays["inc"][y, x]
            n = math.floor(rays["num_steps"][y, x])
            assert n >= 16 and depth[y, x] == n
            e = _expect_1d(st, inc, n)
            np.testing.assert_allclose(f[y, x], e, rtol=1e-9, atol=1e-12)
            worst_mutant = max(worst_mutant, float(np.abs(_expect_1d(st, inc, n, post_step=False) - f[y, x]).max()))
    if direction != "across_gradient":  # across it the gradient is the same at every sample
        assert worst_mutant > 1e-3  # a gradient taken before the step would be caught


# ---------------------------------------------------------------------------
# 2D caster: the log-gradient table row (V/CUDA_vtkCuda2DVolumeMapper_renderAlgo.cuh:263-283,
# V/vtkCuda2DTransferFunctionInformationHandler.cxx:209-222) and per-cell shading
N2 = 256
GMAX = 100.0


def _row_of(G):
    """The handler's gradient axis for gmin = 0, goff = 0.8 gmax, in closed
    form: v = log2(1 + G^2/goff) / (log2(1 + gmax^2/goff) + 1)."""
    goff = 0.8 * GMAX
    v = math.log2(1.0 + G * G / goff) / (math.log2(1.0 + GMAX * GMAX / goff) + 1.0)
    assert abs(N2 * v - round(N2 * v)) > 0.01  # not on a row boundary
    return min(N2 - 1, max(0, math.floor(N2 * v)))


def _row_coded_tf():
    t = {k: np.zeros((N2, N2), np.float32) for k in orc.TF2D_NAMES}
    t["alpha"][:] = 0.125
    t["red"][:] = ((np.arange(N2) + 1) / 512.0)[:, None]  # row g -> (g + 1)/512, exact
    t["ambient"][:] = 1.0
    t["specular_power"][:] = 1.0
    return {"tables": t, "params": configs.tf2d_params_for(0.0, 4095.0, 0.0, GMAX)}


@pytest.mark.parametrize("slope", [10.0, 60.0])
def test_2d_log_gradient_row(slope):
    """A ramp of slope s along y has |grad| = s everywhere inside; rays along
    x composite 20 identical samples of the row the handler's axis puts s
    in.  Catches log2|grad| for log2|grad|^2, a wrong offset or scale."""
    dims = (14, 20, 16)
    j = np.arange(dims[1], dtype=np.float64)
    vol = np.ascontiguousarray(np.broadcast_to((slope * j)[None, :, None], dims[::-1]), np.float32)
    W, H = 8, 8
    view = affine_view(W, H, (2.0, 4.25, 3.25), (0, 0, 8), (0, 10, 0), (10, 0, 0))
    march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 0, "skip_rule": 0, "precise": 1}
    _, f, depth = orc.cast2d(vol, (1.0, 1.0, 1.0), _row_coded_tf(), view, march, "f64")
    row = _row_of(slope)
    A = 1.0 - (1.0 - 0.125) ** 20
    assert np.all(depth == 20)
    np.testing.assert_allclose(f[..., 3], A, rtol=1e-12)
    np.testing.assert_allclose(f[..., 0], (row + 1) / 512.0 * A, rtol=1e-12)  # PL = 0: ambient only


def test_2d_log_gradient_row_from_gradient_magnitude_volume():
    """The u8 gradient-magnitude path: |grad| = q * scale, q = 100."""
    dims = (14, 20, 16)
    vol = np.full(dims[::-1], 1000.0, np.float32)
    gm = np.full(dims[::-1], 100, np.uint8)
    scale = float(np.float32(GMAX) / np.float32(255.0))
    tf = dict(_row_coded_tf(), gradmag=gm, gradmag_scale=scale)
    view = affine_view(8, 8, (2.0, 4.25, 3.25), (0, 0, 8), (0, 10, 0), (10, 0, 0))
    march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 0, "skip_rule": 0, "precise": 1}
    _, f, _ = orc.cast2d(vol, (1.0, 1.0, 1.0), tf, view, march, "f64")
    row = _row_of(100.0 * scale)
    A = 1.0 - (1.0 - 0.125) ** 20
    np.testing.assert_allclose(f[..., 0], (row + 1) / 512.0 * A, rtol=1e-12)  # head-light: PL = 1, diffuse 0


def test_2d_per_cell_shading_oblique():
    """Every sample of a 45-degree ray through a ramp along z lands in one
    table row (|grad| = 20) with PL = cos 45 deg; that row's ambient,
    diffuse, specular and power give sat((a + d PL) c + s PL^p) per sample.
    Catches spec * PL for spec * PL^p, diffuse not scaled by PL, or cells of
    another row."""
    dims = (16, 12, 24)
    k = np.arange(dims[2], dtype=np.float64)
    vol = np.ascontiguousarray(np.broadcast_to((20.0 * k)[:, None, None], dims[::-1]), np.float32)
    row = _row_of(20.0)
    t = {n: np.zeros((N2, N2), np.float32) for n in orc.TF2D_NAMES}
    t["alpha"][row] = 0.125
    t["ambient"][row] = 0.25
    t["diffuse"][row] = 0.5
    t["specular"][row] = 0.25
    t["specular_power"][row] = 3.0
    t["red"][row], t["green"][row], t["blue"][row] = 0.5, 0.25, 1.0
    tf = {"tables": t, "params": configs.tf2d_params_for(0.0, 4095.0, 0.0, GMAX)}
    view = affine_view(8, 6, (2.0, 2.75, 3.0), (2, 0, 0), (0, 6, 0), (8, 0, 8))
    march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 0, "skip_rule": 0, "precise": 1}
    _, f, _ = orc.cast2d(vol, (1.0, 1.0, 1.0), tf, view, march, "f64")
    rays = orc.form_rays(dims, (1.0, 1.0, 1.0), view, march, "f64")
    col = np.array([0.5, 0.25, 1.0])
    for y in range(6):
        for x in range(8):
            inc = rays["inc"][y, x]
            n = math.floor(rays["num_steps"][y, x])
            pl = abs(inc[2]) / np.linalg.norm(inc)
            assert pl == pytest.approx(math.sqrt(0.5), rel=1e-12)
            c = np.clip((0.25 + 0.5 * pl) * col + 0.25 * pl ** 3, 0, 1)
            A = 1.0 - 0.875 ** n
            np.testing.assert_allclose(f[y, x, :3], c * A, rtol=1e-9)
            assert f[y, x, 3] == pytest.approx(A, rel=1e-12)


# ---------------------------------------------------------------------------
# user clipping planes (V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:130-197) and
# the z-buffer end depth (:320-389): DRR of a uniform box = mu x the chord
def _clip_segment(S, D, halfspaces):
    """Cyrus-Beck: the part t in [0, 1] of S + t D with a.p + b >= 0 for every
    (a, b), as (t0, t1), or None."""
    t0, t1 = 0.0, 1.0
    for a, b in halfspaces:
        a = np.asarray(a, np.float64)
        num, den = float(a @ S + b), float(a @ D)
        if den == 0.0:
            if num < 0.0:
                return None
            continue
        t = -num / den
        if den > 0.0:
            t0 = max(t0, t)
        else:
            t1 = min(t1, t)
    return (t0, t1) if t1 > t0 else None


def _box(lo, hi):
    return [((1, 0, 0), -lo[0]), ((-1, 0, 0), hi[0]), ((0, 1, 0), -lo[1]), ((0, -1, 0), hi[1]),
            ((0, 0, 1), -lo[2]), ((0, 0, -1), hi[2])]


def _uniform_drr_expect(view, dims, spacing, mu, halfspaces, step, depth_of=None):
    sp = np.asarray(spacing, np.float64)
    out = np.zeros((view["H"], view["W"]))
    ok = np.ones_like(out, bool)
    vol_box = _box((1, 1, 1), (dims[0] - 2, dims[1] - 2, dims[2] - 2))
    for y in range(view["H"]):
        for x in range(view["W"]):
            S, D = pixel_start(view, x, y)
            if depth_of is not None:
                D = D * depth_of(x, y)
            seg = _clip_segment(S, D, halfspaces + vol_box)
            if seg is None:
                continue
            ns = (seg[1] - seg[0]) * float(np.linalg.norm(D * sp)) / step
            ok[y, x] = abs(ns - round(ns)) > 1e-6
            out[y, x] = math.floor(ns) * step * mu
    return out, ok


def test_drr_uniform_box_under_six_clip_planes():
    """Five axis planes and a tilted one keep a convex region; the DRR of a
    uniform volume is mu times the sampled chord through it.  Catches a
    flipped plane orientation, clipping the wrong end of the ray, and the
    reject rule for rays outside a plane."""
    dims, spacing = (24, 20, 28), (1.0, 1.25, 0.75)
    hu = 250
    vol = np.full(dims[::-1], hu, np.int16)
    mu = 0.02 + 2e-5 * hu
    planes = [((1, 0, 0), -6.25), ((-1, 0, 0), 17.5), ((0, 1, 0), -4.125), ((0, -1, 0), 15.625),
              ((0, 0, 1), -5.25), ((-0.25, 0, -1), 22.0)]
    view = affine_view(16, 12, (1.5, 1.75, -2.0), (20, 0, 0), (0, 16, 0), (2.0, 1.5, 32.0),
                       clip=[(*a, b) for a, b in planes])
    drr = {"intercept": 0.02, "slope": 2e-5, "offset": 1.0, "output": orc.DRR_F32_L, "ref_termination": 0}
    march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 0, "skip_rule": 0, "precise": 1}
    L = orc.drr(vol, spacing, view, march, drr, "f64")
    step = 0.5 * min(spacing)
    expect, ok = _uniform_drr_expect(view, dims, spacing, float(np.float32(0.02)) + float(np.float32(2e-5)) * hu,
                                     planes, step)
    assert ok.sum() > 0.9 * ok.size
    hit = expect > 0
    assert 20 < hit.sum() < hit.size - 20  # some rays pass through the region, some miss it
    np.testing.assert_allclose(L[ok], expect[ok], rtol=2e-6, atol=1e-12)
    # without the planes the chords are longer wherever the region cut them
    L0 = orc.drr(vol, spacing, dict(view, clip=[]), march, drr, "f64")
    assert np.all(L0 >= L - 1e-9) and (L0 > L + 1e-3).sum() > 0.5 * hit.sum()
    del mu


@pytest.mark.parametrize("zb_w,flipped", [(16, 0), (8, 0), (16, 1)])
def test_drr_zbuffer_end_depth(zb_w, flipped):
    """Rays end at the depth the z-buffer holds at (1 - vx, vy), point-sampled
    with clamping: with a buffer that rises along x, each pixel's chord tells
    which texel it read.  Catches reading at vx instead of 1 - vx, and the
    texel index rounding."""
    dims, spacing = (22, 20, 28), (1.0, 1.0, 1.0)
    vol = np.full(dims[::-1], 250, np.int16)
    W, H, zb_h = 16, 8, 4
    zi, zj = np.meshgrid(np.arange(zb_w), np.arange(zb_h))
    # exact in float32, and no chord an integer number of steps
    zb = (0.25 + 0.03125 * zi + 0.0625 * zj + 1.0 / 256.0).astype(np.float32)
    view = affine_view(W, H, (2.5, 2.5, 1.0), (16, 0, 0), (0, 14, 0), (0, 0, 24), flipped=flipped, zbuffer=zb)
    drr = {"intercept": 0.02, "slope": 2e-5, "offset": 1.0, "output": orc.DRR_F32_L, "ref_termination": 0}
    march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 0, "skip_rule": 0, "precise": 1}
    L = orc.drr(vol, spacing, view, march, drr, "f64")

    def depth_of(x, y):
        vx = x / W if flipped else 1.0 - x / W
        u, v = 1.0 - vx, y / H
        return float(zb[min(max(math.floor(v * zb_h), 0), zb_h - 1), min(max(math.floor(u * zb_w), 0), zb_w - 1)])

    mu = float(np.float32(0.02)) + float(np.float32(2e-5)) * 250
    expect, ok = _uniform_drr_expect(view, dims, spacing, mu, [], 0.5, depth_of)
    assert ok.all()
    np.testing.assert_allclose(L, expect, rtol=2e-6)
    assert len(np.unique(np.round(L[0], 9))) >= 4  # the texel read changes along x


# ---------------------------------------------------------------------------
# cel and distance shading at non-zero darkness (V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:449-512)
def test_post_shade_hand_values():
    """A 4x3 image with a depth step between columns 1 and 2 (normalised
    depth 0.2 | 0.6).  Hand values, cel (darkness 0.5, start 0.1, 1/range 2)
    and distance (darkness 0.4, start 0.1, 1/range 2.5):
      x=0: |grad| 0   -> cel 1;   distance sat(0.25) -> 0.9;  factor 0.9
      x=1: |grad| 0.4 -> cel 0.7; distance 0.9;                factor 0.63
      x=2: |grad| 0   -> cel 1;   distance sat(1.25) -> 0.6;  factor 0.6
      x=3: clamped neighbour (itself) -> as x=2
    RGB (201, 101, 51) -> truncated products; alpha kept.  Catches the x/y
    neighbours swapped, the distance factor read at a neighbour, min depth
    added with the wrong sign, and rounding instead of truncation."""
    W, H = 4, 3
    mind = np.zeros((H, W), np.float32)
    maxd = np.full((H, W), 10.0, np.float32)
    depth = np.array([[2, 2, 6, 6]] * H, np.float32)
    # one pixel with min depth 1 and one with no ray (max depth 0 -> 1)
    mind[2, 0], depth[2, 0] = 1.0, 1.0         # (1 + 1) / 10 = 0.2 as its neighbours
    maxd[2, 3] = 0.0                            # normalised depth 1
    rgba = np.zeros((H, W, 4), np.uint8)
    rgba[..., 0], rgba[..., 1], rgba[..., 2], rgba[..., 3] = 201, 101, 51, 77
    shade = {"celr": 0.5, "cela": 0.1, "celc": 2.0, "disr": 0.4, "disa": 0.1, "disc": 2.5}
    dn, out = orc.post_shade(shade, depth, mind, maxd, rgba)
    np.testing.assert_allclose(dn[0], [0.2, 0.2, 0.6, 0.6], rtol=1e-6)
    assert dn[2, 3] == 1.0 and dn[2, 0] == pytest.approx(0.2, rel=1e-6)
    # row 1, x = 3: the pixel below has depth 1, |grad| 0.4 -> cel 0.7 -> 0.42
    fac = {(0, 0): 0.9, (1, 0): 0.63, (2, 0): 0.6, (3, 0): 0.6, (0, 1): 0.9, (1, 1): 0.63, (2, 1): 0.6, (3, 1): 0.42}
    for (x, y), f in fac.items():
        expect = [int(201 * f), int(101 * f), int(51 * f), 77]
        assert list(out[y, x]) == expect, (x, y, list(out[y, x]), expect)
    # row 2, x = 2: below is clamped (itself); right is the no-ray pixel at depth 1:
    # |grad| = 0.4 -> cel 0.7; distance 0.6 -> 0.42
    assert list(out[2, 2]) == [int(201 * 0.42), int(101 * 0.42), int(51 * 0.42), 77]
