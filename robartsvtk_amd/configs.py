"""Seeded workloads for the five casting configurations.

c1  1D-TF compositing: 128^3 u8 blobs, 256-entry TF, 512^2 RGBA8, step 0.5, stop at T < 0.01
c2  DRR:               256^3 i16 CT phantom, cone beam SID 1000 / SOD 600, 1024^2 f32 exp(-L), step 0.5
c3  2D-TF compositing: 512^3 u16 blobs+Perlin with u8 gradient magnitude, 256^2 TF, 1920x1080 RGBA8
c4  batched DRR:       512x512x400 i16 phantom, 64 rigid poses, detectors of 256..1024 px (409.6 mm FOV)
c5  registration step: 256^3 f32 mu, 16 views of 512^2, perturbed poses, 1 - NCC loss

`scale` shrinks volume and image sizes (tests use small scales; 1.0 is the
benchmarked size).  Everything is derived from fixed seeds through the
library's own PRNG, so tests, benchmark and oracle see the same bits.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

SID = 1000.0  # source to detector, mm
SOD = 600.0   # source to isocentre, mm
FOV = 409.6   # detector side, mm (1024 px at 0.4 mm)


@dataclass
class Scene:
    name: str
    volume: np.ndarray            # [z, y, x]
    spacing: tuple
    views: list                   # abi.View
    march: dict                   # step_scale, term_T, jitter, skip_rule
    drr: dict | None = None       # intercept, slope, offset, output, ref_termination, tint
    tf1d: dict | None = None      # alpha, galpha, r, g, b, params
    tf2d: dict | None = None      # tables, params, gradmag_max
    extra: dict = field(default_factory=dict)

    @property
    def dims(self):
        return (self.volume.shape[2], self.volume.shape[1], self.volume.shape[0])


def march_params(m: dict) -> abi.MarchParams:
    return abi.MarchParams(float(m["step_scale"]), float(m["term_T"]), int(m["jitter"]), int(m["skip_rule"]),
                           int(m.get("precise", 1)))


def drr_params(d: dict) -> abi.DrrParams:
    p = abi.DrrParams()
    p.ct_intercept, p.ct_slope, p.ct_offset = d["intercept"], d["slope"], d["offset"]
    p.output, p.ref_termination = d["output"], d["ref_termination"]
    p.tint[:] = list(d.get("tint", (0, 0, 0, 0)))
    return p


def tf1d_params(p: dict) -> abi.Tf1dParams:
    return abi.Tf1dParams(*[float(p[k]) for k in ("intensity_low", "intensity_multiplier", "gradient_low",
                                                   "gradient_multiplier", "ambient", "diffuse", "specular",
                                                   "specular_power")])


def tf2d_params(p: dict) -> abi.Tf2dParams:
    return abi.Tf2dParams(*[float(p[k]) for k in ("intensity_low", "intensity_multiplier", "gradient_low",
                                                   "gradient_multiplier", "gradient_offset")])


def view_dict(v: abi.View) -> dict:
    """Oracle-side description of a view (plain data)."""
    d = {"m": np.array(v.view_to_voxels[:], np.float32), "W": v.width, "H": v.height, "flipped": v.flipped,
         "clip": [tuple(v.clip_planes[4 * i:4 * i + 4]) for i in range(v.n_clip_planes)],
         "keyhole": [tuple(v.keyhole_planes[4 * i:4 * i + 4]) for i in range(v.n_keyhole_planes)]}
    return d


def _rng(seed: int, n: int) -> np.ndarray:
    return abi.synth_uniform(seed, n)


def _unit_dir(u1: float, u2: float) -> np.ndarray:
    z = 2.0 * u1 - 1.0
    t = 2.0 * math.pi * u2
    r = math.sqrt(max(0.0, 1.0 - z * z))
    return np.array([r * math.cos(t), r * math.sin(t), z])


def _up_for(d: np.ndarray) -> np.ndarray:
    up = np.array([0.0, 0.0, 1.0])
    if abs(float(np.dot(up, d))) > 0.9:
        up = np.array([0.0, 1.0, 0.0])
    return up


def volume_center(dims, spacing) -> np.ndarray:
    return np.array([(dims[a] - 1) * spacing[a] * 0.5 for a in range(3)])


def get_table(xs, ys, lo: float, hi: float, n: int) -> np.ndarray:
    """Piecewise-linear function sampled at lo + i (hi-lo)/(n-1), clamped
    outside its points (the sampling vtkPiecewiseFunction::GetTable does)."""
    x = lo + np.arange(n, dtype=np.float64) * ((hi - lo) / (n - 1))
    return np.interp(x, np.asarray(xs, np.float64), np.asarray(ys, np.float64)).astype(np.float32)


# ---------------------------------------------------------------------------
def scene_c1(scale: float = 1.0, seed: int = 1) -> Scene:
    n = max(8, int(round(128 * scale)))
    img = max(8, int(round(512 * scale)))
    dims = (n, n, n)
    vol = abi.synth_blobs(dims, seed, n_blobs=8, amp=(80.0, 255.0), sigma=(6.0 * scale, 20.0 * scale), noise=10.0,
                          vmax=255.0)
    r = _rng(seed + 1000, 32)
    x1 = 30.0 + 40.0 * r[0]
    x2 = 100.0 + 80.0 * r[1]
    xs = [0.0, x1, x2, 255.0]
    alpha = [0.0, 0.0, 0.02 + 0.18 * r[2], 0.3 + 0.6 * r[3]]
    cols = r[4:16].reshape(4, 3)
    N = 256
    lo, hi = 0.0, 255.0
    tf = {"alpha": get_table(xs, alpha, lo, hi, N), "galpha": np.ones(N, np.float32),
          "r": get_table(xs, cols[:, 0], lo, hi, N), "g": get_table(xs, cols[:, 1], lo, hi, N),
          "b": get_table(xs, cols[:, 2], lo, hi, N),
          "params": {"intensity_low": float(np.float32(lo)), "intensity_multiplier": float(np.float32(1.0 / (hi - lo))),
                     "gradient_low": 0.0, "gradient_multiplier": float("inf"), "ambient": 1.0, "diffuse": 0.0,
                     "specular": 0.0, "specular_power": 1.0}}
    spacing = (1.0, 1.0, 1.0)
    c = volume_center(dims, spacing)
    d = _unit_dir(r[16], r[17])
    dist = 400.0 * scale
    half = 0.5 * math.sqrt(3.0) * n
    cam = abi.camera(c + dist * d, c, _up_for(d), 30.0, (max(1e-3, dist - half - 4), dist + half + 4))
    view = abi.compute_view(cam, img, img, spacing=spacing)
    march = {"step_scale": 0.5, "term_T": 0.01, "jitter": 1, "skip_rule": 1, "precise": 1}
    return Scene("c1", vol, spacing, [view], march, tf1d=tf, extra={"camera": cam})


def drr_camera(center: np.ndarray, angle: float, size_px: int, sod: float = SOD, sid: float = SID, fov: float = FOV,
               half_extent: float = 230.0):
    d = np.array([math.cos(angle), math.sin(angle), 0.0])
    view_angle = 2.0 * math.degrees(math.atan(0.5 * fov / sid))
    return abi.camera(center + sod * d, center, (0.0, 0.0, 1.0), view_angle,
                      (max(1.0, sod - half_extent), sod + half_extent))


def scene_c2(scale: float = 1.0, seed: int = 2) -> Scene:
    n = max(8, int(round(256 * scale)))
    img = max(8, int(round(1024 * scale)))
    dims = (n, n, n)
    spacing = (1.0 / scale if scale < 1.0 else 1.0,) * 3  # keep the physical size (256 mm) at any scale
    vol = abi.synth_ct_phantom(dims, spacing, seed, n_bones=6)
    c = volume_center(dims, spacing)
    r = _rng(seed + 1000, 4)
    cam = drr_camera(c, 2.0 * math.pi * r[0], img, half_extent=0.5 * math.sqrt(3.0) * n * spacing[0] + 4)
    view = abi.compute_view(cam, img, img, spacing=spacing)
    march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 1, "skip_rule": 1, "precise": 1}
    drr = {"intercept": 0.02, "slope": 2e-5, "offset": 1.0, "output": abi.DRR_F32_EXP_NEG_L, "ref_termination": 0,
           "tint": (0, 0, 0, 0)}
    return Scene("c2", vol, spacing, [view], march, drr=drr, extra={"camera": cam})


def tf2d_bumps(seed: int, N: int = 256, n_bumps: int = 5):
    r = _rng(seed, 8 * n_bumps)
    ii, gg = np.meshgrid(np.arange(N) / (N - 1.0), np.arange(N) / (N - 1.0))  # [g, i]
    alpha = np.zeros((N, N))
    col = np.zeros((N, N, 3))
    wsum = np.zeros((N, N))
    for b in range(n_bumps):
        ci, cg = 0.2 + 0.7 * r[8 * b], 0.1 + 0.8 * r[8 * b + 1]
        si, sg = 0.03 + 0.08 * r[8 * b + 2], 0.08 + 0.2 * r[8 * b + 3]
        a = 0.05 + 0.6 * r[8 * b + 4]
        w = np.exp(-0.5 * (((ii - ci) / si) ** 2 + ((gg - cg) / sg) ** 2))
        alpha += a * w
        col += w[..., None] * np.array(r[8 * b + 5:8 * b + 8])[None, None, :]
        wsum += w
    alpha = np.clip(alpha, 0.0, 1.0)
    alpha[alpha < 0.004] = 0.0  # empty space stays exactly empty
    col = col / np.maximum(wsum, 1e-12)[..., None]
    tables = {"alpha": alpha, "red": col[..., 0], "green": col[..., 1], "blue": col[..., 2],
              "ambient": np.ones((N, N)), "diffuse": np.zeros((N, N)), "specular": np.zeros((N, N)),
              "specular_power": np.ones((N, N))}
    return {k: np.ascontiguousarray(v, np.float32) for k, v in tables.items()}


def tf2d_params_for(imin: float, imax: float, gmin: float, gmax: float) -> dict:
    """vtkCuda2DTransferFunctionInformationHandler::UpdateTransferFunction's
    scalars (V/vtkCuda2DTransferFunctionInformationHandler.cxx:209-222)."""
    goff = gmax * 0.8
    gmax_l = (math.log(gmax * gmax + goff) - math.log(goff)) / math.log(2.0) + 1.0
    gmin_l = (math.log(gmin * gmin + goff) - math.log(goff)) / math.log(2.0)
    f32 = lambda x: float(np.float32(x))
    return {"intensity_low": f32(imin), "intensity_multiplier": f32(1.0 / (imax - imin)),
            "gradient_multiplier": f32(1.0 / (gmax_l - gmin_l)), "gradient_offset": f32(goff),
            "gradient_low": f32(-gmin_l - math.log(goff) / math.log(2.0))}


def scene_c3(scale: float = 1.0, seed: int = 3) -> Scene:
    n = max(8, int(round(512 * scale)))
    W, H = max(8, int(round(1920 * scale))), max(8, int(round(1080 * scale)))
    dims = (n, n, n)
    spacing = (1.0 / scale if scale < 1.0 else 1.0,) * 3
    vol = abi.synth_blobs_perlin_u16(dims, seed, n_blobs=5, base_period=48.0 * n / 512.0, octaves=4, vmax=4095.0)
    N = 256
    gmax = 512.0 * scale if scale < 1.0 else 512.0
    tables = tf2d_bumps(seed + 1000, N)
    params = tf2d_params_for(0.0, 4095.0, 0.0, gmax)
    c = volume_center(dims, spacing)
    r = _rng(seed + 2000, 4)
    d = _unit_dir(r[0], r[1])
    ext = n * spacing[0]
    dist = 2.2 * ext
    half = 0.5 * math.sqrt(3.0) * ext
    cam = abi.camera(c + dist * d, c, _up_for(d), 30.0, (dist - half - 4, dist + half + 4))
    view = abi.compute_view(cam, W, H, spacing=spacing)
    march = {"step_scale": 0.5, "term_T": 0.01, "jitter": 1, "skip_rule": 1, "precise": 1}
    tf = {"tables": tables, "params": params, "gradmag_max": gmax}
    return Scene("c3", vol, spacing, [view], march, tf2d=tf, extra={"camera": cam})


def random_poses(seed: int, n: int, rot_deg: float, trans_mm: float) -> np.ndarray:
    r = _rng(seed, 6 * n).reshape(n, 6)
    poses = np.empty((n, 6))
    poses[:, :3] = np.radians((2.0 * r[:, :3] - 1.0) * rot_deg)
    poses[:, 3:] = (2.0 * r[:, 3:] - 1.0) * trans_mm
    return poses


def scene_c4(scale: float = 1.0, seed: int = 4, n_poses: int = 64, with_derivatives: bool = False) -> Scene:
    dims = (max(8, int(round(512 * scale))), max(8, int(round(512 * scale))), max(8, int(round(400 * scale))))
    spacing = (1.0 / scale if scale < 1.0 else 1.0,) * 3
    vol = abi.synth_ct_phantom(dims, spacing, seed, n_bones=6)
    c = volume_center(dims, spacing)
    poses = random_poses(seed + 1000, n_poses, 10.0, 20.0)
    sizes_r = _rng(seed + 2000, n_poses)
    choices = [256, 512, 768, 1024]
    views, sizes, dms, cams = [], [], [], []
    half = 0.5 * math.sqrt(sum((dims[a] * spacing[a]) ** 2 for a in range(3))) + 4
    for i in range(n_poses):
        s = max(4, int(round(choices[min(3, int(sizes_r[i] * 4))] * scale)))
        cam = drr_camera(c, 0.0, s, half_extent=half)
        if with_derivatives:
            v, dm = abi.posed_view(cam, poses[i], c, s, s, spacing=spacing, with_derivatives=True)
            views.append(v)
            dms.append(dm)
        else:
            views.append(abi.posed_view(cam, poses[i], c, s, s, spacing=spacing))
        sizes.append(s)
        cams.append(cam)
    # the batched, differentiable DRR marches without the skip rule: which
    # samples it skips jumps with the pose (the single-view c2 keeps it)
    march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 1, "skip_rule": 0, "precise": 1}
    drr = {"intercept": 0.02, "slope": 2e-5, "offset": 1.0, "output": abi.DRR_F32_EXP_NEG_L, "ref_termination": 0,
           "tint": (0, 0, 0, 0)}
    return Scene("c4", vol, spacing, views, march, drr=drr,
                 extra={"poses": poses, "sizes": sizes, "center": c, "cameras": cams,
                        "dview": np.stack(dms) if with_derivatives else None})


def mu_from_hu(hu: np.ndarray, mu_water: float = 0.02) -> np.ndarray:
    return np.maximum(0.0, mu_water * (1.0 + hu.astype(np.float32) / 1000.0)).astype(np.float32)


def scene_c5(scale: float = 1.0, seed: int = 5, n_views: int = 16, noise_sigma: float = 0.01) -> Scene:
    n = max(8, int(round(256 * scale)))
    img = max(8, int(round(512 * scale)))
    dims = (n, n, n)
    spacing = (1.0 / scale if scale < 1.0 else 1.0,) * 3
    mu = mu_from_hu(abi.synth_ct_phantom(dims, spacing, seed, n_bones=6))
    c = volume_center(dims, spacing)
    half = 0.5 * math.sqrt(3.0) * n * spacing[0] + 4
    angles = [math.pi * i / n_views for i in range(n_views)]
    cams = [drr_camera(c, a, img, half_extent=half) for a in angles]
    gt = random_poses(seed + 1000, n_views, 2.0, 4.0)
    pert = random_poses(seed + 2000, n_views, 5.0, 10.0)
    poses = gt + pert
    march = {"step_scale": 0.5, "term_T": 0.0, "jitter": 1, "skip_rule": 0, "precise": 1}
    drr = {"intercept": 0.0, "slope": 1.0, "offset": 1.0, "output": abi.DRR_F32_EXP_NEG_L, "ref_termination": 0,
           "tint": (0, 0, 0, 0)}
    views, dms = [], []
    for i in range(n_views):
        v, dm = abi.posed_view(cams[i], poses[i], c, img, img, spacing=spacing, with_derivatives=True)
        views.append(v)
        dms.append(dm)
    return Scene("c5", mu, spacing, views, march, drr=drr,
                 extra={"cameras": cams, "gt_poses": gt, "poses": poses, "center": c, "noise_sigma": noise_sigma,
                        "img": img, "dview": np.stack(dms)})


SCENES = {"c1": scene_c1, "c2": scene_c2, "c3": scene_c3, "c4": scene_c4, "c5": scene_c5}
