"""Shared test plumbing: run a scene through the HIP casters (C-ABI) and
through the CPU oracle, and compare."""
from __future__ import annotations

import numpy as np

from robartsvtk_amd import abi, configs


# ---------------------------------------------------------------------------
# GPU side (C-ABI through ctypes; outputs are torch tensors on the device)
def gpu_volume(ctx, vol: np.ndarray, spacing):
    return abi.Volume.from_array(ctx, vol, spacing)


def gpu_drr(ctx, vol_h, view: abi.View, march: dict, drr: dict) -> np.ndarray:
    import torch
    W, H = view.width, view.height
    if drr["output"] == abi.DRR_RGBA8_REF:
        out = torch.full((H, W, 4), 77, dtype=torch.uint8, device="cuda")
    else:
        out = torch.full((H, W), float("nan"), dtype=torch.float32, device="cuda")
    abi.render_drr(ctx, vol_h, view, configs.march_params(march), configs.drr_params(drr), out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def gpu_1d(ctx, vol_h, tf: dict, view: abi.View, march: dict, with_aux: bool = False):
    import torch
    W, H = view.width, view.height
    tfh = abi.TF1D(ctx, tf["alpha"], np.stack([tf["r"], tf["g"], tf["b"]], 1), configs.tf1d_params(tf["params"]),
                   galpha=tf["galpha"])
    out = torch.full((H, W, 4), 77, dtype=torch.uint8, device="cuda")
    aux = None
    bufs = None
    if with_aux:
        bufs = [torch.zeros((H, W), dtype=torch.float32, device="cuda") for _ in range(3)]
        aux = abi.AuxOut(*[b.data_ptr() for b in bufs])
    abi.render_1d(ctx, vol_h, tfh, view, configs.march_params(march), out, aux)
    torch.cuda.synchronize()
    tfh.close()
    if with_aux:
        return out.cpu().numpy(), [b.cpu().numpy() for b in bufs]
    return out.cpu().numpy()


def gpu_2d(ctx, vol_h, tf: dict, view: abi.View, march: dict, use_gradmag: bool, with_aux: bool = False):
    import torch
    W, H = view.width, view.height
    tfh = abi.TF2D(ctx, tf["tables"], configs.tf2d_params(tf["params"]), tf.get("keyhole_tables"))
    out = torch.full((H, W, 4), 77, dtype=torch.uint8, device="cuda")
    aux = None
    bufs = None
    if with_aux:
        bufs = [torch.zeros((H, W), dtype=torch.float32, device="cuda") for _ in range(3)]
        aux = abi.AuxOut(*[b.data_ptr() for b in bufs])
    abi.render_2d(ctx, vol_h, tfh, view, configs.march_params(march), out, use_gradmag, aux)
    torch.cuda.synchronize()
    tfh.close()
    if with_aux:
        return out.cpu().numpy(), [b.cpu().numpy() for b in bufs]
    return out.cpu().numpy()


def _aux(with_aux, H, W):
    import torch
    if not with_aux:
        return None, None
    bufs = [torch.zeros((H, W), dtype=torch.float32, device="cuda") for _ in range(3)]
    return abi.AuxOut(*[b.data_ptr() for b in bufs]), bufs


def gpu_2d_inex(ctx, vol_h, tf: dict, inex: np.ndarray, black_keyhole: bool, view: abi.View, march: dict,
                with_aux: bool = False):
    import torch
    W, H = view.width, view.height
    tfh = abi.TF2D(ctx, tf["tables"], configs.tf2d_params(tf["params"]))
    tfh.set_inex(inex)
    out = torch.full((H, W, 4), 77, dtype=torch.uint8, device="cuda")
    aux, bufs = _aux(with_aux, H, W)
    abi.render_2d_inex(ctx, vol_h, tfh, view, configs.march_params(march), out, black_keyhole, aux)
    torch.cuda.synchronize()
    tfh.close()
    if with_aux:
        return out.cpu().numpy(), [b.cpu().numpy() for b in bufs]
    return out.cpu().numpy()


def dual_params(p: dict) -> abi.DualParams:
    """Component 1 from the intensity fields, component 2 from the gradient
    fields of an oracle-style params dict (the oracle's dual caster reads them
    so)."""
    return abi.DualParams(p["intensity_low"], p["intensity_multiplier"], p["gradient_low"],
                          p["gradient_multiplier"])


def gpu_dual(ctx, v1_h, v2_h, tf: dict, view: abi.View, march: dict, with_aux: bool = False):
    import torch
    W, H = view.width, view.height
    tfh = abi.TF2D(ctx, tf["tables"], configs.tf2d_params(tf["params"]), tf.get("keyhole_tables"))
    out = torch.full((H, W, 4), 77, dtype=torch.uint8, device="cuda")
    aux, bufs = _aux(with_aux, H, W)
    abi.render_dual(ctx, v1_h, v2_h, tfh, dual_params(tf["params"]), view, configs.march_params(march), out, aux)
    torch.cuda.synchronize()
    tfh.close()
    if with_aux:
        return out.cpu().numpy(), [b.cpu().numpy() for b in bufs]
    return out.cpu().numpy()


# ---------------------------------------------------------------------------
# comparisons
def rel_err(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


def lsb_stats(a: np.ndarray, b: np.ndarray):
    """(fraction of pixels whose every RGBA channel is within 1 LSB, max |diff|)."""
    d = np.abs(a.astype(np.int32) - b.astype(np.int32))
    per_px = d.max(axis=-1)
    return float(np.mean(per_px <= 1)), int(d.max())


def gradmag_u8(vol: np.ndarray, spacing, gmax: float) -> np.ndarray:
    """Host restatement of the u8 gradient-magnitude volume (central
    differences at voxel centres with clamped neighbours, 0.5/spacing)."""
    v = vol.astype(np.float32)
    sr = [np.float32(0.5 / s) for s in spacing]  # x, y, z

    def d(axis_np, s):
        p = np.concatenate([np.take(v, [0], axis_np), v, np.take(v, [-1], axis_np)], axis=axis_np)
        n = v.shape[axis_np]
        hi = np.take(p, np.arange(2, n + 2), axis_np)
        lo = np.take(p, np.arange(0, n), axis_np)
        return (hi - lo) * s

    gx, gy, gz = d(2, sr[0]), d(1, sr[1]), d(0, sr[2])
    q = np.sqrt(gx * gx + gy * gy + gz * gz) * (np.float32(255.0) / np.float32(gmax))
    return np.minimum(np.rint(q), 255.0).astype(np.uint8)
