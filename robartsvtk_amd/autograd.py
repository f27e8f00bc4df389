"""A thin torch.autograd.Function over the batched DRR plan (gradient harness).

`DrrRenderer` owns a context, a float32 attenuation volume and an
`rvk_drr_plan` for a fixed set of cameras and detector sizes; the rigid poses
(one 6-vector per view: rx, ry, rz in radians, tx, ty, tz in mm, about
`center`, as rvk_pose_matrix defines them) and optionally the volume itself are
the differentiable inputs.

    r = DrrRenderer(volume_zyx, spacing, cameras, sizes, center, march, drr)
    images = drr_render(r, poses)              # packed float32, like the plan
    loss = some_function(images)
    loss.backward()                            # poses.grad from the HIP adjoint

The forward is rvk_drr_plan_render and the backward rvk_drr_plan_backward
(pose adjoint and, when the volume requires grad, the voxel-driven volume
adjoint).  Nothing here computes on the CPU except the per-view camera
matrices and their pose derivatives (rvk_posed_view, double precision).
"""
from __future__ import annotations

import numpy as np
import torch

from . import abi


class DrrRenderer:
    def __init__(self, volume, spacing, cameras, sizes, center, march: abi.MarchParams, drr: abi.DrrParams,
                 ctx: abi.Context | None = None):
        """volume: float32 [z, y, x] numpy array or CUDA tensor (mu per voxel);
        cameras: abi.Camera per view; sizes: (width, height) per view."""
        self.ctx = ctx or abi.Context(0)
        self.spacing = tuple(float(s) for s in spacing)
        self.cameras = list(cameras)
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        self.center = np.asarray(center, np.float64)
        if isinstance(volume, np.ndarray):
            self.vol = abi.Volume.from_array(self.ctx, volume.astype(np.float32, copy=False), self.spacing)
            self.dims_zyx = volume.shape
        else:
            t = volume.detach().contiguous()
            if t.dtype != torch.float32 or not t.is_cuda:
                raise ValueError("volume tensor must be float32 on the HIP device")
            self.dims_zyx = tuple(t.shape)
            self.vol = abi.Volume(self.ctx, (t.shape[2], t.shape[1], t.shape[0]), self.spacing, np.float32)
            self.vol.set_frame(0, t)
        views, dview = self.views_for(np.zeros((len(self.cameras), 6)))
        self.plan = abi.DrrPlan(self.ctx, self.vol, views, march, drr, dview=dview)

    @property
    def n_views(self) -> int:
        return len(self.cameras)

    def views_for(self, poses: np.ndarray):
        views, dms = [], []
        for cam, (w, h), pose in zip(self.cameras, self.sizes, poses):
            v, dm = abi.posed_view(cam, pose, self.center, w, h, spacing=self.spacing, with_derivatives=True)
            views.append(v)
            dms.append(dm)
        return views, np.stack(dms)

    def close(self) -> None:
        self.plan.close()
        self.vol.close()


class DrrFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, poses: torch.Tensor, volume: torch.Tensor | None, renderer: DrrRenderer):
        p = poses.detach().to("cpu", torch.float64).numpy().reshape(renderer.n_views, 6)
        views, dview = renderer.views_for(p)
        renderer.plan.update_views(views, dview)
        if volume is not None:
            renderer.vol.set_frame(0, volume.detach().contiguous())
        out = torch.empty(renderer.plan.total, dtype=torch.float32, device="cuda")
        renderer.plan.render(out)
        ctx.renderer = renderer
        ctx.poses_meta = (poses.dtype, poses.device, poses.shape)
        ctx.has_volume = volume is not None
        ctx.save_for_backward(out)
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (out,) = ctx.saved_tensors
        r: DrrRenderer = ctx.renderer
        want_pose = ctx.needs_input_grad[0]
        want_vol = ctx.has_volume and ctx.needs_input_grad[1]
        gp = torch.empty(6 * r.n_views, dtype=torch.float32, device="cuda") if want_pose else None
        gv = torch.empty(int(np.prod(r.dims_zyx)), dtype=torch.float32, device="cuda") if want_vol else None
        if gp is not None or gv is not None:
            r.plan.backward(out, grad_out.contiguous().to(torch.float32), gp, gv)
        dtype, device, shape = ctx.poses_meta
        g_pose = gp.view(shape).to(device=device, dtype=dtype) if gp is not None else None
        g_vol = gv.view(r.dims_zyx) if gv is not None else None
        return g_pose, g_vol, None


def drr_render(renderer: DrrRenderer, poses: torch.Tensor, volume: torch.Tensor | None = None) -> torch.Tensor:
    """Packed DRR images of all views at `poses` (n_views x 6); differentiable
    in `poses` and, when given, in `volume` (float32 [z, y, x] CUDA tensor
    that replaces the renderer's volume for this call)."""
    return DrrFunction.apply(poses, volume, renderer)
