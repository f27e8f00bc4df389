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
(pose adjoint and, when the volume requires grad, the brick-driven volume
adjoint).  The plan holds one set of views and one volume frame; every forward
records which poses and volume it rendered, and its backward loads them back
into the plan when another forward has replaced them since (several renders in
one loss, line searches).  Nothing here computes on the CPU except the
per-view camera matrices and their pose derivatives (rvk_posed_view, double
precision).
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
            base = torch.from_numpy(np.ascontiguousarray(volume, np.float32)).cuda()
        else:
            base = volume.detach()
            if base.dtype != torch.float32 or not base.is_cuda:
                raise ValueError("volume tensor must be float32 on the HIP device")
            base = base.contiguous().clone()
        # the volume renders fall back to when a forward passes none
        self.base_volume = base
        self.dims_zyx = tuple(base.shape)
        self.vol = abi.Volume(self.ctx, (base.shape[2], base.shape[1], base.shape[0]), self.spacing, np.float32)
        self.vol.set_frame(0, base)
        self._loaded_volume = base
        self._loaded_poses = np.zeros((len(self.cameras), 6))
        views, dview = self.views_for(self._loaded_poses)
        self.plan = abi.DrrPlan(self.ctx, self.vol, views, march, drr, dview=dview)
        self.generation = 0  # bumped whenever the plan's views or frame change

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

    def load(self, poses: np.ndarray, volume: torch.Tensor) -> int:
        """Put `poses` (n_views x 6, float64) and `volume` into the plan; returns
        the generation that identifies this state."""
        poses = np.array(poses, np.float64).reshape(self.n_views, 6)
        if not np.array_equal(poses, self._loaded_poses):
            views, dview = self.views_for(poses)
            self.plan.update_views(views, dview)
            self._loaded_poses = poses
            self.generation += 1
        if volume is not self._loaded_volume:
            self.vol.set_frame(0, volume)
            self._loaded_volume = volume
            self.generation += 1
        return self.generation

    def close(self) -> None:
        self.plan.close()
        self.vol.close()


class DrrFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, poses: torch.Tensor, volume: torch.Tensor | None, renderer: DrrRenderer):
        p = poses.detach().to("cpu", torch.float64).numpy().reshape(renderer.n_views, 6).copy()
        vol = renderer.base_volume if volume is None else volume.detach().contiguous()
        gen = renderer.load(p, vol)
        out = torch.empty(renderer.plan.total, dtype=torch.float32, device="cuda")
        renderer.plan.render(out)
        ctx.renderer = renderer
        ctx.poses_meta = (poses.dtype, poses.device, poses.shape)
        ctx.state = (gen, p, vol)
        ctx.has_volume = volume is not None
        ctx.save_for_backward(out)
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (out,) = ctx.saved_tensors
        r: DrrRenderer = ctx.renderer
        gen, p, vol = ctx.state
        if r.generation != gen:
            # another forward has rendered other poses or another volume since
            r.load(p, vol)
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
    rendered in place of the renderer's own volume for this call)."""
    return DrrFunction.apply(poses, volume, renderer)
