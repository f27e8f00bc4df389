"""ctypes view of librvkhip.so's C-ABI (include/rvk.h, rvk_geometry.h, rvk_synth.h).

The library is loaded from this package directory only; a missing library is
an error (there is no CPU fallback of the casters).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "librvkhip.so"

RVK_OK = 0
RVK_MAX_FRAMES = 30
DTYPES = {"u8": 0, "i8": 1, "u16": 2, "i16": 3, "u32": 4, "i32": 5, "f32": 6}
NP_DTYPES = {np.dtype(np.uint8): 0, np.dtype(np.int8): 1, np.dtype(np.uint16): 2, np.dtype(np.int16): 3,
             np.dtype(np.uint32): 4, np.dtype(np.int32): 5, np.dtype(np.float32): 6}
DRR_RGBA8_REF, DRR_F32_EXP_NEG_L, DRR_F32_L = 0, 1, 2


class RvkError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"rvk status {status}: {message}")
        self.status = status


class View(C.Structure):
    _fields_ = [("view_to_voxels", C.c_float * 16), ("width", C.c_int32), ("height", C.c_int32),
                ("flipped", C.c_int32), ("n_clip_planes", C.c_int32), ("clip_planes", C.c_float * 24),
                ("n_keyhole_planes", C.c_int32), ("keyhole_planes", C.c_float * 24),
                ("zbuffer", C.c_void_p), ("zb_width", C.c_int32), ("zb_height", C.c_int32)]


class MarchParams(C.Structure):
    _fields_ = [("step_scale", C.c_float), ("term_T", C.c_float), ("jitter", C.c_int32), ("skip_rule", C.c_int32),
                ("precise", C.c_int32)]


class DrrParams(C.Structure):
    _fields_ = [("ct_intercept", C.c_float), ("ct_slope", C.c_float), ("ct_offset", C.c_float),
                ("output", C.c_int32), ("ref_termination", C.c_int32), ("tint", C.c_uint8 * 4)]


class Tf1dParams(C.Structure):
    _fields_ = [("intensity_low", C.c_float), ("intensity_multiplier", C.c_float),
                ("gradient_low", C.c_float), ("gradient_multiplier", C.c_float),
                ("ambient", C.c_float), ("diffuse", C.c_float), ("specular", C.c_float),
                ("specular_power", C.c_float)]


class Tf2dParams(C.Structure):
    _fields_ = [("intensity_low", C.c_float), ("intensity_multiplier", C.c_float),
                ("gradient_low", C.c_float), ("gradient_multiplier", C.c_float),
                ("gradient_offset", C.c_float)]


class DualParams(C.Structure):
    _fields_ = [("intensity1_low", C.c_float), ("intensity1_multiplier", C.c_float),
                ("intensity2_low", C.c_float), ("intensity2_multiplier", C.c_float)]


class AuxOut(C.Structure):
    _fields_ = [("depth", C.c_void_p), ("min_depth", C.c_void_p), ("max_depth", C.c_void_p)]


class ShadeParams(C.Structure):
    _fields_ = [("celr", C.c_float), ("cela", C.c_float), ("celc", C.c_float),
                ("disr", C.c_float), ("disa", C.c_float), ("disc", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("focal_point", C.c_double * 3), ("view_up", C.c_double * 3),
                ("view_angle", C.c_double), ("clipping_range", C.c_double * 2),
                ("window_center", C.c_double * 2), ("parallel_projection", C.c_int32),
                ("parallel_scale", C.c_double), ("use_horizontal_view_angle", C.c_int32)]


class PoseOptimizer(C.Structure):
    _fields_ = [("kind", C.c_int32), ("shared_pose", C.c_int32), ("lr_rotation", C.c_float),
                ("lr_translation", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float)]


OPT_GD, OPT_ADAM = 0, 1
NCC_CENTERED, NCC_UNCENTERED = 0, 1


_P = C.c_void_p
_I = C.c_int
_SIGS = {
    "rvk_last_error_string": (C.c_char_p, []),
    "rvk_status_string": (C.c_char_p, [_I]),
    "rvk_version": (_I, []),
    "rvk_device_count": (_I, [C.POINTER(_I)]),
    "rvk_ctx_create": (_I, [_I, C.POINTER(_P)]),
    "rvk_ctx_create_on_stream": (_I, [_I, _P, C.POINTER(_P)]),
    "rvk_ctx_destroy": (_I, [_P]),
    "rvk_ctx_device": (_I, [_P, C.POINTER(_I)]),
    "rvk_ctx_stream": (_I, [_P, C.POINTER(_P)]),
    "rvk_ctx_set_stream": (_I, [_P, _P]),
    "rvk_ctx_sync": (_I, [_P]),
    "rvk_device_alloc": (_I, [_P, C.c_size_t, C.POINTER(_P)]),
    "rvk_device_free": (_I, [_P, _P]),
    "rvk_copy_to_device": (_I, [_P, _P, _P, C.c_size_t]),
    "rvk_copy_to_host": (_I, [_P, _P, _P, C.c_size_t]),
    "rvk_volume_create": (_I, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_double), _I, _I, C.POINTER(_P)]),
    "rvk_volume_set_frame": (_I, [_P, _I, _P, _I]),
    "rvk_volume_compute_gradmag": (_I, [_P, _I, C.c_float]),
    "rvk_volume_destroy": (_I, [_P]),
    "rvk_volume_set_layout": (_I, [_P, _I]),
    "rvk_volume_get_layout": (_I, [_P, C.POINTER(_I)]),
    "rvk_volume_relayout": (_I, [_P, _I]),
    "rvk_volume_dims": (_I, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(_I), C.POINTER(_I)]),
    "rvk_tf1d_create": (_I, [_P, _I, _P, _P, _P, _P, _P, C.POINTER(Tf1dParams), C.POINTER(_P)]),
    "rvk_tf1d_destroy": (_I, [_P]),
    "rvk_tf2d_create": (_I, [_P, _I, C.POINTER(_P), C.POINTER(_P), C.POINTER(Tf2dParams), C.POINTER(_P)]),
    "rvk_tf2d_destroy": (_I, [_P]),
    "rvk_tf2d_set_inex": (_I, [_P, _P]),
    "rvk_render_1d": (_I, [_P, _P, _I, _P, C.POINTER(View), C.POINTER(MarchParams), _P, C.POINTER(AuxOut)]),
    "rvk_render_2d": (_I, [_P, _P, _I, _P, C.POINTER(View), C.POINTER(MarchParams), _I, _P, C.POINTER(AuxOut)]),
    "rvk_render_drr": (_I, [_P, _P, _I, C.POINTER(View), C.POINTER(MarchParams), C.POINTER(DrrParams), _P]),
    "rvk_render_2d_inex": (_I, [_P, _P, _I, _P, C.POINTER(View), C.POINTER(MarchParams), _I, _P,
                                C.POINTER(AuxOut)]),
    "rvk_render_dual": (_I, [_P, _P, _P, _I, _P, C.POINTER(DualParams), C.POINTER(View), C.POINTER(MarchParams), _P,
                             C.POINTER(AuxOut)]),
    "rvk_post_shade": (_I, [_P, _I, _I, C.POINTER(ShadeParams), _P, _P, _P, _P]),
    "rvk_post_shade_inex": (_I, [_P, _I, _I, C.POINTER(ShadeParams), _P, _P, _P, _P]),
    "rvk_post_shade_noalloc": (_I, [_P, _I, _I, C.POINTER(ShadeParams), _P, _P, _P, _P]),
    "rvk_count_ray_steps": (_I, [_P, _P, C.POINTER(View), C.POINTER(MarchParams), C.POINTER(C.c_int64)]),
    "rvk_stream_read": (_I, [_P, _P, C.c_size_t, _P]),
    "rvk_drr_plan_create": (_I, [_P, _P, _I, _I, C.POINTER(View), _P, C.POINTER(MarchParams), C.POINTER(DrrParams),
                                 C.POINTER(_P)]),
    "rvk_drr_plan_destroy": (_I, [_P]),
    "rvk_drr_plan_update_views": (_I, [_P, C.POINTER(View), _P]),
    "rvk_drr_plan_offsets": (_I, [_P, C.POINTER(C.c_int64)]),
    "rvk_drr_plan_set_schedule": (_I, [_P, _I, _I]),
    "rvk_drr_plan_get_schedule": (_I, [_P, C.POINTER(_I), C.POINTER(_I)]),
    "rvk_drr_plan_autotune": (_I, [_P, _P, _I, C.POINTER(C.c_float)]),
    "rvk_drr_plan_autotune_layout": (_I, [_P, _P, _P, _I, C.POINTER(C.c_float), C.POINTER(_I)]),
    "rvk_drr_plan_set_work_order": (_I, [_P, _I]),
    "rvk_drr_plan_get_work_order": (_I, [_P, C.POINTER(_I)]),
    "rvk_drr_plan_set_tile_shape": (_I, [_P, _I]),
    "rvk_drr_plan_get_tile_shape": (_I, [_P, C.POINTER(_I)]),
    "rvk_drr_plan_render": (_I, [_P, _P]),
    "rvk_drr_plan_capture": (_I, [_P, _P, _P, _P, _P]),
    "rvk_drr_plan_ncc": (_I, [_P, _P, _P, _P, _P]),
    "rvk_drr_plan_backward": (_I, [_P, _P, _P, _P, _P]),
    "rvk_drr_plan_register": (_I, [_P, _P, _P, _P, _P]),
    "rvk_drr_plan_set_ncc": (_I, [_P, _I]),
    "rvk_drr_plan_set_poses": (_I, [_P, _P, C.POINTER(C.c_double), _P, C.POINTER(C.c_double),
                                    C.POINTER(PoseOptimizer)]),
    "rvk_drr_plan_get_poses": (_I, [_P, C.POINTER(C.c_double)]),
    "rvk_drr_plan_step": (_I, [_P, _P, _P, _P, _P]),
    "rvk_drr_plan_capture_step": (_I, [_P, _P, _P, _P, _P]),
    "rvk_compute_view": (_I, [C.POINTER(Camera), _P, C.POINTER(C.c_double), C.POINTER(C.c_double), _I, _I, _I,
                              C.POINTER(View)]),
    "rvk_compute_matrices": (_I, [C.POINTER(Camera), _P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "rvk_prebake_view_to_voxels": (None, [C.POINTER(C.c_double), C.POINTER(C.c_float)]),
    "rvk_pose_matrix": (None, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), _P]),
    "rvk_posed_view": (_I, [C.POINTER(Camera), C.POINTER(C.c_double), C.POINTER(C.c_double),
                            C.POINTER(C.c_double), C.POINTER(C.c_double), _I, _I, _I, C.POINTER(View), _P]),
    "rvk_mat4_mul": (None, [C.POINTER(C.c_double)] * 3),
    "rvk_mat4_invert": (_I, [C.POINTER(C.c_double)] * 2),
    "rvk_synth_blobs": (_I, [C.POINTER(C.c_int32), C.c_uint64, _I, C.c_double, C.c_double, C.c_double, C.c_double,
                             C.c_double, C.c_double, _I, _P]),
    "rvk_synth_ct_phantom": (_I, [C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_uint64, _I, _P]),
    "rvk_synth_blobs_perlin_u16": (_I, [C.POINTER(C.c_int32), C.c_uint64, _I, C.c_double, _I, C.c_double, _P]),
    "rvk_synth_uniform": (None, [C.c_uint64, C.c_int64, C.POINTER(C.c_double)]),
}

_lib = None


def lib() -> C.CDLL:
    """Load librvkhip.so from the package directory (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RvkError(-1, f"{LIB_PATH} is missing: run __graft_entry__.build() (or robartsvtk_amd/build.py)")
        # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7.
        # Loading it first makes this library bind to the same copy (same
        # soname); the other order leaves torch on a second runtime that
        # cannot see the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(_lib, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def check(rc: int) -> None:
    if rc != RVK_OK:
        msg = lib().rvk_last_error_string().decode(errors="replace")
        raise RvkError(rc, msg)


def dptr(t) -> int:
    """Device pointer of a torch tensor (or an int)."""
    return t if isinstance(t, int) else t.data_ptr()


def dvec(v, n=3):
    return (C.c_double * n)(*[float(x) for x in v])


# ---------------------------------------------------------------------------
class Context:
    """rvk_ctx on a device; by default it borrows torch's current stream so
    torch-allocated outputs are ordered with the casters."""

    def __init__(self, device: int = 0, stream: int | None = None, own_stream: bool = False):
        h = C.c_void_p()
        if own_stream:
            check(lib().rvk_ctx_create(device, C.byref(h)))
        else:
            if stream is None:
                import torch
                stream = torch.cuda.current_stream(device).cuda_stream
            check(lib().rvk_ctx_create_on_stream(device, C.c_void_p(stream), C.byref(h)))
        self.handle = h
        self.device = device

    @property
    def stream(self) -> int:
        s = C.c_void_p()
        check(lib().rvk_ctx_stream(self.handle, C.byref(s)))
        return s.value or 0

    def set_stream(self, stream: int | None) -> None:
        check(lib().rvk_ctx_set_stream(self.handle, C.c_void_p(stream) if stream else None))

    def sync(self) -> None:
        check(lib().rvk_ctx_sync(self.handle))

    def close(self) -> None:
        if self.handle:
            lib().rvk_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Volume:
    """rvk_volume: native-type cell volumes (one per frame) on the device."""

    def __init__(self, ctx: Context, dims, spacing, dtype, n_frames: int = 1):
        self.ctx = ctx
        self.dims = tuple(int(d) for d in dims)
        self.spacing = tuple(float(s) for s in spacing)
        self.dtype = DTYPES[dtype] if isinstance(dtype, str) else NP_DTYPES[np.dtype(dtype)]
        h = C.c_void_p()
        check(lib().rvk_volume_create(ctx.handle, (C.c_int32 * 3)(*self.dims), dvec(self.spacing), self.dtype,
                                      n_frames, C.byref(h)))
        self.handle = h

    @classmethod
    def from_array(cls, ctx: Context, arr: np.ndarray, spacing=(1.0, 1.0, 1.0), layout: int | None = None) -> "Volume":
        """arr is indexed [z, y, x] (x fastest in memory, VTK point order);
        layout None keeps the library's default cell order."""
        arr = np.ascontiguousarray(arr)
        v = cls(ctx, (arr.shape[2], arr.shape[1], arr.shape[0]), spacing, arr.dtype)
        if layout is not None:
            v.set_layout(layout)
        v.set_frame(0, arr)
        return v

    def set_frame(self, frame: int, data) -> None:
        if isinstance(data, np.ndarray):
            data = np.ascontiguousarray(data)
            if NP_DTYPES[data.dtype] != self.dtype or data.size != self.dims[0] * self.dims[1] * self.dims[2]:
                raise ValueError("frame data does not match the volume's dtype/size")
            check(lib().rvk_volume_set_frame(self.handle, frame, data.ctypes.data_as(C.c_void_p), 0))
        else:  # torch tensor on the device
            check(lib().rvk_volume_set_frame(self.handle, frame, C.c_void_p(data.data_ptr()), 1))

    def set_layout(self, layout: int) -> None:
        """CELLS_PLAIN, CELLS_BRICK4 or CELLS_PAIR8; before the first set_frame."""
        check(lib().rvk_volume_set_layout(self.handle, int(layout)))

    def relayout(self, layout: int) -> None:
        """Re-encode the loaded frames in another cell layout on the device."""
        check(lib().rvk_volume_relayout(self.handle, int(layout)))

    def layout(self) -> int:
        out = C.c_int()
        check(lib().rvk_volume_get_layout(self.handle, C.byref(out)))
        return out.value

    def compute_gradmag(self, frame: int, max_gradient: float) -> None:
        check(lib().rvk_volume_compute_gradmag(self.handle, frame, float(max_gradient)))

    def close(self) -> None:
        if self.handle:
            lib().rvk_volume_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TF1D:
    def __init__(self, ctx: Context, alpha, rgb, params: Tf1dParams, galpha=None):
        alpha = np.ascontiguousarray(alpha, dtype=np.float32)
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        n = alpha.shape[0]
        r, g, b = (np.ascontiguousarray(rgb[:, i]) for i in range(3))
        ga = None if galpha is None else np.ascontiguousarray(galpha, dtype=np.float32)
        self._keep = (alpha, r, g, b, ga)
        h = C.c_void_p()
        check(lib().rvk_tf1d_create(ctx.handle, n, alpha.ctypes.data, None if ga is None else ga.ctypes.data,
                                    r.ctypes.data, g.ctypes.data, b.ctypes.data, C.byref(params), C.byref(h)))
        self.handle = h
        self.params = params
        self.n = n

    def close(self):
        if self.handle:
            lib().rvk_tf1d_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


TF2D_TABLES = ("alpha", "red", "green", "blue", "ambient", "diffuse", "specular", "specular_power")


class TF2D:
    """tables: dict name -> (n, n) float array indexed [g, i]."""

    def __init__(self, ctx: Context, tables: dict, params: Tf2dParams, keyhole_tables: dict | None = None):
        arrs = [np.ascontiguousarray(tables[k], dtype=np.float32) for k in TF2D_TABLES]
        n = arrs[0].shape[0]
        ptrs = (C.c_void_p * 8)(*[a.ctypes.data for a in arrs])
        karrs, kptrs = None, None
        if keyhole_tables is not None:
            karrs = [np.ascontiguousarray(keyhole_tables[k], dtype=np.float32) for k in TF2D_TABLES]
            kptrs = (C.c_void_p * 8)(*[a.ctypes.data for a in karrs])
        self._keep = (arrs, karrs)
        h = C.c_void_p()
        check(lib().rvk_tf2d_create(ctx.handle, n, ptrs, kptrs, C.byref(params), C.byref(h)))
        self.handle = h
        self.n = n

    def set_inex(self, table) -> None:
        """In/Ex logic table ((n, n) floats indexed [g, i]) for render_2d_inex;
        None removes it."""
        if table is None:
            check(lib().rvk_tf2d_set_inex(self.handle, None))
            return
        t = np.ascontiguousarray(table, dtype=np.float32)
        if t.shape != (self.n, self.n):
            raise ValueError(f"in/ex table must be ({self.n}, {self.n}), got {t.shape}")
        check(lib().rvk_tf2d_set_inex(self.handle, t.ctypes.data))

    def close(self):
        if self.handle:
            lib().rvk_tf2d_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_drr(ctx: Context, vol: Volume, view: View, march: MarchParams, drr: DrrParams, out, frame: int = 0):
    check(lib().rvk_render_drr(ctx.handle, vol.handle, frame, C.byref(view), C.byref(march), C.byref(drr),
                               C.c_void_p(dptr(out))))


def render_1d(ctx: Context, vol: Volume, tf: TF1D, view: View, march: MarchParams, out, aux: AuxOut | None = None,
              frame: int = 0):
    check(lib().rvk_render_1d(ctx.handle, vol.handle, frame, tf.handle, C.byref(view), C.byref(march),
                              C.c_void_p(dptr(out)), C.byref(aux) if aux is not None else None))


def render_2d(ctx: Context, vol: Volume, tf: TF2D, view: View, march: MarchParams, out, use_gradmag: bool = False,
              aux: AuxOut | None = None, frame: int = 0):
    check(lib().rvk_render_2d(ctx.handle, vol.handle, frame, tf.handle, C.byref(view), C.byref(march),
                              int(use_gradmag), C.c_void_p(dptr(out)), C.byref(aux) if aux is not None else None))


def render_2d_inex(ctx: Context, vol: Volume, tf: TF2D, view: View, march: MarchParams, out,
                   black_keyhole: bool = False, aux: AuxOut | None = None, frame: int = 0):
    check(lib().rvk_render_2d_inex(ctx.handle, vol.handle, frame, tf.handle, C.byref(view), C.byref(march),
                                   int(black_keyhole), C.c_void_p(dptr(out)),
                                   C.byref(aux) if aux is not None else None))


def render_dual(ctx: Context, vol1: Volume, vol2: Volume, tf: TF2D, params: DualParams, view: View,
                march: MarchParams, out, aux: AuxOut | None = None, frame: int = 0):
    check(lib().rvk_render_dual(ctx.handle, vol1.handle, vol2.handle, frame, tf.handle, C.byref(params),
                                C.byref(view), C.byref(march), C.c_void_p(dptr(out)),
                                C.byref(aux) if aux is not None else None))


def stream_read(ctx: Context, buf, sink=None) -> None:
    """rvk_stream_read over a torch tensor (device), on the context stream."""
    check(lib().rvk_stream_read(ctx.handle, C.c_void_p(dptr(buf)), buf.numel() * buf.element_size(),
                                None if sink is None else C.c_void_p(dptr(sink))))


def count_ray_steps(ctx: Context, vol: Volume, view: View, march: MarchParams) -> int:
    n = C.c_int64(0)
    check(lib().rvk_count_ray_steps(ctx.handle, vol.handle, C.byref(view), C.byref(march), C.byref(n)))
    return int(n.value)


SCHED_GRID, SCHED_PERSISTENT = 1, 2
TILE_8X8, TILE_16X4, TILE_4X16 = 0, 1, 2
ORDER_VIEWS, ORDER_REGIONS = 0, 1
CELLS_PLAIN, CELLS_BRICK4, CELLS_PAIR8 = 0, 1, 2


class DrrPlan:
    """rvk_drr_plan: many views of one volume frame rendered in one launch,
    plus NCC, the pose/volume adjoint and the fused registration step.
    Buffers are torch tensors (or device pointers); images are packed."""

    def __init__(self, ctx: Context, vol: Volume, views, march: MarchParams, drr: DrrParams, dview=None,
                 frame: int = 0):
        n = len(views)
        arr = (View * n)(*views)
        self._dview = None if dview is None else np.ascontiguousarray(dview, np.float32).reshape(n, 96)
        h = C.c_void_p()
        check(lib().rvk_drr_plan_create(ctx.handle, vol.handle, frame, n, arr,
                                        None if self._dview is None else self._dview.ctypes.data,
                                        C.byref(march), C.byref(drr), C.byref(h)))
        self.handle = h
        self.n_views = n
        self.sizes = [(v.height, v.width) for v in views]
        off = (C.c_int64 * (n + 1))()
        check(lib().rvk_drr_plan_offsets(h, off))
        self.offsets = [int(o) for o in off]
        self.total = self.offsets[-1]

    def update_views(self, views, dview=None):
        arr = (View * len(views))(*views)
        dv = None if dview is None else np.ascontiguousarray(dview, np.float32).reshape(len(views), 96)
        check(lib().rvk_drr_plan_update_views(self.handle, arr, None if dv is None else dv.ctypes.data))

    def set_schedule(self, schedule: int, waves_per_simd: int = 4):
        check(lib().rvk_drr_plan_set_schedule(self.handle, schedule, waves_per_simd))

    def schedule(self):
        s, w = C.c_int(), C.c_int()
        check(lib().rvk_drr_plan_get_schedule(self.handle, C.byref(s), C.byref(w)))
        return s.value, w.value

    def set_work_order(self, order: int):
        check(lib().rvk_drr_plan_set_work_order(self.handle, order))

    def work_order(self) -> int:
        o = C.c_int()
        check(lib().rvk_drr_plan_get_work_order(self.handle, C.byref(o)))
        return o.value

    def set_tile_shape(self, shape: int):
        check(lib().rvk_drr_plan_set_tile_shape(self.handle, shape))

    def tile_shape(self) -> int:
        s = C.c_int()
        check(lib().rvk_drr_plan_get_tile_shape(self.handle, C.byref(s)))
        return s.value

    def autotune(self, out, reps: int = 5) -> float:
        ms = C.c_float()
        check(lib().rvk_drr_plan_autotune(self.handle, C.c_void_p(dptr(out)), reps, C.byref(ms)))
        return ms.value

    def autotune_layout(self, vol: "Volume", out, reps: int = 5):
        """autotune in every cell layout of vol (the plan's volume); returns
        (best ms, CELLS_* chosen)."""
        ms, lay = C.c_float(), C.c_int()
        check(lib().rvk_drr_plan_autotune_layout(self.handle, vol.handle, C.c_void_p(dptr(out)), reps,
                                                 C.byref(ms), C.byref(lay)))
        return ms.value, lay.value

    def render(self, out):
        check(lib().rvk_drr_plan_render(self.handle, C.c_void_p(dptr(out))))

    def capture(self, out=None, targets=None, loss=None, grad_pose=None):
        p = lambda t: None if t is None else C.c_void_p(dptr(t))
        check(lib().rvk_drr_plan_capture(self.handle, p(out), p(targets), p(loss), p(grad_pose)))

    def ncc(self, images, targets, loss, dloss_dimage=None):
        p = lambda t: None if t is None else C.c_void_p(dptr(t))
        check(lib().rvk_drr_plan_ncc(self.handle, p(images), p(targets), p(loss), p(dloss_dimage)))

    def backward(self, images, dloss_dimage, grad_pose=None, grad_volume=None):
        p = lambda t: None if t is None else C.c_void_p(dptr(t))
        check(lib().rvk_drr_plan_backward(self.handle, p(images), p(dloss_dimage), p(grad_pose), p(grad_volume)))

    def register(self, out, targets, loss, grad_pose):
        p = lambda t: C.c_void_p(dptr(t))
        check(lib().rvk_drr_plan_register(self.handle, p(out), p(targets), p(loss), p(grad_pose)))

    def set_ncc(self, mode: int):
        check(lib().rvk_drr_plan_set_ncc(self.handle, mode))

    def set_poses(self, cameras, center, poses, optimizer: PoseOptimizer, origin=None):
        """Device-resident poses (one per view, or one shared): cameras per
        view, rotation centre (mm), poses n x 6 (rx, ry, rz rad, t mm)."""
        cams = (Camera * len(cameras))(*cameras)
        ps = np.ascontiguousarray(poses, np.float64).ravel()
        check(lib().rvk_drr_plan_set_poses(self.handle, cams, dvec(center), None if origin is None else dvec(origin),
                                           ps.ctypes.data_as(C.POINTER(C.c_double)), C.byref(optimizer)))
        self.n_pose_slots = 1 if optimizer.shared_pose else self.n_views

    def poses(self) -> np.ndarray:
        out = np.zeros(6 * self.n_pose_slots, np.float64)
        check(lib().rvk_drr_plan_get_poses(self.handle, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(-1, 6)

    def step(self, out, targets, loss, grad_pose):
        p = lambda t: C.c_void_p(dptr(t))
        check(lib().rvk_drr_plan_step(self.handle, p(out), p(targets), p(loss), p(grad_pose)))

    def capture_step(self, out, targets, loss, grad_pose):
        p = lambda t: C.c_void_p(dptr(t))
        check(lib().rvk_drr_plan_capture_step(self.handle, p(out), p(targets), p(loss), p(grad_pose)))

    def split(self, packed):
        """Per-view [H, W] views of a packed torch tensor."""
        return [packed[self.offsets[v]:self.offsets[v + 1]].view(*self.sizes[v]) for v in range(self.n_views)]

    def close(self):
        if self.handle:
            lib().rvk_drr_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# geometry helpers
def camera(position, focal_point, view_up=(0.0, 1.0, 0.0), view_angle=30.0, clipping_range=(0.1, 1000.0),
           parallel=False, parallel_scale=1.0) -> Camera:
    c = Camera()
    c.position[:] = list(map(float, position))
    c.focal_point[:] = list(map(float, focal_point))
    c.view_up[:] = list(map(float, view_up))
    c.view_angle = float(view_angle)
    c.clipping_range[:] = list(map(float, clipping_range))
    c.window_center[:] = [0.0, 0.0]
    c.parallel_projection = int(parallel)
    c.parallel_scale = float(parallel_scale)
    c.use_horizontal_view_angle = 0
    return c


def compute_view(cam: Camera, width: int, height: int, origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0),
                 user_matrix=None, flipped: bool = False) -> View:
    v = View()
    um = None if user_matrix is None else (C.c_double * 16)(*np.asarray(user_matrix, float).ravel())
    check(lib().rvk_compute_view(C.byref(cam), um, dvec(origin), dvec(spacing), width, height, int(flipped),
                                 C.byref(v)))
    return v


def posed_view(cam: Camera, pose, center, width: int, height: int, origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0),
               flipped: bool = False, with_derivatives: bool = False):
    v = View()
    dm = np.zeros((6, 16), np.float32) if with_derivatives else None
    check(lib().rvk_posed_view(C.byref(cam), dvec(pose, 6), dvec(center), dvec(origin), dvec(spacing), width, height,
                               int(flipped), C.byref(v), None if dm is None else dm.ctypes.data))
    return (v, dm) if with_derivatives else v


def compute_matrices(cam: Camera, aspect: float, origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0), user_matrix=None):
    m = (C.c_double * 16)()
    mi = (C.c_double * 16)()
    um = None if user_matrix is None else (C.c_double * 16)(*np.asarray(user_matrix, float).ravel())
    check(lib().rvk_compute_matrices(C.byref(cam), um, dvec(origin), dvec(spacing), float(aspect), m, mi))
    return np.array(m[:]).reshape(4, 4), np.array(mi[:]).reshape(4, 4)


def pose_matrix(pose, center):
    out = (C.c_double * 16)()
    d = np.zeros((6, 16), np.float64)
    lib().rvk_pose_matrix(dvec(pose, 6), dvec(center), out, d.ctypes.data)
    return np.array(out[:]).reshape(4, 4), d.reshape(6, 4, 4)


def view_matrix(view: View) -> np.ndarray:
    return np.array(view.view_to_voxels[:], np.float32).reshape(4, 4)


# ---------------------------------------------------------------------------
# synthetic inputs
def synth_blobs(dims, seed, n_blobs=8, amp=(80.0, 255.0), sigma=(6.0, 20.0), noise=10.0, vmax=255.0, u16=False):
    out = np.empty((dims[2], dims[1], dims[0]), np.uint16 if u16 else np.uint8)
    check(lib().rvk_synth_blobs((C.c_int32 * 3)(*dims), seed, n_blobs, amp[0], amp[1], sigma[0], sigma[1], noise, vmax,
                                int(u16), out.ctypes.data))
    return out


def synth_ct_phantom(dims, spacing, seed, n_bones=6):
    out = np.empty((dims[2], dims[1], dims[0]), np.int16)
    check(lib().rvk_synth_ct_phantom((C.c_int32 * 3)(*dims), dvec(spacing), seed, n_bones, out.ctypes.data))
    return out


def synth_blobs_perlin_u16(dims, seed, n_blobs=5, base_period=48.0, octaves=4, vmax=4095.0):
    out = np.empty((dims[2], dims[1], dims[0]), np.uint16)
    check(lib().rvk_synth_blobs_perlin_u16((C.c_int32 * 3)(*dims), seed, n_blobs, base_period, octaves, vmax,
                                           out.ctypes.data))
    return out


def synth_uniform(seed, n):
    out = np.empty(n, np.float64)
    lib().rvk_synth_uniform(seed, n, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def device_count() -> int:
    n = C.c_int(0)
    check(lib().rvk_device_count(C.byref(n)))
    return n.value

