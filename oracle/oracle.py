"""ctypes wrapper of the CPU oracle (oracle/librvk_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to test against.  The product package never
imports it.

Views are plain dicts so that the oracle stays independent of the product's
structs:  {"m": 16 floats (pre-baked ViewToVoxels), "W", "H", "flipped",
"zbuffer": (h, w) float array or None, "clip": [(a,b,c,d)...], "keyhole": [...]}.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# RVK_ORACLE_LIB selects another build of the same source (the CPU baseline
# times a -march=native build made on the host it runs on)
LIB_PATH = Path(os.environ.get("RVK_ORACLE_LIB", str(HERE / "librvk_oracle.so")))

U8, I8, U16, I16, U32, I32, F32 = range(7)
NP2ORC = {np.dtype(np.uint8): U8, np.dtype(np.int8): I8, np.dtype(np.uint16): U16, np.dtype(np.int16): I16,
          np.dtype(np.uint32): U32, np.dtype(np.int32): I32, np.dtype(np.float32): F32}
DRR_RGBA8_REF, DRR_F32_EXP, DRR_F32_L = 0, 1, 2


class OrcVolume(C.Structure):
    _fields_ = [("data", C.c_void_p), ("dtype", C.c_int), ("dims", C.c_int * 3), ("spacing", C.c_double * 3)]


class OrcView(C.Structure):
    _fields_ = [("m", C.c_float * 16), ("width", C.c_int), ("height", C.c_int), ("flipped", C.c_int),
                ("zbuffer", C.c_void_p), ("zb_width", C.c_int), ("zb_height", C.c_int),
                ("n_clip", C.c_int), ("clip", C.c_float * 24), ("n_keyhole", C.c_int), ("keyhole", C.c_float * 24),
                ("m64", C.c_void_p)]


class OrcMarch(C.Structure):
    _fields_ = [("step_scale", C.c_float), ("term_T", C.c_float), ("jitter", C.c_int), ("skip_rule", C.c_int),
                ("precise", C.c_int), ("f32_control", C.c_int)]


class OrcDrr(C.Structure):
    _fields_ = [("intercept", C.c_float), ("slope", C.c_float), ("offset", C.c_float), ("output", C.c_int),
                ("ref_termination", C.c_int), ("tint", C.c_uint8 * 4)]


class OrcTf1d(C.Structure):
    _fields_ = [("n", C.c_int), ("alpha", C.c_void_p), ("galpha", C.c_void_p), ("r", C.c_void_p),
                ("g", C.c_void_p), ("b", C.c_void_p), ("intensity_low", C.c_float), ("intensity_mult", C.c_float),
                ("gradient_low", C.c_float), ("gradient_mult", C.c_float), ("ambient", C.c_float),
                ("diffuse", C.c_float), ("specular", C.c_float), ("spec_power", C.c_float)]


class OrcTf2d(C.Structure):
    _fields_ = [("n", C.c_int)] + [(k, C.c_void_p) for k in (
        "alpha", "r", "g", "b", "ambient", "diffuse", "specular", "spec_power",
        "k_alpha", "k_r", "k_g", "k_b", "k_ambient", "k_diffuse", "k_specular", "k_spec_power")] + [
        ("intensity_low", C.c_float), ("intensity_mult", C.c_float), ("gradient_low", C.c_float),
        ("gradient_mult", C.c_float), ("gradient_offset", C.c_float), ("gradmag", C.c_void_p),
        ("gradmag_scale", C.c_float)]


class OrcShade(C.Structure):
    _fields_ = [("celr", C.c_float), ("cela", C.c_float), ("celc", C.c_float),
                ("disr", C.c_float), ("disa", C.c_float), ("disc", C.c_float)]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        for name, args, res in [
            ("orc_jitter_table", [], C.POINTER(C.c_float)),
            ("orc_form_rays_f32", [P, P, P, P, P, P], C.c_int),
            ("orc_form_rays_f64", [P, P, P, P, P, P], C.c_int),
            ("orc_drr_f32", [P, P, P, P, P, C.c_int], C.c_int),
            ("orc_drr_f64", [P, P, P, P, P, C.c_int], C.c_int),
            ("orc_1d_f32", [P, P, P, P, P, P, P, C.c_int], C.c_int),
            ("orc_1d_f64", [P, P, P, P, P, P, P, C.c_int], C.c_int),
            ("orc_2d_f32", [P, P, P, P, P, P, P, C.c_int], C.c_int),
            ("orc_2d_f64", [P, P, P, P, P, P, P, C.c_int], C.c_int),
            ("orc_post_shade_f32", [P, C.c_int, C.c_int, P, P, P, P], C.c_int),
            ("orc_post_shade_inex_f32", [P, C.c_int, C.c_int, P, P, P, P], C.c_int),
            ("orc_inex_f32", [P, P, P, C.c_int, P, P, P, P, P, C.c_int], C.c_int),
            ("orc_inex_f64", [P, P, P, C.c_int, P, P, P, P, P, C.c_int], C.c_int),
            ("orc_dual_f32", [P, P, P, P, P, P, P, P, C.c_int], C.c_int),
            ("orc_dual_f64", [P, P, P, P, P, P, P, P, C.c_int], C.c_int),
            ("orc_count_samples", [P, P, P, P, P, C.c_int], C.c_longlong),
            ("orc_count_drr_evaluated", [P, P, P, P, C.c_int], C.c_longlong),
            ("orc_unsafe_rays", [C.c_int], C.c_longlong),
            ("orc_sample_f64", [P, C.c_double, C.c_double, C.c_double], C.c_double),
            ("orc_sample_f32", [P, C.c_float, C.c_float, C.c_float], C.c_float),
            ("orc_tf1d_lookup_f32", [P, C.c_int, C.c_float], C.c_float),
            ("orc_tf2d_lookup_f32", [P, C.c_int, C.c_float, C.c_float], C.c_float),
        ]:
            fn = getattr(_lib, name)
            fn.argtypes = args
            fn.restype = res
    return _lib


def use_library(path) -> None:
    """Load the oracle from another build of the same source (e.g. a
    -march=native build); the next call binds it."""
    global _lib, LIB_PATH
    LIB_PATH = Path(path)
    _lib = None


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


# ---------------------------------------------------------------------------
def _vol(arr: np.ndarray, spacing) -> tuple[OrcVolume, np.ndarray]:
    arr = np.ascontiguousarray(arr)
    v = OrcVolume()
    v.data = arr.ctypes.data
    v.dtype = NP2ORC[arr.dtype]
    v.dims[:] = [arr.shape[2], arr.shape[1], arr.shape[0]]
    v.spacing[:] = [float(s) for s in spacing]
    return v, arr


def _view(view: dict) -> tuple[OrcView, object]:
    v = OrcView()
    v.m[:] = [float(x) for x in np.asarray(view["m"], np.float32).ravel()]
    v.width, v.height = int(view["W"]), int(view["H"])
    v.flipped = int(view.get("flipped", 0))
    keep = None
    if view.get("m64") is not None:
        m64 = np.ascontiguousarray(view["m64"], np.float64).ravel()
        v.m64 = m64.ctypes.data
        keep = [m64]
    zb = view.get("zbuffer")
    if zb is not None:
        zba = np.ascontiguousarray(zb, np.float32)
        keep = (keep or []) + [zba]
        v.zbuffer = zba.ctypes.data
        v.zb_height, v.zb_width = zba.shape
    clip = view.get("clip") or []
    key = view.get("keyhole") or []
    v.n_clip = len(clip)
    v.n_keyhole = len(key)
    for i, p in enumerate(clip):
        v.clip[4 * i:4 * i + 4] = [float(x) for x in p]
    for i, p in enumerate(key):
        v.keyhole[4 * i:4 * i + 4] = [float(x) for x in p]
    return v, keep


def _march(m: dict) -> OrcMarch:
    o = OrcMarch()
    o.step_scale = float(m.get("step_scale", 1.0))
    o.term_T = float(m.get("term_T", 0.0))
    o.jitter = int(m.get("jitter", 1))
    o.skip_rule = int(m.get("skip_rule", 1))
    o.precise = int(m.get("precise", 0))
    o.f32_control = int(m.get("f32_control", 0))
    return o


def _drr(d: dict) -> OrcDrr:
    o = OrcDrr()
    o.intercept = float(d.get("intercept", 0.022))
    o.slope = float(d.get("slope", 0.000022))
    o.offset = float(d.get("offset", 1.0))
    o.output = int(d.get("output", DRR_F32_EXP))
    o.ref_termination = int(d.get("ref_termination", 0))
    o.tint[:] = list(d.get("tint", (0, 0, 0, 0)))
    return o


def _tf1d(tf: dict):
    keep = [np.ascontiguousarray(tf[k], np.float32) for k in ("alpha", "galpha", "r", "g", "b")]
    o = OrcTf1d()
    o.n = keep[0].shape[0]
    o.alpha, o.galpha, o.r, o.g, o.b = (a.ctypes.data for a in keep)
    p = tf["params"]
    o.intensity_low, o.intensity_mult = p["intensity_low"], p["intensity_multiplier"]
    o.gradient_low, o.gradient_mult = p["gradient_low"], p["gradient_multiplier"]
    o.ambient, o.diffuse, o.specular, o.spec_power = p["ambient"], p["diffuse"], p["specular"], p["specular_power"]
    return o, keep


TF2D_NAMES = ("alpha", "red", "green", "blue", "ambient", "diffuse", "specular", "specular_power")
_TF2D_FIELDS = ("alpha", "r", "g", "b", "ambient", "diffuse", "specular", "spec_power")


def _tf2d(tf: dict):
    o = OrcTf2d()
    keep = [np.ascontiguousarray(tf["tables"][k], np.float32) for k in TF2D_NAMES]
    o.n = keep[0].shape[0]
    for f, a in zip(_TF2D_FIELDS, keep):
        setattr(o, f, a.ctypes.data)
    if tf.get("keyhole_tables") is not None:
        kk = [np.ascontiguousarray(tf["keyhole_tables"][k], np.float32) for k in TF2D_NAMES]
        keep += kk
        for f, a in zip(_TF2D_FIELDS, kk):
            setattr(o, "k_" + f, a.ctypes.data)
    p = tf["params"]
    o.intensity_low, o.intensity_mult = p["intensity_low"], p["intensity_multiplier"]
    o.gradient_low, o.gradient_mult, o.gradient_offset = p["gradient_low"], p["gradient_multiplier"], p["gradient_offset"]
    if tf.get("gradmag") is not None:
        gm = np.ascontiguousarray(tf["gradmag"], np.uint8)
        keep.append(gm)
        o.gradmag = gm.ctypes.data
        o.gradmag_scale = float(tf["gradmag_scale"])
    return o, keep


# ---------------------------------------------------------------------------
def jitter_table() -> np.ndarray:
    p = lib().orc_jitter_table()
    return np.array([p[i] for i in range(256)], np.float32)


def drr(vol: np.ndarray, spacing, view: dict, march: dict, drr_params: dict, precision: str = "f32",
        nthreads: int | None = None) -> np.ndarray:
    v, _a = _vol(vol, spacing)
    vw, _z = _view(view)
    m, d = _march(march), _drr(drr_params)
    W, H = vw.width, vw.height
    if d.output == DRR_RGBA8_REF:
        out = np.zeros((H, W, 4), np.uint8)
    else:
        out = np.zeros((H, W), np.float32 if precision == "f32" else np.float64)
    fn = lib().orc_drr_f32 if precision == "f32" else lib().orc_drr_f64
    rc = fn(C.byref(v), C.byref(vw), C.byref(m), C.byref(d), out.ctypes.data, nthreads or default_threads())
    assert rc == 0, rc
    return out


def cast1d(vol: np.ndarray, spacing, tf: dict, view: dict, march: dict, precision: str = "f32",
           nthreads: int | None = None):
    """Returns (rgba8 [H,W,4], rgba float [H,W,4], depth [H,W])."""
    v, _a = _vol(vol, spacing)
    vw, _z = _view(view)
    t, _k = _tf1d(tf)
    m = _march(march)
    W, H = vw.width, vw.height
    ft = np.float32 if precision == "f32" else np.float64
    rgba8 = np.zeros((H, W, 4), np.uint8)
    rgbaf = np.zeros((H, W, 4), ft)
    depth = np.zeros((H, W), ft)
    fn = lib().orc_1d_f32 if precision == "f32" else lib().orc_1d_f64
    rc = fn(C.byref(v), C.byref(t), C.byref(vw), C.byref(m), rgba8.ctypes.data, rgbaf.ctypes.data,
            depth.ctypes.data, nthreads or default_threads())
    assert rc == 0, rc
    return rgba8, rgbaf, depth


def cast2d(vol: np.ndarray, spacing, tf: dict, view: dict, march: dict, precision: str = "f32",
           nthreads: int | None = None):
    v, _a = _vol(vol, spacing)
    vw, _z = _view(view)
    t, _k = _tf2d(tf)
    m = _march(march)
    W, H = vw.width, vw.height
    ft = np.float32 if precision == "f32" else np.float64
    rgba8 = np.zeros((H, W, 4), np.uint8)
    rgbaf = np.zeros((H, W, 4), ft)
    depth = np.zeros((H, W), ft)
    fn = lib().orc_2d_f32 if precision == "f32" else lib().orc_2d_f64
    rc = fn(C.byref(v), C.byref(t), C.byref(vw), C.byref(m), rgba8.ctypes.data, rgbaf.ctypes.data,
            depth.ctypes.data, nthreads or default_threads())
    assert rc == 0, rc
    return rgba8, rgbaf, depth


def cast2d_inex(vol: np.ndarray, spacing, tf: dict, inex: np.ndarray, black_keyhole: bool, view: dict, march: dict,
                precision: str = "f32", nthreads: int | None = None):
    """In/Ex-logic 2D caster; tf as for cast2d (no keyhole tables), inex the
    (n, n) in/ex table.  Returns (rgba8, rgba float, depth)."""
    v, _a = _vol(vol, spacing)
    vw, _z = _view(view)
    t, _k = _tf2d(tf)
    ie = np.ascontiguousarray(inex, np.float32)
    m = _march(march)
    W, H = vw.width, vw.height
    ft = np.float32 if precision == "f32" else np.float64
    rgba8 = np.zeros((H, W, 4), np.uint8)
    rgbaf = np.zeros((H, W, 4), ft)
    depth = np.zeros((H, W), ft)
    fn = lib().orc_inex_f32 if precision == "f32" else lib().orc_inex_f64
    rc = fn(C.byref(v), C.byref(t), ie.ctypes.data, int(black_keyhole), C.byref(vw), C.byref(m), rgba8.ctypes.data,
            rgbaf.ctypes.data, depth.ctypes.data, nthreads or default_threads())
    assert rc == 0, rc
    return rgba8, rgbaf, depth


def cast_dual(vol1: np.ndarray, vol2: np.ndarray, spacing, tf: dict, view: dict, march: dict, precision: str = "f32",
              nthreads: int | None = None):
    """Dual-image caster over two co-registered volumes; tf["params"] gives
    intensity_low/multiplier (component 1) and gradient_low/multiplier
    (component 2).  Returns (rgba8, rgba float, depth)."""
    v1, _a1 = _vol(vol1, spacing)
    v2, _a2 = _vol(vol2, spacing)
    vw, _z = _view(view)
    t, _k = _tf2d(tf)
    m = _march(march)
    W, H = vw.width, vw.height
    ft = np.float32 if precision == "f32" else np.float64
    rgba8 = np.zeros((H, W, 4), np.uint8)
    rgbaf = np.zeros((H, W, 4), ft)
    depth = np.zeros((H, W), ft)
    fn = lib().orc_dual_f32 if precision == "f32" else lib().orc_dual_f64
    rc = fn(C.byref(v1), C.byref(v2), C.byref(t), C.byref(vw), C.byref(m), rgba8.ctypes.data, rgbaf.ctypes.data,
            depth.ctypes.data, nthreads or default_threads())
    assert rc == 0, rc
    return rgba8, rgbaf, depth


def form_rays(vol_dims, spacing, view: dict, march: dict, precision: str = "f64"):
    """Returns dict of per-pixel arrays (start, inc [H,W,3], num_steps, exclude_start/end, min/max depth)."""
    dummy = np.zeros((1, 1, 1), np.uint8)
    v, _a = _vol(dummy, spacing)
    v.dims[:] = list(vol_dims)
    vw, _z = _view(view)
    m = _march(march)
    W, H = vw.width, vw.height
    ft = np.float32 if precision == "f32" else np.float64
    out9 = np.zeros((9, H, W), ft)
    mind = np.zeros((H, W), ft)
    maxd = np.zeros((H, W), ft)
    fn = lib().orc_form_rays_f32 if precision == "f32" else lib().orc_form_rays_f64
    rc = fn(C.byref(v), C.byref(vw), C.byref(m), out9.ctypes.data, mind.ctypes.data, maxd.ctypes.data)
    assert rc == 0, rc
    return {"start": np.moveaxis(out9[0:3], 0, -1), "inc": np.moveaxis(out9[3:6], 0, -1), "num_steps": out9[6],
            "exclude_start": out9[7], "exclude_end": out9[8], "min_depth": mind, "max_depth": maxd}


def count_drr_evaluated(vol: np.ndarray, spacing, view: dict, march: dict, drr_params: dict,
                        nthreads: int | None = None) -> int:
    v, _a = _vol(vol, spacing)
    vw, _z = _view(view)
    m, d = _march(march), _drr(drr_params)
    n = lib().orc_count_drr_evaluated(C.byref(v), C.byref(vw), C.byref(m), C.byref(d), nthreads or default_threads())
    assert n >= 0
    return int(n)


def count_samples(vol: np.ndarray, spacing, view: dict, march: dict, tf1d: dict | None = None,
                  tf2d: dict | None = None, nthreads: int | None = None) -> int:
    v, _a = _vol(vol, spacing)
    vw, _z = _view(view)
    m = _march(march)
    t1 = t2 = None
    if tf1d is not None:
        t1, _k1 = _tf1d(tf1d)
    if tf2d is not None:
        t2, _k2 = _tf2d(tf2d)
    n = lib().orc_count_samples(C.byref(v), C.byref(vw), C.byref(m), C.byref(t1) if t1 else None,
                                C.byref(t2) if t2 else None, nthreads or default_threads())
    assert n >= 0
    return int(n)


def post_shade(shade: dict, depth: np.ndarray, min_depth: np.ndarray, max_depth: np.ndarray, rgba8: np.ndarray):
    s = OrcShade(*[float(shade[k]) for k in ("celr", "cela", "celc", "disr", "disa", "disc")])
    depth = np.ascontiguousarray(depth, np.float32).copy()
    rgba = np.ascontiguousarray(rgba8, np.uint8).copy()
    H, W = depth.shape
    rc = lib().orc_post_shade_f32(C.byref(s), W, H, depth.ctypes.data,
                                  np.ascontiguousarray(min_depth, np.float32).ctypes.data,
                                  np.ascontiguousarray(max_depth, np.float32).ctypes.data, rgba.ctypes.data)
    assert rc == 0
    return depth, rgba


def post_shade_inex(shade: dict, depth: np.ndarray, min_depth: np.ndarray, max_depth: np.ndarray,
                    rgba8: np.ndarray):
    s = OrcShade(*[float(shade[k]) for k in ("celr", "cela", "celc", "disr", "disa", "disc")])
    depth = np.ascontiguousarray(depth, np.float32).copy()
    rgba = np.ascontiguousarray(rgba8, np.uint8).copy()
    H, W = depth.shape
    rc = lib().orc_post_shade_inex_f32(C.byref(s), W, H, depth.ctypes.data,
                                       np.ascontiguousarray(min_depth, np.float32).ctypes.data,
                                       np.ascontiguousarray(max_depth, np.float32).ctypes.data, rgba.ctypes.data)
    assert rc == 0
    return depth, rgba


def sample(vol: np.ndarray, spacing, p, precision="f64") -> float:
    v, _a = _vol(vol, spacing)
    fn = lib().orc_sample_f64 if precision == "f64" else lib().orc_sample_f32
    return float(fn(C.byref(v), *map(float, p)))


def tf1d_lookup(table: np.ndarray, u: float) -> float:
    t = np.ascontiguousarray(table, np.float32)
    return float(lib().orc_tf1d_lookup_f32(t.ctypes.data, t.shape[0], float(u)))


def tf2d_lookup(table: np.ndarray, u: float, v: float) -> float:
    t = np.ascontiguousarray(table, np.float32)
    return float(lib().orc_tf2d_lookup_f32(t.ctypes.data, t.shape[0], float(u), float(v)))


def unsafe_rays(reset: bool = True) -> int:
    """Rays the precise DRR march took no samples of because their sample
    range left the cell range [0, n + 1] (expected 0 for volume-clipped rays)."""
    return int(lib().orc_unsafe_rays(int(reset)))
