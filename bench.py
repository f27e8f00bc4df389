"""Benchmark: G ray-samples/s of the gfx950 DRR caster on one MI355X, with the
compositing casters, batched forward+backward and the CPU baseline beside it.

Headline workload (N=1): 256^3 int16 CT phantom, cone beam SID 1000 / SOD 600
mm, 1024x1024 float32 exp(-L) detector, step 0.5 voxel, skip rule on
(config c2).  One step = one full-detector render with the volume resident in
HBM.  `value` = N_samples / (seconds per render) / 1e9, with N_samples the
plain-march ray-sample count (sum over pixels of floor(num_steps - ray
offset)).

`roofline` is the binding resource of the DRR kernel, VALU issue: its VALU
instructions per launch (SQ_INSTS_VALU, committed rocprofv3 counter pass,
profiles/kernel_counters.json) over the live kernel time, against
256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction.  The HBM
traffic the counters measure and the gather-equivalent bytes (8 voxels per
sample) are reported beside it.  `cpu_baseline` times the CPU oracle (FP32,
a -march=native build made on this host, every core of the process's
affinity) on each configuration and reports every configuration's error
against the FP64 oracle.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# wave64 VALU instructions per second: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4
COUNTERS = ROOT / "profiles" / "kernel_counters.json"
DRR_KERNEL = "drr_kernel<Cells<HalfPair,0,true>,true,true,false,false>"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def kernel_counters(kernel: str):
    """The committed counter record of `kernel` (profiles/kernel_counters.json), or None."""
    try:
        rec = json.loads(COUNTERS.read_text())
    except (OSError, ValueError):
        return None
    return rec.get("kernels", {}).get(kernel)


def cpu_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def host_info() -> dict:
    model = platform.processor() or ""
    try:
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cores": cpu_threads()}


def _events(torch):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def _time(fn, reps: int) -> float:
    """Mean milliseconds of fn() over reps runs (after one warm-up), HIP events
    on torch's current stream (the stream the library's context uses)."""
    import torch
    fn()
    torch.cuda.synchronize()
    a, b = _events(torch)
    s = torch.cuda.current_stream()
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


# ---------------------------------------------------------------------------
def stream_read_probe(gib: float = 4.0, reps: int = 10) -> dict:
    """Attainable HBM read bandwidth: rvk_stream_read over a 4 GiB buffer."""
    import torch

    from robartsvtk_amd import abi

    ctx = abi.Context(0)
    n = int(gib * (1 << 30))
    buf = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    ms = _time(lambda: abi.stream_read(ctx, buf), reps)
    del buf
    ctx.close()
    gbs = n / (ms / 1e3) / 1e9
    return {"GB_per_s": round(gbs, 1), "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4), "bytes": n, "ms": round(ms, 4)}


def bench_drr(scene, steps: int, warmup: int, march=None):
    import torch

    from robartsvtk_amd import abi, configs

    ctx = abi.Context(0)
    vol = abi.Volume.from_array(ctx, scene.volume, scene.spacing)
    view = scene.views[0]
    mp = configs.march_params(march or scene.march)
    dp = configs.drr_params(scene.drr)
    out = torch.empty((view.height, view.width), dtype=torch.float32, device="cuda")
    n_samples = abi.count_ray_steps(ctx, vol, view, mp)
    for _ in range(warmup):
        abi.render_drr(ctx, vol, view, mp, dp, out)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [_events(torch) for _ in range(steps)]
    t0, t1 = _events(torch)
    torch.cuda.synchronize()
    t0.record(stream)
    for a, b in ev:
        a.record(stream)
        abi.render_drr(ctx, vol, view, mp, dp, out)
        b.record(stream)
    t1.record(stream)
    torch.cuda.synchronize()
    total_ms = t0.elapsed_time(t1)
    kern_ms = statistics.fmean(a.elapsed_time(b) for a, b in ev)
    img = out.cpu().numpy()
    vol.close()
    ctx.close()
    return {"n_samples": n_samples, "ms_per_step": total_ms / steps, "kernel_ms": kern_ms,
            "pixels": view.width * view.height, "image": img}


def bench_compositing(scene, steps: int, warmup: int, gradmag: bool = True):
    """1D (c1) or 2D (c3) compositing caster.  N_samples is the plain march
    (skip rule off) up to opacity termination: one render with the skip rule
    off writes the per-pixel sample count into the steps-taken side output,
    and its sum is the count.  The timed renders use the configured march
    (skip rule on)."""
    import numpy as np
    import torch

    from robartsvtk_amd import abi, configs

    ctx = abi.Context(0)
    vol = abi.Volume.from_array(ctx, scene.volume, scene.spacing)
    view = scene.views[0]
    W, H = view.width, view.height
    two_d = scene.tf2d is not None
    if two_d:
        tf = scene.tf2d
        if gradmag:
            vol.compute_gradmag(0, tf["gradmag_max"])
        tfh = abi.TF2D(ctx, tf["tables"], configs.tf2d_params(tf["params"]))

        def render(mp, o, aux=None):
            abi.render_2d(ctx, vol, tfh, view, mp, o, gradmag, aux)
    else:
        tf = scene.tf1d
        tfh = abi.TF1D(ctx, tf["alpha"], np.stack([tf["r"], tf["g"], tf["b"]], 1), configs.tf1d_params(tf["params"]),
                       galpha=tf["galpha"])

        def render(mp, o, aux=None):
            abi.render_1d(ctx, vol, tfh, view, mp, o, aux)
    out = torch.empty((H, W, 4), dtype=torch.uint8, device="cuda")
    steps_taken = torch.zeros((H, W), dtype=torch.float32, device="cuda")
    plain = configs.march_params(dict(scene.march, skip_rule=0))
    render(plain, out, abi.AuxOut(steps_taken.data_ptr(), None, None))
    torch.cuda.synchronize()
    n_samples = int(steps_taken.double().sum().item())
    mp = configs.march_params(scene.march)
    for _ in range(warmup):
        render(mp, out)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [_events(torch) for _ in range(steps)]
    for a, b in ev:
        a.record(stream)
        render(mp, out)
        b.record(stream)
    torch.cuda.synchronize()
    kern_ms = statistics.fmean(a.elapsed_time(b) for a, b in ev)
    img = out.cpu().numpy()
    tfh.close()
    vol.close()
    ctx.close()
    return {"n_samples": n_samples, "pixels": W * H, "ms_per_render": round(kern_ms, 4),
            "G_samples_per_s": round(n_samples / (kern_ms / 1e3) / 1e9, 3),
            "covered_pixel_fraction": round(float((img[..., 3] > 0).mean()), 4)}, img


def bench_batched(scene, reps: int = 5, volume_gradient: bool = False, loop: bool = False):
    """Batched forward and forward+backward (registration step) against the
    plain per-pose loop: one single-view launch per pose for the forward, and
    13 single-view renders per pose (1 + 12 central differences of the 6-DoF
    pose) for the gradient.  The plain loop's loss evaluations are left out,
    which favours the plain loop.  volume_gradient: also time the volume
    adjoint; loop: also time the device-resident registration step (render,
    NCC, pose gradient, Adam update and new views, one replayed graph)."""
    import numpy as np
    import torch

    from robartsvtk_amd import abi, configs

    ctx = abi.Context(0)
    vol = abi.Volume.from_array(ctx, scene.volume, scene.spacing)
    mp, dp = configs.march_params(scene.march), configs.drr_params(scene.drr)
    views = scene.views
    n = len(views)
    plan = abi.DrrPlan(ctx, vol, views, mp, dp, dview=scene.extra["dview"])
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    singles = [torch.empty((v.height, v.width), dtype=torch.float32, device="cuda") for v in views]
    n_samples = sum(abi.count_ray_steps(ctx, vol, v, mp) for v in views)

    def plain_fwd():
        for v, o in zip(views, singles):
            abi.render_drr(ctx, vol, v, mp, dp, o)

    def plain_fd():
        for v, o in zip(views, singles):
            for _ in range(13):
                abi.render_drr(ctx, vol, v, mp, dp, o)

    t_plain = _time(plain_fwd, reps)
    t_plain_fd = _time(plain_fd, max(1, reps // 2))
    tuned_ms = plan.autotune(out, reps=3)
    sched = plan.schedule()
    plan.capture(out)
    t_batch = _time(lambda: plan.render(out), reps)
    plan.render(out)
    g = torch.Generator(device="cuda").manual_seed(7)
    tgt = out + 0.01 * torch.randn(out.shape, device="cuda", generator=g)
    loss = torch.empty(n, dtype=torch.float32, device="cuda")
    grad = torch.empty(6 * n, dtype=torch.float32, device="cuda")
    plan.capture(out, tgt, loss, grad)
    t_reg = _time(lambda: plan.register(out, tgt, loss, grad), reps)
    plan.render(out)
    torch.cuda.synchronize()
    assert torch.isfinite(loss).all() and torch.isfinite(grad).all()
    res = {"poses": n, "pixels": int(sum(v.width * v.height for v in views)), "n_samples": n_samples,
           "march": "plain (no skip rule)" if not scene.march["skip_rule"] else "skip rule",
           "plain_fwd_ms": round(t_plain, 3), "batched_fwd_ms": round(t_batch, 3),
           "fwd_speedup": round(t_plain / t_batch, 3),
           "batched_fwd_G_samples_per_s": round(n_samples / (t_batch / 1e3) / 1e9, 2),
           "schedule": {"kind": "persistent" if sched[0] == abi.SCHED_PERSISTENT else "grid",
                        "waves_per_simd": sched[1],
                        "tile": {abi.TILE_8X8: "8x8", abi.TILE_16X4: "16x4", abi.TILE_4X16: "4x16"}[plan.tile_shape()],
                        "work_order": "regions" if plan.work_order() == abi.ORDER_REGIONS else "views",
                        "autotune_ms": round(tuned_ms, 3)},
           "plain_fwd_fd_ms": round(t_plain_fd, 3), "batched_fwd_bwd_ms": round(t_reg, 3),
           "plain_fd_poses_per_s": round(n / (t_plain_fd / 1e3), 2),
           "batched_fwd_bwd_poses_per_s": round(n / (t_reg / 1e3), 2),
           "fwd_bwd_speedup": round(t_plain_fd / t_reg, 3)}
    images = out.cpu().numpy()
    if volume_gradient:
        dldi = torch.empty_like(out)
        plan.ncc(out, tgt, loss, dldi)
        gvol = torch.empty(scene.volume.size, dtype=torch.float32, device="cuda")
        t_vg = _time(lambda: plan.backward(out, dldi, grad_volume=gvol), max(1, reps // 2))
        assert bool(torch.isfinite(gvol).all())
        res["volume_gradient_ms"] = round(t_vg, 3)
        res["volume_gradient_note"] = (f"rvk_drr_plan_backward(grad_volume) alone: ray records + brick-driven "
                                       f"fixed-point scatter over {scene.volume.size} voxels x {n} views")
    if loop:
        opt = abi.PoseOptimizer(kind=abi.OPT_ADAM, shared_pose=0, lr_rotation=0.005, lr_translation=0.5, beta1=0.9,
                                beta2=0.999, eps=1e-8)
        plan.set_poses(scene.extra["cameras"], scene.extra["center"], scene.extra["poses"], opt)
        plan.capture_step(out, tgt, loss, grad)
        t_step = _time(lambda: plan.step(out, tgt, loss, grad), reps)
        res["device_step_ms"] = round(t_step, 3)
        res["device_step_note"] = ("one replayed hipGraph: batched render + NCC + pose adjoint + Adam update + "
                                   "views recomputed from the poses on the device")
    plan.close()
    vol.close()
    ctx.close()
    return res, images


# ---------------------------------------------------------------------------
# CPU baseline and accuracy (the oracle is what the GPU results are tested
# against, and the CPU port)
def _median_time(fn, runs: int = 5) -> float:
    fn()
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return statistics.median(ts)


def _oracle_native() -> str:
    """Build the oracle for this host (-march=native) and load it; falls back
    to the portable x86-64-v3 build if the native one does not build."""
    from oracle import oracle as orc
    try:
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "native"], check=True, capture_output=True,
                       timeout=300)
        orc.use_library(ROOT / "oracle" / "_native" / "librvk_oracle.so")
        orc.lib()
        return "-O3 -march=native"
    except Exception as e:  # noqa: BLE001
        log(f"native oracle build failed ({e}); timing the portable build")
        orc.use_library(ROOT / "oracle" / "librvk_oracle.so")
        return "-O3 -march=x86-64-v3"


def _rel(a, b):
    import numpy as np
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30)))


def _lsb(g, o):
    import numpy as np
    d = np.abs(g.astype(np.int32) - o.astype(np.int32)).max(axis=-1)
    return round(float((d <= 1).mean()), 6), int(d.max())


def cpu_baseline(scenes: dict, gpu: dict) -> dict:
    """Times the FP32 oracle on every host core of this process, per config,
    and reports each config's GPU error against the FP64 oracle (replaying
    the FP32 branch decisions, and unconstrained)."""
    import numpy as np

    from oracle import oracle as orc
    from robartsvtk_amd import abi, configs

    import_helpers = str(ROOT / "tests")
    if import_helpers not in sys.path:
        sys.path.insert(0, import_helpers)
    import helpers as H

    build = _oracle_native()
    th = cpu_threads()
    per, acc = {}, {}

    # c2 (headline)
    s = scenes["c2"]
    vd = configs.view_dict(s.views[0])
    n2 = gpu["c2"]["n_samples"]
    t = _median_time(lambda: orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f32", nthreads=th))
    per["c2"] = {"G_samples_per_s": round(n2 / t / 1e9, 4), "s": round(t, 4), "sample": "full frame"}
    log(f"cpu c2 {t:.3f} s")
    o64 = orc.drr(s.volume, s.spacing, vd, dict(s.march, f32_control=1), s.drr, "f64", nthreads=th)
    o64u = orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f64", nthreads=th)
    acc["c2"] = {"max_rel_err_f64_f32_decisions": _rel(gpu["c2"]["image"], o64),
                 "max_rel_err_f64_unconstrained": _rel(gpu["c2"]["image"], o64u), "tolerance": 1e-4,
                 "n_samples_oracle_f64": orc.count_samples(s.volume, s.spacing, vd, s.march, nthreads=th),
                 "n_samples_gpu": n2,
                 "evaluated_samples_skip_rule": orc.count_drr_evaluated(s.volume, s.spacing, vd, s.march, s.drr,
                                                                        nthreads=th)}
    # c1
    s = scenes["c1"]
    vd = configs.view_dict(s.views[0])
    t = _median_time(lambda: orc.cast1d(s.volume, s.spacing, s.tf1d, vd, s.march, "f32", nthreads=th))
    per["c1"] = {"G_samples_per_s": round(gpu["c1"]["n_samples"] / t / 1e9, 4), "s": round(t, 4),
                 "sample": "full frame"}
    o64, _, _ = orc.cast1d(s.volume, s.spacing, s.tf1d, vd, dict(s.march, f32_control=1), "f64", nthreads=th)
    o64u, _, _ = orc.cast1d(s.volume, s.spacing, s.tf1d, vd, s.march, "f64", nthreads=th)
    f, m = _lsb(gpu["c1"]["image"], o64)
    fu, mu = _lsb(gpu["c1"]["image"], o64u)
    acc["c1"] = {"within_1_lsb_f64_f32_decisions": f, "max_lsb": m, "within_1_lsb_f64_unconstrained": fu,
                 "max_lsb_unconstrained": mu, "tolerance": ">= 0.999 within 1 LSB"}
    log(f"cpu c1 {t:.3f} s")
    # c3
    s = scenes["c3"]
    vd = configs.view_dict(s.views[0])
    gmax = s.tf2d["gradmag_max"]
    tf = dict(s.tf2d, gradmag=H.gradmag_u8(s.volume, s.spacing, gmax),
              gradmag_scale=float(np.float32(gmax) / np.float32(255.0)))
    t = _median_time(lambda: orc.cast2d(s.volume, s.spacing, tf, vd, s.march, "f32", nthreads=th), runs=5)
    per["c3"] = {"G_samples_per_s": round(gpu["c3"]["n_samples"] / t / 1e9, 4), "s": round(t, 4),
                 "sample": "full frame"}
    log(f"cpu c3 {t:.3f} s")
    o64, _, _ = orc.cast2d(s.volume, s.spacing, tf, vd, dict(s.march, f32_control=1), "f64", nthreads=th)
    o64u, _, _ = orc.cast2d(s.volume, s.spacing, tf, vd, s.march, "f64", nthreads=th)
    f, m = _lsb(gpu["c3"]["image"], o64)
    fu, mu = _lsb(gpu["c3"]["image"], o64u)
    acc["c3"] = {"within_1_lsb_f64_f32_decisions": f, "max_lsb": m, "within_1_lsb_f64_unconstrained": fu,
                 "max_lsb_unconstrained": mu, "tolerance": ">= 0.999 within 1 LSB"}
    # c4: one pose of each detector size (4 of 64)
    s = scenes["c4"]
    sizes = s.extra["sizes"]
    pick = sorted({n: v for v, n in reversed(list(enumerate(sizes)))}.values())
    off = np.concatenate([[0], np.cumsum([v.width * v.height for v in s.views])])
    vds = [configs.view_dict(s.views[v]) for v in pick]
    t = _median_time(lambda: [orc.drr(s.volume, s.spacing, d, s.march, s.drr, "f32", nthreads=th) for d in vds],
                     runs=5)
    ctx = abi.Context(0)
    vol = abi.Volume.from_array(ctx, s.volume, s.spacing)
    mp = configs.march_params(s.march)
    ns = sum(abi.count_ray_steps(ctx, vol, s.views[v], mp) for v in pick)
    vol.close()
    ctx.close()
    per["c4"] = {"G_samples_per_s": round(ns / t / 1e9, 4), "poses_per_s": round(len(pick) / t, 3),
                 "fwd_fd_poses_per_s": round(len(pick) / t / 13.0, 4), "s": round(t, 4),
                 "sample": f"poses {pick} (one per detector size), forward; fwd+FD = 13 renders per pose"}
    e64, e64u = 0.0, 0.0
    for v, d in zip(pick, vds):
        g = gpu["c4"]["images"][off[v]:off[v + 1]].reshape(sizes[v], sizes[v])
        e64 = max(e64, _rel(g, orc.drr(s.volume, s.spacing, d, dict(s.march, f32_control=1), s.drr, "f64",
                                       nthreads=th)))
        e64u = max(e64u, _rel(g, orc.drr(s.volume, s.spacing, d, s.march, s.drr, "f64", nthreads=th)))
    acc["c4"] = {"max_rel_err_f64_f32_decisions": e64, "max_rel_err_f64_unconstrained": e64u, "views": pick,
                 "tolerance": 1e-4}
    log(f"cpu c4 {t:.3f} s")
    # c5: the 16 views (forward), and one view's pose gradient against FP64 FD
    s = scenes["c5"]
    img = s.extra["img"]
    vds = [configs.view_dict(v) for v in s.views]
    t = _median_time(lambda: [orc.drr(s.volume, s.spacing, d, s.march, s.drr, "f32", nthreads=th) for d in vds],
                     runs=5)
    per["c5"] = {"G_samples_per_s": round(gpu["c5"]["n_samples"] / t / 1e9, 4), "views_per_s": round(16 / t, 3),
                 "fwd_fd_poses_per_s": round(16 / t / 13.0, 4), "s": round(t, 4),
                 "sample": "all 16 views, forward; fwd+FD = 13 renders per pose"}
    e64 = max(_rel(gpu["c5"]["images"][v * img * img:(v + 1) * img * img].reshape(img, img),
                   orc.drr(s.volume, s.spacing, vds[v], dict(s.march, f32_control=1), s.drr, "f64", nthreads=th))
              for v in (0, 9))
    acc["c5"] = {"max_rel_err_f64_f32_decisions": e64, "views": [0, 9], "tolerance": 1e-4}
    acc["c5"].update(_c5_gradient_accuracy(s, gpu["c5"]["grad_view0"], gpu["c5"]["target_view0"], th))
    log(f"cpu c5 {t:.3f} s")
    head = per["c2"]
    return {"value": head["G_samples_per_s"], "unit": "G ray-samples/s", "cores": th, "kind": "port",
            "sample": f"full c2 frame (1024x1024 rays, {gpu['c2']['n_samples']} samples), FP32 oracle built "
                      f"{build}, {th} threads, median of 5 after 1 warm-up, {head['s']:.3f} s each",
            "host": dict(host_info(), build=build), "per_config": per, "accuracy_vs_fp64": acc}


def _c5_gradient_accuracy(s, grad, target, th) -> dict:
    """View 0 of c5: the registration step's pose gradient against FP64
    central differences of 1 - NCC at h = 1e-3 and 1e-6."""
    import numpy as np

    from oracle import oracle as orc
    from robartsvtk_amd import abi

    c, img = s.extra["center"], s.extra["img"]
    cam = s.extra["cameras"][0]

    def baked64(pose):
        u, _ = abi.pose_matrix(pose, c)
        m, _ = abi.compute_matrices(cam, 1.0, spacing=s.spacing, user_matrix=u)
        return np.array([[m[r, 0] * -2.0, m[r, 1] * 2.0, m[r, 2], m[r, 3] + m[r, 0] - m[r, 1]] for r in range(4)])

    def loss(pose):
        vd = {"m": np.zeros(16, np.float32), "W": img, "H": img, "m64": baked64(pose)}
        a = orc.drr(s.volume, s.spacing, vd, s.march, s.drr, "f64", nthreads=th).ravel()
        b = target.astype(np.float64)
        a, b = a - a.mean(), b - b.mean()
        return 1.0 - float(a @ b / np.sqrt((a @ a) * (b @ b)))

    out = {}
    for h in (1e-3, 1e-6):
        fd = np.zeros(6)
        for i in range(6):
            p1 = np.array(s.extra["poses"][0], float)
            p2 = p1.copy()
            p1[i] += h
            p2[i] -= h
            fd[i] = (loss(p1) - loss(p2)) / (2 * h)
        er = float(np.linalg.norm(grad[:3] - fd[:3]) / np.linalg.norm(fd[:3]))
        et = float(np.linalg.norm(grad[3:] - fd[3:]) / np.linalg.norm(fd[3:]))
        out[f"pose_grad_rel_err_fd_h{h:g}"] = {"rotation": er, "translation": et}
    out["pose_grad_tolerance"] = 1e-3
    return out


# ---------------------------------------------------------------------------
def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    # 50 untimed steps (~24 ms of c2) bring the clocks to their steady state;
    # after 5 the timed steps still run ~2 % slower
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the c4/c5 batching measurements")
    ap.add_argument("--no-casters", action="store_true", help="skip the c1/c3 compositing-caster measurements")
    ap.add_argument("--no-probe", action="store_true", help="skip the streaming-read bandwidth probe")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build runs every kernel on a single MI355X and has no "
              "multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import torch

    from robartsvtk_amd import configs

    if args.copy_mib:
        n = args.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log(f"copied {args.copy_mib} MiB twice")
    log("building c2 scene")
    scenes = {"c2": configs.scene_c2(scale=1.0)}
    log("timing DRR caster (c2, skip rule on)")
    r = bench_drr(scenes["c2"], args.steps, args.warmup)
    gpu = {"c2": r}
    sec = r["ms_per_step"] / 1e3
    value = r["n_samples"] / sec / 1e9
    log("timing DRR caster (c2, plain march)")
    rp = bench_drr(scenes["c2"], max(5, args.steps // 2), min(args.warmup, 2),
                   march=dict(scenes["c2"].march, skip_rule=0))
    kern_s = r["kernel_ms"] / 1e3
    alg_bytes = r["n_samples"] * 8 * 2 + r["pixels"] * 4
    cnt = kernel_counters(DRR_KERNEL)
    roof = {"bound": "valu", "unit": "G VALU wave-instructions/s", "peak": VALU_PEAK_GINST,
            "kernel": DRR_KERNEL, "kernel_ms": round(r["kernel_ms"], 4)}
    if cnt:
        ach = cnt["SQ_INSTS_VALU"] / kern_s / 1e9
        hbm = cnt["hbm_bytes"] / kern_s / 1e9
        roof.update({"achieved": round(ach, 1), "frac": round(ach / VALU_PEAK_GINST, 4), "traffic": cnt["hbm_bytes"],
                     "valu_insts_per_launch": cnt["SQ_INSTS_VALU"], "valu_busy_counters": cnt.get("valu_busy"),
                     "hbm": {"achieved_GBs": round(hbm, 1), "peak": HBM_PEAK_GBS, "frac": round(hbm / HBM_PEAK_GBS, 4)}})
    else:
        roof.update({"achieved": None, "frac": None, "traffic": None})
    roof["gather_equivalent"] = {"bytes_per_launch": alg_bytes, "GB_per_s": round(alg_bytes / kern_s / 1e9, 1),
                                 "x_hbm_peak": round(alg_bytes / kern_s / 1e9 / HBM_PEAK_GBS, 4),
                                 "note": "8 voxels x 2 B per sample + 4 B per pixel; served mostly from L2"}
    line = {
        "metric": "G ray-samples/s + batched DRR fwd+bwd poses/s vs plain per-pose, 1 MI355X",
        "value": round(value, 3),
        "unit": "G ray-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "c2: DRR of a 256^3 int16 CT phantom (seeded), cone beam SID 1000 / SOD 600 mm, "
                               "1024x1024 float32 exp(-L), step 0.5 voxel, skip rule on, precise march",
                   "n_samples": r["n_samples"], "voxel_bytes": 2},
        "roofline": roof,
        "c2_plain_march": {"G_samples_per_s": round(rp["n_samples"] / (rp["ms_per_step"] / 1e3) / 1e9, 3),
                           "ms_per_step": round(rp["ms_per_step"], 4),
                           "note": "skip rule off: every one of the n_samples samples is evaluated"},
    }
    if not args.no_probe:
        log("streaming-read probe")
        line["hbm_stream_read"] = stream_read_probe()
    if not args.no_casters:
        log("timing c1: 1D TF caster, 128^3 u8, 512x512")
        scenes["c1"] = configs.scene_c1(scale=1.0)
        c1, gpu_img = bench_compositing(scenes["c1"], args.steps, args.warmup)
        gpu["c1"] = dict(c1, image=gpu_img)
        log("timing c3: 2D TF caster, 512^3 u16 + u8 gradient magnitude, 1920x1080")
        scenes["c3"] = configs.scene_c3(scale=1.0)
        c3, gpu_img = bench_compositing(scenes["c3"], max(5, args.steps // 5), min(args.warmup, 2))
        gpu["c3"] = dict(c3, image=gpu_img)
        line["casters"] = {"c1": c1, "c3": c3}
        for k, name in (("c1", "cast1d_kernel<Cells<unsigned char,0,false>,true,false,true>"),
                        ("c3", "cast2d_kernel<Cells<unsigned short,2,true>,true,false,true,true>")):
            cc = kernel_counters(name)
            if cc:
                ks = line["casters"][k]["ms_per_render"] / 1e3
                line["casters"][k]["counters"] = {
                    "kernel": name, "valu_busy": cc.get("valu_busy"), "valu_lane_use": cc.get("valu_lane_use"),
                    "hbm_bytes": cc["hbm_bytes"], "hbm_GBs": round(cc["hbm_bytes"] / ks / 1e9, 1),
                    "hbm_frac": round(cc["hbm_bytes"] / ks / 1e9 / HBM_PEAK_GBS, 4)}
    if not args.no_batched:
        log("timing c4: 64 poses, uneven detectors, batched vs per-pose")
        scenes["c4"] = configs.scene_c4(scale=1.0, with_derivatives=True)
        c4, imgs = bench_batched(scenes["c4"], volume_gradient=True)
        gpu["c4"] = {"images": imgs}
        log("timing c5: 16-view registration step")
        scenes["c5"] = configs.scene_c5(scale=1.0)
        c5, imgs = bench_batched(scenes["c5"], volume_gradient=True, loop=True)
        gpu["c5"] = {"images": imgs, "n_samples": c5["n_samples"]}
        line["batched"] = {"c4": c4, "c5": c5,
                           "note": "fwd_bwd = one replayed hipGraph of batched render + NCC + pose adjoint; plain = "
                                   "13 single-view renders per pose (1 + 12 central differences); both marches "
                                   "plain (no skip rule)"}
    if not args.no_cpu_baseline and not args.no_casters and not args.no_batched:
        # the c5 accuracy leg needs one view's gradient against a fixed target
        gpu["c5"].update(_c5_view0_gradient(scenes["c5"]))
        log("timing CPU oracle baseline")
        line["cpu_baseline"] = cpu_baseline(scenes, gpu)
    elif not args.no_cpu_baseline:
        log("timing CPU oracle baseline (c2 only)")
        line["cpu_baseline"] = cpu_baseline_c2(scenes["c2"], gpu["c2"])
    print(json.dumps(line), flush=True)
    return 0


def _c5_view0_gradient(s) -> dict:
    """The registration step's pose gradient for c5 view 0 against a target
    DRR at the ground-truth pose plus noise (GPU render of the target)."""
    import numpy as np
    import torch

    from robartsvtk_amd import abi, configs

    ctx = abi.Context(0)
    vol = abi.Volume.from_array(ctx, s.volume, s.spacing)
    c, img = s.extra["center"], s.extra["img"]
    gt_view = abi.posed_view(s.extra["cameras"][0], s.extra["gt_poses"][0], c, img, img, spacing=s.spacing)
    t = torch.empty((img, img), dtype=torch.float32, device="cuda")
    abi.render_drr(ctx, vol, gt_view, configs.march_params(s.march), configs.drr_params(s.drr), t)
    g = torch.Generator(device="cuda").manual_seed(5)
    tgt = (t.ravel() + 0.01 * torch.randn(img * img, device="cuda", generator=g)).contiguous()
    plan = abi.DrrPlan(ctx, vol, s.views[:1], configs.march_params(s.march), configs.drr_params(s.drr),
                       dview=s.extra["dview"][:1])
    out = torch.empty(plan.total, dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    grad = torch.empty(6, dtype=torch.float32, device="cuda")
    plan.register(out, tgt, loss, grad)
    torch.cuda.synchronize()
    res = {"grad_view0": grad.cpu().numpy().astype(np.float64), "target_view0": tgt.cpu().numpy()}
    plan.close()
    vol.close()
    ctx.close()
    return res


def cpu_baseline_c2(scene, r) -> dict:
    from oracle import oracle as orc
    from robartsvtk_amd import configs

    build = _oracle_native()
    th = cpu_threads()
    vd = configs.view_dict(scene.views[0])
    t = _median_time(lambda: orc.drr(scene.volume, scene.spacing, vd, scene.march, scene.drr, "f32", nthreads=th))
    return {"value": round(r["n_samples"] / t / 1e9, 4), "unit": "G ray-samples/s", "cores": th, "kind": "port",
            "sample": f"full c2 frame, FP32 oracle built {build}, {th} threads, median of 5, {t:.3f} s each",
            "host": dict(host_info(), build=build)}


if __name__ == "__main__":
    sys.exit(main())
