"""Benchmark: G ray-samples/s of the gfx950 DRR caster on one MI355X.

Workload (N=1): 256^3 int16 CT phantom, cone beam SID 1000 / SOD 600 mm,
1024x1024 float32 exp(-L) detector, step 0.5 voxel (config c2).  One step =
one full-detector render with the volume already resident in HBM.

Prints ONE JSON line.  `value` = N_samples / (seconds per render) / 1e9, where
N_samples is the plain-march ray-sample count (sum over pixels of
floor(num_steps - ray offset)).  `roofline` prices the caster kernel's
gather-equivalent bytes (8 voxels x 2 B per sample + 4 B per pixel) against
the 8 TB/s HBM peak, and `roofline.traffic` is the kernel's real HBM bytes per
launch from the committed rocprofv3 counter passes; `cpu_baseline` times the
CPU oracle (FP32, all granted host threads) on the same frame.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed counter passes
    (profiles/drr_kernel_c2_traffic.json: FETCH_SIZE and WRITE_SIZE scaled by
    a 256 MiB copy measured in the same runs), or None."""
    f = ROOT / "profiles" / "drr_kernel_c2_traffic.json"
    try:
        rec = json.loads(f.read_text())
    except (OSError, ValueError):
        return None
    return int(rec["bytes_per_launch"]) if rec.get("kernel") == kernel else None


def cpu_threads() -> int:
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    return max(1, min(16, n))


def bench_drr(scene, steps: int, warmup: int):
    import torch

    from robartsvtk_amd import abi, configs

    ctx = abi.Context(0)
    vol = abi.Volume.from_array(ctx, scene.volume, scene.spacing)
    view = scene.views[0]
    mp = configs.march_params(scene.march)
    dp = configs.drr_params(scene.drr)
    out = torch.empty((view.height, view.width), dtype=torch.float32, device="cuda")
    n_samples = abi.count_ray_steps(ctx, vol, view, mp)
    for _ in range(warmup):
        abi.render_drr(ctx, vol, view, mp, dp, out)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
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
    return {"n_samples": n_samples, "ms_per_step": total_ms / steps, "kernel_ms": kern_ms, "pixels": view.width * view.height,
            "image_mean": float(img.mean())}


def bench_compositing(scene, steps: int, warmup: int, gradmag: bool = True):
    """1D (c1) or 2D (c3) compositing caster.  N_samples is the plain march
    (skip rule off) up to opacity termination: one render with the skip rule
    off writes the per-pixel sample count into the steps-taken side output,
    and its sum is the count.  The timed renders use the configured march
    (skip rule on).  gradmag=False renders the 2D caster with gradients from
    six gathers instead of the u8 gradient-magnitude volume."""
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
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in ev:
        a.record(stream)
        render(mp, out)
        b.record(stream)
    torch.cuda.synchronize()
    kern_ms = statistics.fmean(a.elapsed_time(b) for a, b in ev)
    coverage = float((out[..., 3] > 0).float().mean().item())
    tfh.close()
    vol.close()
    ctx.close()
    vbytes = scene.volume.dtype.itemsize
    per_sample = 8 * vbytes + (8 if two_d else 0)  # + the u8 gradient-magnitude footprint
    alg_bytes = n_samples * per_sample + W * H * 4
    achieved = alg_bytes / (kern_ms / 1e3) / 1e9
    return {"n_samples": n_samples, "pixels": W * H, "ms_per_render": round(kern_ms, 4),
            "G_samples_per_s": round(n_samples / (kern_ms / 1e3) / 1e9, 3),
            "bytes_per_sample": per_sample, "covered_pixel_fraction": round(coverage, 4),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4)}}


def _time(fn, reps: int) -> float:
    """Mean milliseconds of fn() over reps runs (after one warm-up), HIP events
    on torch's current stream (the stream the library's context uses)."""
    import torch
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s = torch.cuda.current_stream()
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def bench_batched(scene, reps: int = 5):
    """Batched forward and forward+backward (registration step) against the
    plain per-pose loop: one single-view launch per pose for the forward, and
    13 single-view renders per pose (1 + 12 central differences of the 6-DoF
    pose) for the gradient.  The plain loop's loss evaluations are left out,
    which favours the plain loop."""
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
    torch.cuda.synchronize()
    assert torch.isfinite(loss).all() and torch.isfinite(grad).all()
    plan.close()
    vol.close()
    ctx.close()
    return {"poses": n, "pixels": int(sum(v.width * v.height for v in views)), "n_samples": n_samples,
            "plain_fwd_ms": round(t_plain, 3), "batched_fwd_ms": round(t_batch, 3),
            "fwd_speedup": round(t_plain / t_batch, 3),
            "batched_fwd_G_samples_per_s": round(n_samples / (t_batch / 1e3) / 1e9, 2),
            "schedule": {"kind": "persistent" if sched[0] == abi.SCHED_PERSISTENT else "grid",
                         "waves_per_simd": sched[1], "autotune_ms": round(tuned_ms, 3)},
            "plain_fwd_fd_ms": round(t_plain_fd, 3), "batched_fwd_bwd_ms": round(t_reg, 3),
            "plain_fd_poses_per_s": round(n / (t_plain_fd / 1e3), 2),
            "batched_fwd_bwd_poses_per_s": round(n / (t_reg / 1e3), 2),
            "fwd_bwd_speedup": round(t_plain_fd / t_reg, 3)}


def cpu_baseline(scene, n_samples: int, runs: int = 3):
    from oracle import oracle as orc
    from robartsvtk_amd import configs

    th = cpu_threads()
    vd = configs.view_dict(scene.views[0])
    orc.drr(scene.volume, scene.spacing, vd, scene.march, scene.drr, "f32", nthreads=th)  # warm-up
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        orc.drr(scene.volume, scene.spacing, vd, scene.march, scene.drr, "f32", nthreads=th)
        ts.append(time.perf_counter() - t)
    med = statistics.median(ts)
    return {"value": round(n_samples / med / 1e9, 4), "unit": "G ray-samples/s", "cores": th, "kind": "port",
            "sample": f"full c2 frame (1024x1024 rays, {n_samples} samples), FP32 oracle, median of {runs}, "
                      f"{med:.3f} s each"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the c4/c5 batching measurements")
    ap.add_argument("--no-casters", action="store_true", help="skip the c1/c3 compositing-caster measurements")
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
        log(f"copied {args.copy_mib} MiB twice")
    log("building c2 scene")
    scene = configs.scene_c2(scale=1.0)
    log("timing DRR caster")
    r = bench_drr(scene, args.steps, args.warmup)
    sec = r["ms_per_step"] / 1e3
    value = r["n_samples"] / sec / 1e9
    alg_bytes = r["n_samples"] * 8 * 2 + r["pixels"] * 4
    achieved = alg_bytes / (r["kernel_ms"] / 1e3) / 1e9
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
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("drr_kernel<Cells<HalfPair,0,true>,true,true,false,false>"),
                     "kernel": "drr_kernel<Cells<HalfPair,0,true>,true,true,false,false>", "kernel_ms": round(r["kernel_ms"], 4),
                     "alg_bytes_per_launch": alg_bytes},
    }
    if not args.no_casters:
        log("timing c1: 1D TF caster, 128^3 u8, 512x512")
        line["casters"] = {"c1": bench_compositing(configs.scene_c1(scale=1.0), args.steps, args.warmup)}
        log("timing c3: 2D TF caster, 512^3 u16 + u8 gradient magnitude, 1920x1080")
        line["casters"]["c3"] = bench_compositing(configs.scene_c3(scale=1.0), max(5, args.steps // 5),
                                                  min(args.warmup, 2))
    if not args.no_batched:
        log("timing c4: 64 poses, uneven detectors, batched vs per-pose")
        line["batched"] = {"c4": bench_batched(configs.scene_c4(scale=1.0, with_derivatives=True))}
        log("timing c5: 16-view registration step")
        line["batched"]["c5"] = bench_batched(configs.scene_c5(scale=1.0))
        line["batched"]["note"] = ("fwd_bwd = one replayed hipGraph of batched render + NCC + pose adjoint; plain = "
                                   "13 single-view renders per pose (1 + 12 central differences)")
    if not args.no_cpu_baseline:
        log("timing CPU oracle baseline")
        line["cpu_baseline"] = cpu_baseline(scene, r["n_samples"])
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
