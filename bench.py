#!/usr/bin/env python
"""Benchmark of the Gaussian-splatting training step on one MI355X.

Workload (the configuration the headline metric is quoted on): 1,048,576
Gaussians with SH degree 3 (synthetic, seed 5), 8 views per step with uneven
resolutions {640x480, 960x540, 1280x720 x3, 1600x900, 1920x1080 x2} on a ring
of radius 5 around the scene, targets rendered from a second synthetic scene
(seed 6).  One step = forward of the 8 views, L1 loss (weight 1/8 per view),
backward summed over the views, one fused Adam update (constant learning
rates: means 1.6e-4, scales 5e-3, quats 1e-3, opacity 5e-2, SH DC 2.5e-3, SH
rest 1.25e-4), and the densification statistics accumulated per view as
the reference's train_step does.

`value` is batched steps/s: all views in one launch per stage, the step
captured as a HIP graph, tile width and batch depth from the autotuner.  The
plain per-view path (views one after another, no graph, same kernels and
maths) is timed in the same run and reported beside it.

Also reported: per-stage kernel times from HIP events on the library stream
(an instrumented eager pass over the same steps), the roofline of the
dominant kernel, gradient agreement between the batched and per-view paths,
and the CPU baseline (the oracle, the C++ restatement of the reference
pipeline, timed on the host for one full step and compared on gradients).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# the reference's train_step accumulates the densification statistics every
# step (trainer.rs:366-373); the timed steps do the same
from rustscan_amd._lib import RGS_DENSIFY_ACCUMULATE  # noqa: E402

METRIC = "train steps/s @1M Gaussians 720p: batched vs per-view, grad rel err ≤1e-3"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X published vector FP32 (counts packed FP32: two lanes per FMA)
FP32_NONPACKED_TFLOPS = 78.6  # the same without packed FP32, which this build does not emit


# kernel of each timed stage in the batched step (the dominant one is looked up
# in the HBM-traffic profile)
STAGE_KERNELS = {
    "project_fwd": "k_project_fwd<3>", "depth_sort": "k_radix_scatter", "count_scan": "k_scan_down",
    "emit": "k_emit", "tile_sort": "k_radix_scatter", "tile_ranges": "k_tile_ranges",
    "raster_fwd": "k_raster_fwd<{tw}, {bd}, true, false>", "raster_bwd": "k_raster_bwd<{tw}, {bd}, false, true, false>",
    "project_bwd": "k_project_bwd<3>", "adam": "k_adam",
}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_8view_1M.md")


def measured_traffic(kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (calibrated against a plain copy), or None."""
    import re

    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    for name, val in re.findall(r"traffic\[(.+?)\] = (\d+)", text):
        if name == kernel:
            return int(val)
    return None


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / d) if d > 0 else float(np.linalg.norm(a))


def run_eval_render(ctx, state, cam, target, reps: int = 10) -> dict:
    """GaussianRenderer::render_splats / TiledRenderer semantics on the device
    for one view of the scene, and one evaluation frame (render + metrics):
    milliseconds per synchronous call over `reps` calls after one warm-up."""
    import torch

    from rustscan_amd.evaluation import GaussianRenderer, TiledRenderer, evaluate_frame
    from rustscan_amd.rasterizer import DeviceSplats

    ds = DeviceSplats(ctx, state)
    gr = GaussianRenderer(cam.width, cam.height).with_raster_cov_blur(0.3)
    gr.render_splats(ctx, ds, cam)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        gr.render_splats(ctx, ds, cam)
    t_render = (time.perf_counter() - t0) / reps
    evaluate_frame(ctx, ds, cam, target, 0.3)
    t0 = time.perf_counter()
    for _ in range(reps):
        metrics, _ = evaluate_frame(ctx, ds, cam, target, 0.3)
    t_eval = (time.perf_counter() - t0) / reps
    color, depth = TiledRenderer(cam.width, cam.height).with_raster_cov_blur(0.3).render_camera_splats(ctx, ds, cam)
    images = (color.cpu().numpy(), depth.cpu().numpy())
    ds.close()
    px = cam.width * cam.height
    log(f"eval render {cam.width}x{cam.height}: {1e3 * t_render:.2f} ms/frame; evaluate frame "
        f"{1e3 * t_eval:.2f} ms (psnr {metrics['psnr_db']:.3f} dB)")
    return {"view": f"{cam.width}x{cam.height}", "render_ms": round(1e3 * t_render, 3),
            "render_mpix_per_s": round(px / t_render / 1e6, 1), "evaluate_frame_ms": round(1e3 * t_eval, 3),
            "psnr_db": metrics["psnr_db"], "_images": images}


def run_topology_and_files(ctx, opt, state, cams, targets, steps: int = 10) -> dict:
    """One device topology step (densify / prune / reset) on the 1,048,576-
    Gaussian C4 scene after `steps` batched steps that gather the
    densification statistics, then a save and load of the resulting store as
    the reference's ASCII PLY; wall clock per synchronous call.  RustGS's
    defaults except: refine_every 1 (the schedule applies at once), growth
    and split gradient thresholds 1e-6 (ten steps of a synthetic scene
    gather less than the defaults' 1.4e-4 of a trained one) and threshold
    pruning at opacity 0.05, so that the step clones, splits and prunes."""
    import os
    import shutil
    import tempfile

    from rustscan_amd import scene_io
    from rustscan_amd.rasterizer import DeviceSplats, TopologyConfig, Trainer

    sp = DeviceSplats(ctx, state)
    tr = Trainer(ctx, opt, cov_blur=0.3, loss_weight=1.0 / len(cams), densify=RGS_DENSIFY_ACCUMULATE)
    for _ in range(steps):
        tr.train_step(sp, cams, targets, batched=True, graph=True)
    ctx.sync()
    t0 = time.perf_counter()
    cfg = TopologyConfig(refine_every=1, growth_grad_threshold=1e-6, split_grad_threshold=1e-6, prune_mode=0,
                         prune_opacity_threshold=0.05)
    res = sp.topology_step(cfg, 100, len(cams))
    ctx.sync()
    t_topo = time.perf_counter() - t0
    topo = {"ms": round(1e3 * t_topo, 3), "before": int(res["before"]), "after": int(res["after"]),
            "added": int(res["added"]), "pruned": int(res["pruned"]),
            "clone_candidates": int(res["clone_candidates"]), "split_candidates": int(res["split_candidates"])}
    log(f"topology step: {topo}")
    d = tempfile.mkdtemp(prefix="rgs_bench_")
    try:
        path = os.path.join(d, "scene.ply")
        t0 = time.perf_counter()
        scene_io.save_device_splats(ctx, sp, path)
        t_save = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        loaded, _ = scene_io.load_device_splats(ctx, path)
        ctx.sync()
        t_load = time.perf_counter() - t0
        files = {"format": "PLY, ASCII (the reference's writer)", "gaussians": int(sp.n), "ply_bytes": int(size), "save_ms": round(1e3 * t_save, 1),
                 "load_ms": round(1e3 * t_load, 1), "save_mb_per_s": round(size / t_save / 1e6, 1),
                 "load_mb_per_s": round(size / t_load / 1e6, 1)}
        loaded.close()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    log(f"scene file (ASCII PLY): {files}")
    sp.close()
    return {"topology_step": topo, "scene_file": files}


def run_config3(ctx, opt, init, gt, steps: int = 200) -> dict:
    """The reference's training semantics: one 1280x720 view and one Adam
    update per step, cycling over 32 ring cameras, targets rendered from the
    ground-truth scene; steps/s and the L1 loss at the start and the end."""
    import torch

    from rustscan_amd import scene
    from rustscan_amd.rasterizer import DeviceSplats, Trainer, render_forward

    cams = scene.config3_cameras()
    g = DeviceSplats(ctx, gt)
    targets = []
    for c in cams:
        targets.append(render_forward(ctx, g, [c])[0][..., :3].contiguous())
    g.close()
    splats = DeviceSplats(ctx, init)
    trainer = Trainer(ctx, opt, cov_blur=0.3, loss_weight=1.0, densify=RGS_DENSIFY_ACCUMULATE)
    first = [trainer.train_step(splats, [cams[i]], [targets[i]], batched=True, graph=True, read_loss=True)[0]
             for i in range(8)]
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(steps):
        i = it % len(cams)
        trainer.train_step(splats, [cams[i]], [targets[i]], batched=True, graph=True)
    ctx.sync()
    dt = (time.perf_counter() - t0) / steps
    last = [trainer.train_step(splats, [cams[i]], [targets[i]], batched=True, graph=True, read_loss=True)[0]
            for i in range(8)]
    splats.close()
    out = {
        "workload": "1,048,576 Gaussians SH3, 32 ring views 1280x720, one view + one Adam update per step",
        "steps": steps,
        "steps_per_s": round(1.0 / dt, 2),
        "ms_per_step": round(1e3 * dt, 3),
        "l1_first_8_views": round(float(np.mean(first)), 5),
        "l1_after_training": round(float(np.mean(last)), 5),
        "skipped_steps": ctx.skipped_steps(),
    }
    log(f"config3: {out}")
    return out


def run_config5(args, ctx, opt) -> dict:
    """Large batched training step: 4,194,304 Gaussians (1% with scale x20),
    16 views at 1920x1080, uniform random targets; steps/s, intersections and
    device memory in use."""
    import torch

    from rustscan_amd import scene
    from rustscan_amd.rasterizer import DeviceSplats, Trainer

    t0 = time.perf_counter()
    sc = scene.config5_scene()
    cams = scene.config5_cameras()
    gen = torch.Generator(device="cuda").manual_seed(10)
    targets = [torch.rand((c.height, c.width, 3), device="cuda", generator=gen) for c in cams]
    splats = DeviceSplats(ctx, sc)
    trainer = Trainer(ctx, opt, cov_blur=0.3, loss_weight=1.0 / len(cams), densify=RGS_DENSIFY_ACCUMULATE)
    log(f"config5 setup {time.perf_counter() - t0:.1f}s")
    trainer.train_step(splats, cams, targets, batched=True, graph=True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.config5_steps):
        trainer.train_step(splats, cams, targets, batched=True, graph=True)
    ctx.sync()
    dt = (time.perf_counter() - t0) / args.config5_steps
    free, total = torch.cuda.mem_get_info()
    stages = {}
    if not args.no_instrument:
        # where the step's time goes: HIP events around each stage of eager
        # steps (after the timed graph steps; the parameters keep training)
        ctx.set_timing(True)
        ctx.timing(reset=True)
        for _ in range(args.config5_steps):
            trainer.train_step(splats, cams, targets, batched=True, graph=False)
        ctx.sync()
        stages = {k: round(ms / args.config5_steps, 3) for k, (ms, calls) in ctx.timing(reset=True).items() if calls}
        ctx.set_timing(False)
    counts = ctx.counts(len(cams))  # reads the tile ranges and keys back (after the steps)
    out = {
        "workload": "4,194,304 Gaussians SH3 (1% log-scale + ln 20), 16 views 1920x1080, batched, HIP graph",
        "steps_per_s": round(1.0 / dt, 4),
        "ms_per_step": round(1e3 * dt, 2),
        "intersections": int(sum(c[1] for c in counts)),
        "visible_instances": int(sum(c[0] for c in counts)),
        "device_memory_in_use_gb": round((total - free) / 1e9, 2),
        "train_mpix_per_s": round(sum(c.width * c.height for c in cams) / dt / 1e6, 1),
        "skipped_steps": ctx.skipped_steps(),
        "stage_ms_per_step": stages or None,
    }
    log(f"config5: {out}")
    splats.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1048576, help="Gaussians (the workload uses 1,048,576)")
    ap.add_argument("--per-view-steps", type=int, default=None)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-instrument", action="store_true")
    ap.add_argument("--no-config5", action="store_true",
                    help="skip the large configuration (4,194,304 Gaussians with a heavy scale tail, "
                         "16 views at 1920x1080 per batched step)")
    ap.add_argument("--config5-steps", type=int, default=3)
    ap.add_argument("--no-config3", action="store_true",
                    help="skip the single-view training loop (32 views at 1280x720, one view and one Adam "
                         "update per step, 200 steps)")
    ap.add_argument("--no-topology", action="store_true",
                    help="skip the topology step and the PLY save/load (ten more batched steps on a copy of "
                         "the scene, then one densify/prune step)")
    ap.add_argument("--tile", type=int, default=32)
    ap.add_argument("--batch-depth", type=int, default=128)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (more GPUs would split the "
              "views of a step across ranks and reduce the Gaussian gradients; that path is designed, not built)",
              file=sys.stderr)
        return 2

    import torch

    from rustscan_amd import scene
    from rustscan_amd.rasterizer import Context, DeviceSplats, Trainer, render_forward, training_optimizer

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev_name = torch.cuda.get_device_name(0)
    log(f"device: {dev_name}")
    t_setup = time.perf_counter()
    init = scene.gaussian_scene(args.n, seed=5, sh_degree=3)
    gt = scene.gaussian_scene(args.n, seed=6, sh_degree=3)
    cams = scene.config4_cameras()
    ctx = Context(0)
    ctx.set_raster(args.tile, args.batch_depth, True)
    gt_dev = DeviceSplats(ctx, gt)
    imgs = render_forward(ctx, gt_dev, cams)
    targets = [im[..., :3].contiguous() for im in imgs]
    gt_dev.close()
    del imgs
    torch.cuda.synchronize()
    opt = training_optimizer(3, lr_position=1.6e-4, lr_rotation=1e-3, lr_scale=5e-3, lr_opacity=5e-2,
                             lr_sh_dc=2.5e-3, lr_sh_rest=1.25e-4)
    trainer = Trainer(ctx, opt, cov_blur=0.3, loss_weight=1.0 / len(cams), densify=RGS_DENSIFY_ACCUMULATE)
    splats = DeviceSplats(ctx, init)
    log(f"setup {time.perf_counter() - t_setup:.1f}s")

    tuned = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tuned = trainer.autotune(splats, cams, targets, batched=True, graph=True)
        log(f"autotune -> tile {tuned[0]}, batch depth {tuned[1]} ({time.perf_counter() - t0:.1f}s)")
    tw, bd, _ = ctx.raster()

    # ---- batched, graph-captured: the measured value
    for _ in range(args.warmup):
        trainer.train_step(splats, cams, targets, batched=True, graph=True)
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        trainer.train_step(splats, cams, targets, batched=True, graph=True)
    ctx.sync()
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t0
    skipped = ctx.skipped_steps()
    batched_sps = args.steps / t_batched
    log(f"batched+graph: {batched_sps:.2f} steps/s ({1e3 * t_batched / args.steps:.2f} ms/step), skipped {skipped}")

    # ---- plain per-view launches (no graph)
    pv_steps = args.per_view_steps or args.steps
    for _ in range(max(1, args.warmup)):
        trainer.train_step(splats, cams, targets, batched=False, graph=False)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(pv_steps):
        trainer.train_step(splats, cams, targets, batched=False, graph=False)
    ctx.sync()
    t_pv = time.perf_counter() - t0
    pv_sps = pv_steps / t_pv
    log(f"per-view: {pv_sps:.2f} steps/s ({1e3 * t_pv / pv_steps:.2f} ms/step)")

    # ---- the same batched step with RustGS's default loss (L1 0.8 + D-SSIM 0.2,
    # loss.rs:112-142, config.rs:595-608) in place of the configurations' L1
    from rustscan_amd.rasterizer import LossConfig

    state0 = splats.download()
    ssim_trainer = Trainer(ctx, opt, cov_blur=0.3, loss_weight=1.0 / len(cams), densify=RGS_DENSIFY_ACCUMULATE,
                           loss=LossConfig())
    for _ in range(args.warmup):
        ssim_trainer.train_step(splats, cams, targets, batched=True, graph=True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ssim_trainer.train_step(splats, cams, targets, batched=True, graph=True)
    ctx.sync()
    t_ssim = time.perf_counter() - t0
    default_loss = {"loss": "L1 0.8 + D-SSIM 0.2", "steps_per_s": round(args.steps / t_ssim, 2),
                    "ms_per_step": round(1e3 * t_ssim / args.steps, 3)}
    log(f"batched+graph with L1 0.8 + D-SSIM 0.2: {default_loss}")
    splats.upload(state0)

    # ---- gradient agreement batched vs per-view from identical parameters
    state = splats.download()
    splats.upload(state)
    trainer.train_step(splats, cams, targets, batched=True, graph=False)
    gb = splats.grads()
    splats.upload(state)
    trainer.train_step(splats, cams, targets, batched=False, graph=False)
    gp = splats.grads()
    grad_err_bp = max(rel_l2(gb[0], gp[0]), rel_l2(gb[1], gp[1]), rel_l2(gb[2], gp[2]))
    log(f"grad rel-L2 batched vs per-view: {grad_err_bp:.2e}")

    # ---- instrumented eager batched pass: per-stage times and fragment counts
    stages = {}
    frags = {}
    counts = []
    if not args.no_instrument:
        # timing pass: the production kernels, HIP events around each stage
        splats.upload(state)
        ctx.set_timing(True)
        ctx.timing(reset=True)
        for _ in range(args.steps):
            trainer.train_step(splats, cams, targets, batched=True, graph=False)
        stages = ctx.timing(reset=True)
        counts = ctx.counts(len(cams))
        ctx.set_timing(False)
        # counting pass: one step with the fragment-counting kernel variants
        splats.upload(state)
        ctx.set_debug(count_fragments=True)
        ctx.fragment_counts(reset=True)
        trainer.train_step(splats, cams, targets, batched=True, graph=False)
        frags = ctx.fragment_counts(reset=True)
        ctx.set_debug()
        for k, (ms, calls) in stages.items():
            log(f"  stage {k:12s} {ms / max(calls, 1):8.3f} ms/launch  ({calls} launches)")
        log(f"  fragments per step: {frags}")

    roofline = None
    stage_summary = {}
    if stages:
        n = args.n
        V = sum(c[0] for c in counts)
        M = sum(c[1] for c in counts)
        nv = len(cams)
        k = 16
        per_launch = {}
        for name, (ms, calls) in stages.items():
            if calls:
                per_launch[name] = ms / calls
        # algorithmic bytes (HBM-bound stages) / flops (compositing) per launch
        sort_passes_depth = 4
        tile_bits = max(1, int(np.ceil(np.log2(max(2, sum(c[2] for c in counts))))))
        sort_passes_tile = (tile_bits + 7) // 8
        inst_bits = int(np.ceil(np.log2(max(2, n * len(counts)))))
        # two passes with a packed middle array (keys of 9..16 bits): key+value read, one word written,
        # one word read, key+value written, plus one key word read per pass by the histogram
        tile_sort_bytes = (32 * M if 8 < tile_bits <= 16 and tile_bits - 8 + inst_bits <= 32
                           else 20 * M * sort_passes_tile)
        T = sum(c[2] for c in counts)
        work = {
            # parameters read once per Gaussian; ProjSplat + key + value + count written per instance
            "project_fwd": ("hbm", (11 + 3 * k) * 4 * n + (64 + 12) * n * nv),
            "depth_sort": ("hbm", 20 * n * nv * sort_passes_depth + 4 * n * nv),
            "count_scan": ("hbm", 12 * n * nv),
            # ProjSplat + order + offset read per visible instance; (key, value) written per intersection
            "emit": ("hbm", 72 * V + 8 * M),
            "tile_sort": ("hbm", tile_sort_bytes),
            "tile_ranges": ("hbm", 4 * M + 8 * T),
            # 28 FP32 operations per evaluated fragment (sigma 11, exp 1, alpha 2, tests 2, blend 9,
            # transmittance 3); about 95 per contributing fragment in the backward
            "raster_fwd": ("fp32-valu", 28 * frags.get("fwd_evaluated", 0)),
            "raster_bwd": ("fp32-valu", 95 * frags.get("bwd_contributing", 0)),
            # accumulator + key per instance; parameters read, gradients and statistics written per Gaussian
            "project_bwd": ("hbm", 68 * n * nv + ((11 + 3 * k) * 8 + 28) * n),
            "adam": ("hbm", 28 * (11 + 3 * k) * n),
        }
        for name, ms in per_launch.items():
            bound, amount = work[name]
            if bound == "hbm":
                ach = amount / (ms * 1e-3) / 1e9
                stage_summary[name] = dict(ms=round(ms, 4), bound="hbm", achieved=round(ach, 1), unit="GB/s",
                                           frac=round(ach / HBM_PEAK_GBS, 4))
            else:
                ach = amount / (ms * 1e-3) / 1e12
                stage_summary[name] = dict(ms=round(ms, 4), bound="fp32-valu", achieved=round(ach, 2),
                                           unit="TFLOP/s", frac=round(ach / FP32_PEAK_TFLOPS, 4),
                                           frac_nonpacked=round(ach / FP32_NONPACKED_TFLOPS, 4))
        if "raster_fwd" in per_launch and frags:
            stage_summary["raster_fwd"]["blended_fragments_per_s"] = round(
                frags.get("fwd_blended", 0) / (per_launch["raster_fwd"] * 1e-3), 1)
            # share of the 64 lanes of a (wave, splat) iteration whose pixel is still being composited
            stage_summary["raster_fwd"]["active_lane_fraction"] = round(
                frags.get("fwd_evaluated", 0) / max(1, 64 * frags.get("fwd_wave_iters", 0)), 4)
        if "raster_bwd" in per_launch and frags:
            stage_summary["raster_bwd"]["active_lane_fraction"] = round(
                frags.get("bwd_evaluated", 0) / max(1, 64 * frags.get("bwd_wave_iters", 0)), 4)
        if "raster_bwd" in per_launch and frags:
            stage_summary["raster_bwd"]["contributing_fragments_per_s"] = round(
                frags.get("bwd_contributing", 0) / (per_launch["raster_bwd"] * 1e-3), 1)
        dom = max(per_launch, key=lambda s: per_launch[s])
        d = stage_summary[dom]
        kname = STAGE_KERNELS[dom].format(tw=tw, bd=bd)
        traffic = measured_traffic(kname)
        roofline = {
            "kernel": dom,
            "bound": d["bound"],
            "achieved": d["achieved"],
            "peak": HBM_PEAK_GBS if d["bound"] == "hbm" else FP32_PEAK_TFLOPS,
            "unit": d["unit"],
            "frac": d["frac"],
            "traffic": traffic,
            "kernel_name": kname,
            "ms_per_launch": d["ms"],
            "note": "algorithmic bytes per launch / HIP-event launch time" if d["bound"] == "hbm" else
            "FP32 vector ALU (no MFMA: compositing has no dense contraction); peak is the published FP32 figure, "
            f"which counts packed FP32; against the non-packed {FP32_NONPACKED_TFLOPS} TFLOP/s that this build "
            f"can reach, frac = {d.get('frac_nonpacked')}",
        }
        if d["bound"] != "hbm":
            roofline["frac_nonpacked"] = d.get("frac_nonpacked")

    # ---- absolute pixel rates: pixels trained per second (batched step), and
    # pixels rendered per second by the forward stages alone (eager, HIP events)
    pixels = sum(c.width * c.height for c in cams)
    mpix = {"train_step": round(pixels * batched_sps / 1e6, 1)}
    if stages:
        fwd_ms = sum(per_launch.get(k, 0.0) for k in
                     ("project_fwd", "depth_sort", "count_scan", "emit", "tile_sort", "tile_ranges", "raster_fwd"))
        if fwd_ms > 0:
            mpix["forward_render"] = round(pixels / (fwd_ms * 1e-3) / 1e6, 1)
    mpix["pixels_per_step"] = pixels
    log(f"GPU Mpix/s: {mpix}")

    # ---- evaluation render on the device (the reference's CPU renderer
    # semantics, GaussianRenderer::render_splats) and one evaluation frame
    # (render + PSNR and sharpness), wall clock per synchronous call
    eval_render = run_eval_render(ctx, state, cams[2], targets[2])

    # ---- the reference's CPU renderer (render/tiled_renderer.rs restated, single-threaded)
    cpu_render = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        cam = cams[2]  # 1280x720
        t0 = time.perf_counter()
        ref_color, ref_depth = O.tiled_renderer(state, cam, cov_blur=0.3)
        t_r = time.perf_counter() - t0
        cpu_render = {"value": round(cam.width * cam.height / t_r / 1e6, 4), "unit": "Mpix/s", "cores": 1,
                      "kind": "port", "sample": f"one {cam.width}x{cam.height} forward render of the 1,048,576-"
                                                "Gaussian scene by the CPU tiled renderer"}
        log(f"cpu tiled renderer: {t_r:.1f}s for one {cam.width}x{cam.height} view")
        if eval_render is not None:
            # the device render against the reference arithmetic (libm expf): equal bit for bit
            gc, gd = eval_render.pop("_images")
            eval_render["vs_cpu_renderer"] = {
                "max_abs_color": float(np.abs(gc - ref_color).max()),
                "color_values_differing": int((gc != ref_color).sum()),
                "depth_pixels_differing": int((gd != ref_depth).sum()),
                "pixels": int(cam.width * cam.height),
                "scene": "the parameters after the timed training steps",
            }
            eval_render["gpu_over_cpu_renderer"] = round(eval_render["render_mpix_per_s"] / cpu_render["value"], 1)
            log(f"eval render vs cpu renderer: {eval_render['vs_cpu_renderer']}")
    if eval_render is not None:
        eval_render.pop("_images", None)

    # ---- CPU baseline: the oracle for one full step, and gradient agreement with it
    cpu = None
    grad_err_oracle = None
    grad_err_groups = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        tg_host = [t.cpu().numpy() for t in targets]
        splats.upload(state)
        trainer.train_step(splats, cams, targets, batched=True, graph=False)
        g_gpu = splats.grads()
        t0 = time.perf_counter()
        ref_vt = np.zeros((args.n, 10), np.float32)
        ref_vs = np.zeros((args.n, 16, 3), np.float32)
        ref_vo = np.zeros(args.n, np.float32)
        for cam, tg in zip(cams, tg_host):
            r = O.Render(state, cam, cov_blur=0.3)
            _, v = O.l1_loss_grad(r.image(), tg, weight=1.0 / len(cams))
            g = r.backward(v, 0)
            ref_vt += g["v_transforms"]
            ref_vs += g["v_sh"]
            ref_vo += g["v_raw_opac"]
            del r
        p = state.copy()
        m = np.zeros((args.n, 59), np.float32)
        vv = np.zeros((args.n, 59), np.float32)
        flat = np.concatenate([p.transforms, p.raw_opac[:, None], p.sh.reshape(args.n, -1)], axis=1).copy()
        gflat = np.concatenate([ref_vt, ref_vo[:, None], ref_vs.reshape(args.n, -1)], axis=1).copy()
        lr = np.array([1.6e-4] * 3 + [1e-3] * 4 + [5e-3] * 3 + [5e-2] + [2.5e-3] * 3 + [1.25e-4] * 45, np.float32)
        O.adam_group(flat, gflat, m, vv, lr, 0.9, 0.999, 1e-8, 1)
        t_cpu = time.perf_counter() - t0
        grad_err_groups = {
            "means": rel_l2(g_gpu[0][:, 0:3], ref_vt[:, 0:3]),
            "quats": rel_l2(g_gpu[0][:, 3:7], ref_vt[:, 3:7]),
            "log_scales": rel_l2(g_gpu[0][:, 7:10], ref_vt[:, 7:10]),
            "opacity": rel_l2(g_gpu[2], ref_vo),
            "sh": rel_l2(g_gpu[1], ref_vs),
        }
        grad_err_oracle = max(grad_err_groups.values())
        cpu = {
            "value": round(1.0 / t_cpu, 5),
            "unit": "steps/s",
            "cores": int(O.num_threads()),
            "kind": "port",
            "sample": "one full 8-view step (1,048,576 Gaussians, fwd+L1+bwd+Adam) of the oracle, OpenMP over tiles",
        }
        log(f"cpu baseline: {t_cpu:.1f}s per step on {cpu['cores']} threads; grad rel-L2 GPU vs oracle "
            + " ".join(f"{k}={e:.2e}" for k, e in grad_err_groups.items()))

    out = {
        "metric": METRIC,
        "value": round(batched_sps, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_batched / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "config4: 1,048,576 Gaussians SH3, 8 views {640x480, 960x540, 1280x720 x3, 1600x900, "
                        "1920x1080 x2}, fwd+L1+bwd+Adam per step",
            "mode": "batched, HIP graph",
            "tile_width": tw,
            "batch_depth": bd,
        },
        "per_view_steps_per_s": round(pv_sps, 3),
        "batched_over_per_view": round(batched_sps / pv_sps, 3),
        "grad_rel_err_batched_vs_per_view": grad_err_bp,
        "grad_rel_err_vs_oracle": grad_err_oracle,
        "grad_rel_err_vs_oracle_per_group": grad_err_groups,
        "gpu_mpix_per_s": mpix,
        "skipped_steps": skipped,
        "default_loss_step": default_loss,
        "stages": stage_summary,
        "counts": {"visible": int(sum(c[0] for c in counts)), "intersections": int(sum(c[1] for c in counts)),
                   "tiles": int(sum(c[2] for c in counts))} if counts else None,
        "fragments_per_step": frags or None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_tiled_renderer": cpu_render,
        "evaluation_render": eval_render,
        "topology_and_files": None if args.no_topology else run_topology_and_files(ctx, opt, state, cams, targets),
        "config3": None if args.no_config3 else run_config3(ctx, opt, init, gt),
        "config5": None if args.no_config5 else run_config5(args, ctx, opt),
        "device": dev_name,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
