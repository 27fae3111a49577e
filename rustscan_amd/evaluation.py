"""Evaluation path: the reference's CPU renderer and per-frame metrics on the
device, plus the host-side bookkeeping of RustGS's evaluate_splats.

Device (through the C ABI, librgs_hip.so):
  TiledRenderer.render_camera_splats     render/tiled_renderer.rs:174-492
  GaussianRenderer.render_splats         render/renderer.rs:72-98
  SplatEvaluationRenderer.render         training/evaluation/core.rs:349-377
  image_metrics / evaluate_frame         evaluation/core.rs:238-319, 431-449

Host (plain Python over a handful of floats, as in the reference):
  summarize_training_metrics             core.rs:19-37
  select_evaluation_frames               core.rs:178-199 (on a list of frames)
  summarize_psnr_samples                 core.rs:201-238
  worst_frame_metrics                    core.rs:221-236
  scaled_dimensions                      core.rs:321-326
  scaled_camera                          core.rs:503-523 (scaled_camera_for_pose)
  evaluate_splats                        core.rs:402-497
The float32 summaries add left to right in float32 as the reference does.
"""
from __future__ import annotations

import ctypes
import time
from ctypes import c_float
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .rasterizer import Context, DeviceSplats, _torch, to_rgs_camera
from .scene import Camera

DEFAULT_RASTER_COV_BLUR = 0.3  # training::DEFAULT_RASTER_COV_BLUR
MIN_RENDER_SCALE = 0.0625  # evaluation/core.rs:11
F32_MAX = np.float32(3.4028235e38)

_F = np.float32


def _bg(background):
    return (c_float * 3)(*[float(x) for x in background])


class TiledRenderer:
    """TiledRenderer (render/tiled_renderer.rs:40-72): 16x16 tiles, blur floor
    max(raster_cov_blur, 0)."""

    def __init__(self, width: int, height: int):
        self.width, self.height = int(width), int(height)
        self.raster_cov_blur = DEFAULT_RASTER_COV_BLUR

    def with_raster_cov_blur(self, raster_cov_blur: float) -> "TiledRenderer":
        self.raster_cov_blur = max(float(raster_cov_blur), 0.0)
        return self

    def render_camera_splats(self, ctx: Context, splats: DeviceSplats, camera: Camera, background=(0.0, 0.0, 0.0)):
        """(color [H,W,3] float32 in [0,1], depth [H,W], f32::MAX where empty)
        as device tensors; the image size is the renderer's."""
        torch = _torch()
        cam = _sized(camera, self.width, self.height)
        color = torch.empty((self.height, self.width, 3), dtype=torch.float32, device="cuda")
        depth = torch.empty((self.height, self.width), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        c = to_rgs_camera(cam)
        ctx.check(ctx.lib.rgs_tiled_render(ctx.h, splats.h, ctypes.byref(c), _bg(background),
                                           float(self.raster_cov_blur), color.data_ptr(), depth.data_ptr()))
        return color, depth


@dataclass
class RenderOutput:
    """render/renderer.rs:21-33 (normal is never produced)."""

    color: object  # [H,W,3] uint8 device tensor
    depth: object  # [H,W] float32, 0 where empty
    width: int
    height: int
    normal: object = None


class GaussianRenderer:
    """GaussianRenderer (render/renderer.rs:35-98)."""

    def __init__(self, width: int, height: int):
        self.width, self.height = int(width), int(height)
        self.background = (0.0, 0.0, 0.0)
        self.raster_cov_blur = DEFAULT_RASTER_COV_BLUR

    def with_background(self, r: float, g: float, b: float) -> "GaussianRenderer":
        self.background = (float(r), float(g), float(b))
        return self

    def with_raster_cov_blur(self, raster_cov_blur: float) -> "GaussianRenderer":
        self.raster_cov_blur = max(float(raster_cov_blur), 0.0)
        return self

    def render_splats(self, ctx: Context, splats: DeviceSplats, camera: Camera) -> RenderOutput:
        torch = _torch()
        cam = _sized(camera, self.width, self.height)
        color = torch.empty((self.height, self.width, 3), dtype=torch.uint8, device="cuda")
        depth = torch.empty((self.height, self.width), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        c = to_rgs_camera(cam)
        ctx.check(ctx.lib.rgs_render_splats(ctx.h, splats.h, ctypes.byref(c), _bg(self.background),
                                            float(self.raster_cov_blur), color.data_ptr(), depth.data_ptr()))
        return RenderOutput(color, depth, self.width, self.height)


class SplatEvaluationRenderer:
    """SplatEvaluationRenderer (evaluation/core.rs:343-378): GaussianRenderer
    with a black background; render() returns the u8 colour / 255."""

    def __init__(self, render_width: int, render_height: int, raster_cov_blur: float = DEFAULT_RASTER_COV_BLUR):
        self.renderer = GaussianRenderer(render_width, render_height).with_raster_cov_blur(raster_cov_blur)

    def render(self, ctx: Context, splats: DeviceSplats, camera: Camera):
        torch = _torch()
        out = self.renderer.render_splats(ctx, splats, camera)
        # `value as f32 / 255.0` as a table of the 256 correctly rounded
        # quotients (a tensor division by a scalar may multiply by 1/255)
        lut = torch.from_numpy(np.arange(256, dtype=np.float32) / np.float32(255.0)).to(out.color.device)
        return lut[out.color.long()]


def _sized(camera: Camera, width: int, height: int) -> Camera:
    if camera.width == width and camera.height == height:
        return camera
    return Camera(camera.viewmat, camera.fx, camera.fy, camera.cx, camera.cy, width, height, camera.cam_pos)


def _metrics_dict(m: L.RgsFrameMetrics) -> dict:
    return {name: float(getattr(m, name)) for name, _ in L.RgsFrameMetrics._fields_}


def image_metrics(ctx: Context, rendered, target) -> dict:
    """PSNR (compute_psnr_f32) and the gradient / Laplacian sharpness of both
    images with their ratios, for rendered and target [H,W,3] float32 device
    tensors.  Mismatched sizes give PSNR 0 as in the reference (and no
    sharpness)."""
    torch = _torch()
    if tuple(rendered.shape) != tuple(target.shape):
        return {name: 0.0 for name, _ in L.RgsFrameMetrics._fields_}
    h, w = int(rendered.shape[0]), int(rendered.shape[1])
    rendered = rendered.contiguous().float()
    target = target.contiguous().float()
    torch.cuda.synchronize()
    m = L.RgsFrameMetrics()
    ctx.check(ctx.lib.rgs_image_metrics(ctx.h, rendered.data_ptr(), target.data_ptr(), w, h, ctypes.byref(m)))
    return _metrics_dict(m)


def evaluate_frame(ctx: Context, splats: DeviceSplats, camera: Camera, target, raster_cov_blur=DEFAULT_RASTER_COV_BLUR,
                   keep_color: bool = False):
    """Render one frame as SplatEvaluationRenderer does and measure it against
    target [H,W,3] (device, float32): (metrics dict, u8 colour or None)."""
    torch = _torch()
    target = target.contiguous().float()
    h, w = int(target.shape[0]), int(target.shape[1])
    if (int(camera.width), int(camera.height)) != (w, h):
        raise ValueError(f"camera is {camera.width}x{camera.height}, target {w}x{h}")
    cam = camera
    color = torch.empty((h, w, 3), dtype=torch.uint8, device="cuda") if keep_color else None
    torch.cuda.synchronize()
    m = L.RgsFrameMetrics()
    c = to_rgs_camera(cam)
    ctx.check(ctx.lib.rgs_evaluate_frame(ctx.h, splats.h, ctypes.byref(c), float(raster_cov_blur), target.data_ptr(),
                                         color.data_ptr() if color is not None else None, ctypes.byref(m)))
    return _metrics_dict(m), color


# ---------------------------------------------------------------- host logic


def _f32_sum(values) -> np.float32:
    a = np.asarray(values, _F).ravel()
    return _F(0.0) if a.size == 0 else _F(np.cumsum(a, dtype=_F)[-1])


@dataclass
class PsnrSummary:
    mean_db: float
    median_db: float
    min_db: float
    max_db: float
    stddev_db: float


def summarize_psnr_samples(values) -> PsnrSummary:
    """core.rs:201-238: mean, median (average of the middle pair), min, max
    and population standard deviation, in float32."""
    vals = [_F(v) for v in values]
    srt = sorted(vals, key=lambda v: (np.isnan(v), v))
    n = len(vals)
    mean = _f32_sum(vals) / _F(max(n, 1))
    if n == 0:
        median = _F(0.0)
    elif n % 2 == 0:
        median = (srt[n // 2 - 1] + srt[n // 2]) * _F(0.5)
    else:
        median = srt[n // 2]
    lo = srt[0] if n else _F(0.0)
    hi = srt[-1] if n else _F(0.0)
    var = _f32_sum([(v - mean) * (v - mean) for v in vals]) / _F(max(n, 1))
    return PsnrSummary(float(mean), float(median), float(lo), float(hi), float(np.sqrt(_F(var))))


def summarize_training_metrics(loss_history, frame_count: int) -> dict:
    """core.rs:19-37: mean of the last max(frame_count, 1) losses and the
    last loss."""
    if len(loss_history) == 0:
        return dict(final_loss=0.0, final_step_loss=0.0)
    window = min(max(int(frame_count), 1), len(loss_history))
    final = _f32_sum(loss_history[len(loss_history) - window:]) / _F(window)
    return dict(final_loss=float(final), final_step_loss=float(_F(loss_history[-1])))


def select_evaluation_frames(frames: list, max_frames: int = 0, frame_stride: int = 1) -> list:
    """core.rs:178-199: the first max_frames frames (all for 0), every
    frame_stride-th of them."""
    take = frames if max_frames == 0 else frames[:max_frames]
    return take[::max(int(frame_stride), 1)]


def worst_frame_metrics(frame_metrics: list, count: int) -> list:
    """core.rs:221-236: the `count` lowest-PSNR frames, lowest first (stable)."""
    srt = sorted(frame_metrics, key=lambda m: m["psnr_db"])
    return srt[:min(count, len(frame_metrics))]


def scaled_dimensions(width: int, height: int, render_scale: float) -> tuple:
    """core.rs:321-326: round(side * clamp(scale, 1/16, 1)), at least 1."""
    scale = min(max(_F(render_scale), _F(MIN_RENDER_SCALE)), _F(1.0))
    sw = max(_round_half_away(_F(width) * scale), 1.0)
    sh = max(_round_half_away(_F(height) * scale), 1.0)
    return int(sw), int(sh)


def _round_half_away(x) -> float:
    """f32::round: halfway cases away from zero."""
    x = float(x)
    return float(np.copysign(np.floor(abs(x) + 0.5), x))


def scaled_camera(camera: Camera, src_width: int, src_height: int, dst_width: int, dst_height: int) -> Camera:
    """scaled_camera_for_pose (core.rs:503-523): intrinsics scaled by
    dst/src per axis, same pose."""
    sx = _F(dst_width) / _F(src_width)
    sy = _F(dst_height) / _F(src_height)
    return Camera(camera.viewmat, _F(camera.fx) * sx, _F(camera.fy) * sy, _F(camera.cx) * sx, _F(camera.cy) * sy,
                  int(dst_width), int(dst_height), camera.cam_pos)


@dataclass
class SplatEvaluationConfig:
    """core.rs:66-87 defaults."""

    render_scale: float = 0.5
    raster_cov_blur: float = DEFAULT_RASTER_COV_BLUR
    frame_stride: int = 1
    max_frames: int = 0
    worst_frame_count: int = 5


@dataclass
class SplatEvaluationResult:
    summary: dict
    frame_metrics: list = field(default_factory=list)


def evaluate_splats(ctx: Context, splats: DeviceSplats, cameras: list, targets: list,
                    config: SplatEvaluationConfig | None = None, training_metrics: dict | None = None,
                    final_loss: float = 0.0, iterations: int = 0) -> SplatEvaluationResult:
    """evaluate_splats (core.rs:402-497) over full-size cameras and targets:
    frames selected by max_frames / frame_stride, each rendered at
    scaled_dimensions(render_scale) with scaled intrinsics against its target
    resized on the device (bilinear, in place of the reference's image-crate
    resize of the file), then summarised."""
    torch = _torch()
    config = config or SplatEvaluationConfig()
    idx = select_evaluation_frames(list(range(len(cameras))), config.max_frames, config.frame_stride)
    if not idx:
        raise ValueError("evaluation dataset resolved to zero frames")
    W, H = int(cameras[idx[0]].width), int(cameras[idx[0]].height)
    rw, rh = scaled_dimensions(W, H, config.render_scale)
    start = time.perf_counter()
    frames = []
    for k, i in enumerate(idx):
        cam = scaled_camera(cameras[i], W, H, rw, rh)
        tgt = targets[i]
        if tuple(tgt.shape[:2]) != (rh, rw):
            tgt = torch.nn.functional.interpolate(tgt.permute(2, 0, 1)[None].float(), size=(rh, rw),
                                                  mode="bilinear", align_corners=False)[0].permute(1, 2, 0)
        m, _ = evaluate_frame(ctx, splats, cam, tgt, config.raster_cov_blur)
        frames.append(dict(dataset_index=k, frame_id=i, psnr_db=m["psnr_db"],
                           sharpness_grad_ratio=m["sharpness_grad_ratio"],
                           sharpness_lap_ratio=m["sharpness_lap_ratio"]))
    ps = summarize_psnr_samples([f["psnr_db"] for f in frames])
    summary = dict(
        render_scale=config.render_scale, raster_cov_blur=config.raster_cov_blur, render_width=rw, render_height=rh,
        frame_stride=config.frame_stride, max_frames=config.max_frames, frame_count=len(frames),
        splat_iterations=iterations, splat_count=splats.n,
        final_loss=(training_metrics or {}).get("final_loss", final_loss),
        final_step_loss=(training_metrics or {}).get("final_step_loss"),
        elapsed_seconds=time.perf_counter() - start,
        psnr_mean_db=ps.mean_db, psnr_median_db=ps.median_db, psnr_min_db=ps.min_db, psnr_max_db=ps.max_db,
        psnr_std_db=ps.stddev_db,
        sharpness_grad_ratio_mean=float(_f32_sum([f["sharpness_grad_ratio"] for f in frames]) / _F(len(frames))),
        sharpness_lap_ratio_mean=float(_f32_sum([f["sharpness_lap_ratio"] for f in frames]) / _F(len(frames))),
        worst_frames=worst_frame_metrics(frames, config.worst_frame_count),
    )
    return SplatEvaluationResult(summary, frames)
