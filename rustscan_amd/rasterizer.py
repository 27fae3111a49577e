"""Host-side mirror of RustGS's rasterizer/trainer interface over librgs_hip.

Names follow the reference so parity tests read like its own tests:

  render_forward                 training/forward/mod.rs:156-294
  render_splats_with_visibility  training/backward/autodiff.rs:240-251
  RenderBackward.backward        training/backward/autodiff.rs:85-141
  AdamScaled                     training/engine/optimizer.rs:7-163
  Trainer.train_step             training/engine/trainer.rs:253-440
  lr_at                          training/engine/trainer.rs:179-191

Device memory for images and targets is held in torch tensors on the HIP
device; all computation happens in the HIP library.
"""
from __future__ import annotations

import ctypes
import math
from ctypes import c_float, c_int64, c_void_p
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .scene import Camera, Splats, sh_coeffs


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device available: the render path runs only on the GPU")
    return torch


def to_rgs_camera(cam: Camera) -> L.RgsCamera:
    c = L.RgsCamera()
    for i in range(16):
        c.viewmat[i] = float(cam.viewmat[i])
    c.fx, c.fy, c.cx, c.cy = float(cam.fx), float(cam.fy), float(cam.cx), float(cam.cy)
    c.width, c.height = int(cam.width), int(cam.height)
    for i in range(3):
        c.cam_pos[i] = float(cam.cam_pos[i])
    return c


def _cams(cams) -> tuple:
    arr = (L.RgsCamera * len(cams))(*[to_rgs_camera(c) for c in cams])
    return arr, len(cams)


def _ptrs(tensors) -> ctypes.Array:
    return (c_void_p * len(tensors))(*[t.data_ptr() if t is not None else None for t in tensors])


class Context:
    """One HIP stream and workspace (rgs_ctx)."""

    def __init__(self, device: int = 0):
        _torch()
        self.lib = L.lib()
        h = c_void_p()
        L.check(None, self.lib.rgs_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            self.lib.rgs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, code):
        L.check(self.h, code)

    def sync(self):
        self.check(self.lib.rgs_sync(self.h))

    def set_raster(self, tile_width=16, batch_depth=256, persistent=True):
        cfg = L.RgsRasterCfg(tile_width, batch_depth, 1 if persistent else 0)
        self.check(self.lib.rgs_set_raster_config(self.h, ctypes.byref(cfg)))

    def set_launch_limit(self, max_intersections: int):
        """Most intersections per batched launch; larger steps run as groups of views."""
        self.check(self.lib.rgs_set_launch_limit(self.h, int(max_intersections)))

    def raster(self) -> tuple:
        cfg = L.RgsRasterCfg()
        self.check(self.lib.rgs_get_raster_config(self.h, ctypes.byref(cfg)))
        return cfg.tile_width, cfg.batch_depth, bool(cfg.persistent)

    def debug(self, what: int, view: int = 0, dtype=np.uint32) -> np.ndarray:
        cnt = c_int64()
        self.check(self.lib.rgs_debug_get(self.h, what, view, None, ctypes.byref(cnt)))
        out = np.zeros(cnt.value, dtype)
        if cnt.value:
            self.check(self.lib.rgs_debug_get(self.h, what, view, out.ctypes.data, ctypes.byref(cnt)))
        return out

    def counts(self, n_views: int):
        arr = (L.RgsCounts * n_views)()
        self.check(self.lib.rgs_get_counts(self.h, arr, n_views))
        return [(a.num_visible, a.num_intersections, a.num_tiles) for a in arr]

    def skipped_steps(self) -> int:
        return int(self.lib.rgs_skipped_steps(self.h))

    def set_timing(self, enable: bool):
        self.check(self.lib.rgs_set_timing(self.h, 1 if enable else 0))

    def timing(self, reset: bool = True) -> dict:
        """Accumulated per-stage milliseconds and launch counts (eager launches only)."""
        n = len(L.STAGES)
        ms = (ctypes.c_double * n)()
        calls = (c_int64 * n)()
        self.check(self.lib.rgs_get_timing(self.h, ms, calls, 1 if reset else 0))
        return {name: (ms[i], calls[i]) for i, name in enumerate(L.STAGES)}

    def set_debug(self, keep_raster_grads: bool = False, count_fragments: bool = False, plain_sorts: bool = False,
                  no_cull: bool = False):
        """Debug switches (rgs_set_debug): keep the rasterizer accumulators,
        count fragments, plain (unpacked) sorts, no block culling."""
        flags = (1 if keep_raster_grads else 0) | (2 if count_fragments else 0) | (4 if plain_sorts else 0) | \
            (8 if no_cull else 0)
        self.check(self.lib.rgs_set_debug(self.h, flags))

    def fragment_counts(self, reset: bool = True) -> dict:
        out = (c_int64 * 8)()
        self.check(self.lib.rgs_get_fragment_counts(self.h, out, 1 if reset else 0))
        return dict(fwd_evaluated=out[0], fwd_blended=out[1], bwd_evaluated=out[2], bwd_contributing=out[3],
                    fwd_wave_iters=out[4], fwd_wave_iters_exp=out[5], bwd_wave_iters=out[6],
                    bwd_wave_iters_contrib=out[7])


class DeviceSplats:
    """GPU-resident splat store (DeviceSplats, engine/splats.rs:13-63)."""

    def __init__(self, ctx: Context, host: Splats):
        self.ctx = ctx
        self.n = host.n
        self.sh_degree = host.sh_degree
        self.k = sh_coeffs(host.sh_degree)
        h = c_void_p()
        ctx.check(ctx.lib.rgs_splats_create(ctx.h, self.n, self.sh_degree, ctypes.byref(h)))
        self.h = h
        self.upload(host)

    @classmethod
    def adopt(cls, ctx: Context, handle: c_void_p) -> "DeviceSplats":
        """Wraps a store the library created (e.g. rgs_scene_to_splats)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.h = handle
        self.n = int(ctx.lib.rgs_splats_count(handle))
        self.sh_degree = int(ctx.lib.rgs_splats_sh_degree(handle))
        self.k = sh_coeffs(self.sh_degree)
        return self

    def upload(self, host: Splats):
        tf = np.ascontiguousarray(host.transforms, np.float32)
        sh = np.ascontiguousarray(host.sh, np.float32)
        op = np.ascontiguousarray(host.raw_opac, np.float32)
        self.ctx.check(self.ctx.lib.rgs_splats_import_aos(self.ctx.h, self.h, tf.ctypes.data, sh.ctypes.data,
                                                          op.ctypes.data))

    def download(self) -> Splats:
        tf = np.zeros((self.n, 10), np.float32)
        sh = np.zeros((self.n, self.k, 3), np.float32)
        op = np.zeros(self.n, np.float32)
        self.ctx.check(self.ctx.lib.rgs_splats_export_aos(self.ctx.h, self.h, tf.ctypes.data, sh.ctypes.data,
                                                          op.ctypes.data))
        return Splats(tf, sh, op, self.sh_degree)

    def grads(self):
        vt = np.zeros((self.n, 10), np.float32)
        vs = np.zeros((self.n, self.k, 3), np.float32)
        vo = np.zeros(self.n, np.float32)
        st = np.zeros((self.n, 7), np.float32)
        self.ctx.check(self.ctx.lib.rgs_splats_export_grads(self.ctx.h, self.h, vt.ctypes.data, vs.ctypes.data,
                                                            vo.ctypes.data, st.ctypes.data))
        return vt, vs, vo, st

    def adam_moments(self):
        """(moment1, moment2) of AdamState, each (transforms [N,10], sh [N,K,3], raw_opac [N])."""
        arrs = [np.zeros((self.n, 10), np.float32), np.zeros((self.n, self.k, 3), np.float32),
                np.zeros(self.n, np.float32), np.zeros((self.n, 10), np.float32),
                np.zeros((self.n, self.k, 3), np.float32), np.zeros(self.n, np.float32)]
        self.ctx.check(self.ctx.lib.rgs_adam_export_moments(self.ctx.h, self.h, *[a.ctypes.data for a in arrs]))
        return tuple(arrs[:3]), tuple(arrs[3:])

    def set_adam_moments(self, m, v):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (*m, *v)]
        self.ctx.check(self.ctx.lib.rgs_adam_import_moments(self.ctx.h, self.h, *[a.ctypes.data for a in arrs]))

    def densify_stats(self) -> dict:
        """The trainer's densification accumulators (trainer.rs:66-75), [N] each."""
        out = np.zeros((10, self.n), np.float32)
        self.ctx.check(self.ctx.lib.rgs_densify_export(self.ctx.h, self.h, out.ctypes.data))
        return {name: out[i] for i, name in enumerate(L.DENSIFY_FIELDS)}

    def set_densify_stats(self, stats: np.ndarray):
        """Sets the accumulators from a [10, N] array (rgs_densify_export order)."""
        a = np.ascontiguousarray(stats, np.float32).reshape(10, self.n)
        self.ctx.check(self.ctx.lib.rgs_densify_import(self.ctx.h, self.h, a.ctypes.data))

    def reset_densify_stats(self):
        self.ctx.check(self.ctx.lib.rgs_densify_reset(self.ctx.h, self.h))

    def topology_state(self):
        """(birth iteration, invisible windows) per Gaussian (trainer.rs:76-77)."""
        birth = np.zeros(self.n, np.uint32)
        win = np.zeros(self.n, np.uint32)
        self.ctx.check(self.ctx.lib.rgs_topology_export_state(self.ctx.h, self.h, birth.ctypes.data, win.ctypes.data))
        return birth, win

    def set_topology_state(self, birth, invisible_windows):
        b = np.ascontiguousarray(birth, np.uint32)
        w = np.ascontiguousarray(invisible_windows, np.uint32)
        self.ctx.check(self.ctx.lib.rgs_topology_import_state(self.ctx.h, self.h, b.ctypes.data, w.ctypes.data))

    def topology_step(self, cfg: "TopologyConfig", iteration: int, frame_count: int) -> dict:
        """apply_topology_mutations (engine/trainer.rs:792-859) on the device;
        the store may change size."""
        res = L.RgsTopologyResult()
        c = cfg.to_cfg()
        self.ctx.check(self.ctx.lib.rgs_topology_step(self.ctx.h, self.h, ctypes.byref(c), int(iteration),
                                                      int(frame_count), ctypes.byref(res)))
        self.n = int(res.after)
        out = {f: getattr(res, f) for f, _ in L.RgsTopologyResult._fields_ if f != "schedule"}
        out["schedule"] = {f: getattr(res.schedule, f) for f, _ in L.RgsTopologySchedule._fields_}
        return out

    def visible(self) -> np.ndarray:
        out = np.zeros(self.n, np.float32)
        self.ctx.check(self.ctx.lib.rgs_splats_export_visible(self.ctx.h, self.h, out.ctypes.data))
        return out

    def close(self):
        if self.h:
            self.ctx.lib.rgs_splats_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RenderCheckpoint:
    """A forward state of its own (RenderCheckpoint, backward/autodiff.rs:56-71):
    pass it to render_forward and later to render_backward; several may be
    outstanding at once."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = c_void_p()
        ctx.check(ctx.lib.rgs_checkpoint_create(ctx.h, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.rgs_checkpoint_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_forward(ctx: Context, splats: DeviceSplats, cams, background=(0.0, 0.0, 0.0), cov_blur: float = 0.3,
                   checkpoint: RenderCheckpoint | None = None):
    """Renders every camera; returns [H,W,4] float32 torch tensors on the device.
    With a checkpoint the forward state goes there instead of the context."""
    torch = _torch()
    arr, nv = _cams(cams)
    outs = [torch.empty((c.height, c.width, 4), dtype=torch.float32, device="cuda") for c in cams]
    bg = (c_float * 3)(*background)
    torch.cuda.synchronize()
    if checkpoint is None:
        ctx.check(ctx.lib.rgs_render_forward(ctx.h, splats.h, arr, nv, bg, float(cov_blur), _ptrs(outs)))
    else:
        ctx.check(ctx.lib.rgs_render_forward_checkpoint(ctx.h, splats.h, arr, nv, bg, float(cov_blur), _ptrs(outs),
                                                        checkpoint.h))
    return outs


def render_backward(ctx: Context, splats: DeviceSplats, v_outs, checkpoint: RenderCheckpoint | None = None):
    """RenderBackward::backward: image gradients -> splat-store gradients (of
    the context's last forward, or of the checkpoint's)."""
    torch = _torch()
    v_outs = [v.contiguous().float() for v in v_outs]
    torch.cuda.synchronize()
    if checkpoint is None:
        ctx.check(ctx.lib.rgs_render_backward(ctx.h, splats.h, _ptrs(v_outs)))
    else:
        ctx.check(ctx.lib.rgs_render_backward_checkpoint(ctx.h, splats.h, checkpoint.h, _ptrs(v_outs)))
    return splats.grads()


def l1_loss_grad(ctx: Context, pred, target, weight: float = 1.0):
    torch = _torch()
    h, w = int(pred.shape[0]), int(pred.shape[1])
    v_out = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    loss = c_float()
    torch.cuda.synchronize()
    ctx.check(ctx.lib.rgs_l1_loss_grad(ctx.h, pred.data_ptr(), target.contiguous().data_ptr(), w, h, float(weight),
                                       v_out.data_ptr(), ctypes.byref(loss)))
    return loss.value, v_out


@dataclass
class TopologyConfig:
    """The TrainingConfig fields the topology step reads (training/config.rs:
    262-383 LiteGS topology, growth, refine and pruning groups; 666-686
    max_initial_gaussians; 712-721 frame_shuffle_seed; 790-809 iterations),
    with RustGS's defaults.  Enumerations as integers: opacity_reset_mode 0
    decay / 1 reset; split_score_mode 0 baseline / 1 abs / 2 abs-pixel / 3
    abs-pixel-depth; prune_mode 0 threshold / 1 weight / 2 visibility-weight;
    optional epochs -1 for none."""

    iterations: int = 30000
    max_initial_gaussians: int = 100000
    frame_shuffle_seed: int = 0
    refine_every: int = 160
    opacity_reset_interval: int = 10
    opacity_reset_mode: int = 0
    topology_freeze_after_epoch: int = -1
    growth_freeze_after_epoch: int = -1
    prune_until_epoch: int = -1
    split_score_mode: int = 0
    growth_stop_iter: int = 15000
    growth_grad_threshold: float = 0.00014
    split_grad_threshold: float = 0.00014
    depth_scale_gamma: float = 0.37
    growth_select_fraction: float = 0.25
    opacity_decay: float = 0.0
    scale_decay: float = 0.0
    prune_mode: int = 1
    prune_min_age: int = 5
    prune_invisible_epochs: int = 10
    prune_visibility_dry_run: int = 0
    prune_opacity_threshold: float = 1.0 / 255.0
    prune_visibility_threshold: float = 0.05
    prune_high_opacity_threshold: float = 0.8

    def to_cfg(self) -> L.RgsTopologyCfg:
        return L.RgsTopologyCfg(**{f: getattr(self, f) for f, _ in L.RgsTopologyCfg._fields_})

    def should_apply(self, iteration: int, frame_count: int) -> bool:
        """should_apply_topology_step (topology/schedule.rs:97-112)."""
        c = self.to_cfg()
        return bool(L.lib().rgs_topology_should_apply(ctypes.byref(c), int(iteration), int(frame_count)))


@dataclass
class LossConfig:
    """TrainingLossConfig (training/config.rs:566-608) fields used by
    combined_loss_with_kernel (engine/loss.rs:112-142); the defaults are
    RustGS's.  LossConfig.l1() is the plain L1 loss of the benchmark
    configurations."""

    l1_weight: float = 0.8
    ssim_weight: float = 0.2
    gradient_weight: float = 0.0
    robust_delta: float = 0.0
    outlier_threshold: float = 0.0
    outlier_weight: float = 1.0
    combined: bool = True

    @staticmethod
    def l1() -> "LossConfig":
        return LossConfig(1.0, 0.0, 0.0, 0.0, 0.0, 1.0, combined=False)

    def to_cfg(self) -> L.RgsLossCfg:
        return L.RgsLossCfg(1 if self.combined else 0, self.l1_weight, self.ssim_weight, self.gradient_weight,
                            self.robust_delta, self.outlier_threshold, self.outlier_weight)


def image_loss_grad(ctx: Context, pred, target, weight: float = 1.0, cfg: LossConfig | None = None):
    """combined_loss_with_kernel of pred [H,W,4] against target [H,W,3]:
    (weight * loss, weight * d loss / d pred as an [H,W,4] device tensor)."""
    torch = _torch()
    cfg = cfg or LossConfig()
    h, w = int(pred.shape[0]), int(pred.shape[1])
    v_out = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    loss = c_float()
    c = cfg.to_cfg()
    pred = pred.contiguous()
    target = target.contiguous()
    torch.cuda.synchronize()
    ctx.check(ctx.lib.rgs_image_loss_grad(ctx.h, pred.data_ptr(), target.data_ptr(), w, h, float(weight),
                                          ctypes.byref(c), v_out.data_ptr(), ctypes.byref(loss)))
    return loss.value, v_out


@dataclass
class AdamScaledConfig:
    """engine/optimizer.rs:7-24 (lr 1.0 in the trainer, per-column scaling)."""

    lr: float = 1.0
    betas: tuple = (0.9, 0.999)
    eps: float = 1e-8
    weight_decay: float = 0.0


@dataclass
class AdamScaled:
    config: AdamScaledConfig = field(default_factory=AdamScaledConfig)
    transform_scaling: list = field(default_factory=lambda: [1.0] * 10)
    sh_scaling: list = field(default_factory=lambda: [1.0] * 25)
    opacity_scaling: float = 1.0

    def set_transform_scaling(self, s):
        self.transform_scaling = [float(x) for x in s]

    def set_sh_scaling(self, s):
        self.sh_scaling = [float(x) for x in s] + [0.0] * (25 - len(s))

    def set_opacity_scaling(self, s: float):
        self.opacity_scaling = float(s)

    def to_cfg(self) -> L.RgsAdamCfg:
        c = L.RgsAdamCfg()
        c.beta1, c.beta2, c.eps = float(np.float32(self.config.betas[0])), float(np.float32(self.config.betas[1])), \
            float(np.float32(self.config.eps))
        lr = np.float32(self.config.lr)
        for i in range(10):
            c.lr_transforms[i] = float(np.float32(self.transform_scaling[i]) * lr)
        for i in range(25):
            c.lr_sh[i] = float(np.float32(self.sh_scaling[i]) * lr)
        c.lr_opacity = float(np.float32(self.opacity_scaling) * lr)
        c.weight_decay = float(np.float32(self.config.weight_decay))
        return c

    def step_device_splats(self, ctx: Context, splats: DeviceSplats):
        cfg = self.to_cfg()
        ctx.check(ctx.lib.rgs_adam_step(ctx.h, splats.h, ctypes.byref(cfg)))

    def reset(self, ctx: Context, splats: DeviceSplats):
        ctx.check(ctx.lib.rgs_adam_reset(ctx.h, splats.h))

    @staticmethod
    def steps(ctx: Context, splats: DeviceSplats) -> list:
        """AdamState::step of the transforms, SH and opacity groups."""
        out = (ctypes.c_uint32 * 3)()
        ctx.check(ctx.lib.rgs_adam_get_steps(ctx.h, splats.h, out))
        return list(out)

    @staticmethod
    def set_steps(ctx: Context, splats: DeviceSplats, steps):
        arr = (ctypes.c_uint32 * 3)(*[int(x) for x in steps])
        ctx.check(ctx.lib.rgs_adam_set_steps(ctx.h, splats.h, arr))


def lr_at(initial: float, final_value: float, iteration: int, iterations: int, decay_iterations=None) -> float:
    """engine/trainer.rs:179-191, in float32 like the reference."""
    if final_value <= 0.0 or final_value >= initial or iterations == 0:
        return initial
    d = max(decay_iterations if decay_iterations is not None else iterations, 1)
    t = np.float32(min(iteration, d)) / np.float32(d)
    return float(np.float32(initial) * np.exp(np.float32(np.log(np.float32(final_value) / np.float32(initial))) * t,
                                               dtype=np.float32))


def training_optimizer(sh_degree: int, lr_position: float, lr_rotation: float, lr_scale: float, lr_opacity: float,
                       lr_sh_dc: float, lr_sh_rest: float | None = None) -> AdamScaled:
    """The per-column learning rates of WgpuTrainer::new (trainer.rs:116-151);
    lr_sh_rest, when given, sets the scaling of the non-DC SH coefficients."""
    opt = AdamScaled()
    opt.set_transform_scaling([lr_position] * 3 + [lr_rotation] * 4 + [lr_scale] * 3)
    k = sh_coeffs(sh_degree)
    rest = lr_sh_dc if lr_sh_rest is None else lr_sh_rest
    opt.set_sh_scaling([lr_sh_dc] + [rest] * (k - 1))
    opt.set_opacity_scaling(lr_opacity)
    return opt


class Trainer:
    """Fused training step (WgpuTrainer::train_step) over one or more views."""

    def __init__(self, ctx: Context, optimizer: AdamScaled, cov_blur: float = 0.3, background=(0.0, 0.0, 0.0),
                 loss_weight: float = 1.0, check_loss: bool = True, densify: int = 0, loss: LossConfig | None = None):
        """densify: RGS_DENSIFY_* flags (accumulate_gradients each step, trainer.rs:651-728);
        loss: the loss of each view (default: plain L1)."""
        self.densify = densify
        self.loss = loss or LossConfig.l1()
        self.ctx = ctx
        self.optimizer = optimizer
        self.cov_blur = cov_blur
        self.background = background
        self.loss_weight = loss_weight
        self.check_loss = check_loss

    def step_cfg(self, batched: bool, graph: bool) -> L.RgsStepCfg:
        c = L.RgsStepCfg()
        c.mode = L.RGS_MODE_BATCHED if batched else L.RGS_MODE_PER_VIEW
        c.graph = L.RGS_GRAPH_ON if graph else L.RGS_GRAPH_OFF
        c.cov_blur = float(self.cov_blur)
        for i in range(3):
            c.background[i] = float(self.background[i])
        c.loss_weight = float(self.loss_weight)
        c.check_loss = 1 if self.check_loss else 0
        c.adam = self.optimizer.to_cfg()
        c.densify = int(self.densify)
        c.loss = self.loss.to_cfg()
        return c

    def train_step(self, splats: DeviceSplats, cams, targets, batched: bool = True, graph: bool = False,
                   read_loss: bool = False):
        arr, nv = _cams(cams)
        tgts = _ptrs(targets)
        cfg = self.step_cfg(batched, graph)
        loss = (c_float * nv)() if read_loss else None
        self.ctx.check(self.ctx.lib.rgs_train_step(self.ctx.h, splats.h, arr, tgts, nv, ctypes.byref(cfg), loss))
        return list(loss) if read_loss else None

    def autotune(self, splats: DeviceSplats, cams, targets, batched: bool = True, graph: bool = True):
        arr, nv = _cams(cams)
        cfg = self.step_cfg(batched, graph)
        out = L.RgsRasterCfg()
        self.ctx.check(self.ctx.lib.rgs_autotune(self.ctx.h, splats.h, arr, _ptrs(targets), nv, ctypes.byref(cfg),
                                                 ctypes.byref(out)))
        return out.tile_width, out.batch_depth, bool(out.persistent)


def sort_pairs_u32(ctx: Context, keys, values):
    """radix_sort_by_key_u32_primitive on device tensors (int32 views of u32)."""
    torch = _torch()
    n = int(keys.numel())
    ok = torch.empty_like(keys)
    ov = torch.empty_like(values)
    torch.cuda.synchronize()
    ctx.check(ctx.lib.rgs_sort_pairs_u32(ctx.h, keys.data_ptr() if n else None, values.data_ptr() if n else None,
                                         ok.data_ptr() if n else None, ov.data_ptr() if n else None, n))
    return ok, ov


def inclusive_scan_u32(ctx: Context, x):
    torch = _torch()
    n = int(x.numel())
    out = torch.empty_like(x)
    torch.cuda.synchronize()
    ctx.check(ctx.lib.rgs_inclusive_scan_u32(ctx.h, x.data_ptr() if n else None, out.data_ptr() if n else None, n))
    return out
