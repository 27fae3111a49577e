"""Oracle for the training loss (test infrastructure only: the product path
never imports this module).

Restates RustGS/src/training/engine/loss.rs in numpy:
  SsimConfig defaults                 loss.rs:8-27
  ssim_loss_with_kernel               loss.rs:42-76
  combined_loss_with_kernel           loss.rs:112-142
  reconstruction_residual_loss        loss.rs:144-196 (L1, saturating robust,
                                      soft outlier; the dynamic mask is not
                                      restated: it is off in the configurations)
  gradient_difference_loss            loss.rs:213-227
  gaussian_kernel_1d                  loss.rs:233-253
  separable_blur (edge padding)       loss.rs:255-276

forward(...) evaluates in float32 with the reference's operation order, or in
float64 (for finite differences).  grad(...) is the analytic gradient of the
loss with respect to the predicted RGB image, in float64: the reference
obtains it from Burn's autodiff of the same expression.  It is pinned in
tests/test_oracle_loss.py by float64 central differences of forward(), by
closed forms (constant images, identical images) and by the reference's own
loss tests (loss_tests.rs).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class LossConfig:
    """TrainingLossConfig fields the loss uses (config.rs:566-608 defaults
    are l1 0.8, ssim 0.2, gradient 0, robust 0, outlier off)."""

    l1_weight: float = 0.8
    ssim_weight: float = 0.2
    gradient_weight: float = 0.0
    robust_delta: float = 0.0
    outlier_threshold: float = 0.0
    outlier_weight: float = 1.0
    window_size: int = 11
    sigma: float = 1.5
    k1: float = 0.01
    k2: float = 0.03
    data_range: float = 1.0


def gaussian_kernel_1d(window_size: int = 11, sigma: float = 1.5) -> np.ndarray:
    """loss.rs:233-253: exp in f64 cast to f32, sum and division in f32."""
    radius = window_size // 2
    vals = []
    s = np.float32(0.0)
    for off in range(-radius, radius + 1):
        v = np.float32(np.exp(-float(off * off) / (2.0 * sigma * sigma)))
        vals.append(v)
        s = np.float32(s + v)
    den = np.float32(max(s, np.float32(1e-8)))
    return np.array([np.float32(v / den) for v in vals], np.float32)


def separable_blur(x: np.ndarray, k: np.ndarray) -> np.ndarray:
    """loss.rs:255-276 on [..., H, W]: horizontal taps over an edge-padded
    row, then vertical taps over the edge-padded result, accumulated tap by
    tap in the dtype of x (the order of loss_tests.rs:13-59)."""
    dt = x.dtype
    k = k.astype(dt)
    pad = len(k) // 2
    H, W = x.shape[-2], x.shape[-1]
    xp = np.pad(x, [(0, 0)] * (x.ndim - 1) + [(pad, pad)], mode="edge")
    h = xp[..., 0:W] * k[0]
    for i in range(1, len(k)):
        h = h + xp[..., i:i + W] * k[i]
    hp = np.pad(h, [(0, 0)] * (x.ndim - 2) + [(pad, pad), (0, 0)], mode="edge")
    v = hp[..., 0:H, :] * k[0]
    for i in range(1, len(k)):
        v = v + hp[..., i:i + H, :] * k[i]
    return v


def blur_transpose(g: np.ndarray, k: np.ndarray) -> np.ndarray:
    """Adjoint of separable_blur (float64): <blur(x), g> = <x, blur_transpose(g)>."""
    pad = len(k) // 2
    H, W = g.shape[-2], g.shape[-1]
    k = k.astype(np.float64)
    # adjoint of the vertical pass: scatter back onto the padded rows, fold the pads onto the edges
    hp = np.zeros(g.shape[:-2] + (H + 2 * pad, W), np.float64)
    for i in range(len(k)):
        hp[..., i:i + H, :] += g * k[i]
    h = hp[..., pad:pad + H, :].copy()
    h[..., 0, :] += hp[..., 0:pad, :].sum(axis=-2)
    h[..., H - 1, :] += hp[..., pad + H:, :].sum(axis=-2)
    xp = np.zeros(g.shape[:-1] + (W + 2 * pad,), np.float64)
    for i in range(len(k)):
        xp[..., i:i + W] += h * k[i]
    x = xp[..., pad:pad + W].copy()
    x[..., 0] += xp[..., 0:pad].sum(axis=-1)
    x[..., W - 1] += xp[..., pad + W:].sum(axis=-1)
    return x


def _chw(img):
    return np.moveaxis(np.asarray(img)[..., 0:3], -1, 0)


def ssim_map(pred_chw, tgt_chw, cfg: LossConfig, dtype=np.float32):
    """The SSIM map of loss.rs:48-72 and its intermediates."""
    f = dtype
    x = pred_chw.astype(f)
    y = tgt_chw.astype(f)
    k = gaussian_kernel_1d(cfg.window_size, cfg.sigma).astype(f)
    mu_x = separable_blur(x, k)
    mu_y = separable_blur(y, k)
    mu_x_sq = mu_x * mu_x
    mu_y_sq = mu_y * mu_y
    mu_xy = mu_x * mu_y
    sxx = separable_blur(x * x, k) - mu_x_sq
    syy = separable_blur(y * y, k) - mu_y_sq
    sxy = separable_blur(x * y, k) - mu_xy
    c1 = f((cfg.k1 * cfg.data_range) ** 2)
    c2 = f((cfg.k2 * cfg.data_range) ** 2)
    a1 = mu_xy * f(2.0) + c1
    a2 = sxy * f(2.0) + c2
    b1 = np.maximum((mu_x_sq + mu_y_sq) + c1, f(1e-8))
    b2 = np.maximum((sxx + syy) + c2, f(1e-8))
    s = (a1 * a2) / (b1 * b2)
    return dict(s=s, mu_x=mu_x, mu_y=mu_y, a1=a1, a2=a2, b1=b1, b2=b2,
                b1_raw=(mu_x_sq + mu_y_sq) + c1, b2_raw=(sxx + syy) + c2, k=k)


def _residual_terms(d, cfg: LossConfig, dtype):
    """Per-element residual loss r(|d|) of loss.rs:153-176 and dr/dd."""
    f = dtype
    a = np.abs(d)
    sgn = np.sign(d)
    if np.isfinite(cfg.robust_delta) and cfg.robust_delta > 0.0:
        delta = f(max(cfg.robust_delta, 1e-6))
        r = (a * delta) / (a + delta)
        dr = sgn * (delta * delta) / ((a + delta) * (a + delta))
    elif (np.isfinite(cfg.outlier_threshold) and cfg.outlier_threshold > 0.0 and np.isfinite(cfg.outlier_weight)
          and cfg.outlier_weight < 1.0):
        thr = f(max(cfg.outlier_threshold, 1e-6))
        floor = f(min(max(cfg.outlier_weight, 0.0), 1.0))
        w = (a * f(0.0) + floor) + ((a * f(0.0) + (f(1.0) - floor)) * thr) / (a + thr)
        r = a * w
        dr = sgn * (floor + (f(1.0) - floor) * thr * thr / ((a + thr) * (a + thr)))
    else:
        r = a
        dr = sgn
    return r, dr


def residual_loss(pred, target, cfg: LossConfig, dtype=np.float32) -> float:
    """reconstruction_residual_loss (loss.rs:144-196) without the dynamic mask."""
    p = np.asarray(pred)[..., 0:3].astype(dtype)
    t = np.asarray(target)[..., 0:3].astype(dtype)
    r, _ = _residual_terms(p - t, cfg, dtype)
    return float(r.mean(dtype=dtype))


def forward(pred, target, cfg: LossConfig, dtype=np.float32) -> float:
    """combined_loss_with_kernel (loss.rs:112-142) of pred [H,W,>=3] against
    target [H,W,3]: l1 * residual + ssim * (1 - mean SSIM) + gradient * GDL.
    The image-gradient term is evaluated whatever its weight, as in the
    reference: on an image one pixel wide or high one of its means is over
    an empty difference (NaN), and the loss is NaN."""
    f = dtype
    p = np.asarray(pred)[..., 0:3].astype(f)
    t = np.asarray(target)[..., 0:3].astype(f)
    r, _ = _residual_terms(p - t, cfg, f)
    l1 = r.mean(dtype=f)
    sm = ssim_map(_chw(p), _chw(t), cfg, f)["s"]
    ssim = f(1.0) - sm.mean(dtype=f)
    dxp = p[:, 1:, :] - p[:, :-1, :]
    dxt = t[:, 1:, :] - t[:, :-1, :]
    dyp = p[1:, :, :] - p[:-1, :, :]
    dyt = t[1:, :, :] - t[:-1, :, :]
    gx = np.abs(dxp - dxt).mean(dtype=f) if p.shape[1] > 1 else f(np.nan)
    gy = np.abs(dyp - dyt).mean(dtype=f) if p.shape[0] > 1 else f(np.nan)
    gdl = (gx + gy) * f(0.5)
    out = (l1 * f(cfg.l1_weight) + ssim * f(cfg.ssim_weight)) + gdl * f(cfg.gradient_weight)
    return float(out)


def grad(pred, target, cfg: LossConfig) -> np.ndarray:
    """d forward / d pred_rgb in float64, shape [H,W,3]."""
    p = np.asarray(pred)[..., 0:3].astype(np.float64)
    t = np.asarray(target)[..., 0:3].astype(np.float64)
    H, W, _ = p.shape
    n = p.size
    _, dr = _residual_terms(p - t, cfg, np.float64)
    g = dr * (cfg.l1_weight / n)
    if cfg.ssim_weight != 0.0:
        x, y = _chw(p), _chw(t)
        m = ssim_map(x, y, cfg, np.float64)
        s, mu_x, mu_y, a1, a2, b1, b2 = m["s"], m["mu_x"], m["mu_y"], m["a1"], m["a2"], m["b1"], m["b2"]
        on1 = m["b1_raw"] > 1e-8  # clamp_min passes the gradient only where it does not clamp
        on2 = m["b2_raw"] > 1e-8
        # s = a1 a2 / (b1 b2) with a1 = 2 mx my + c1, a2 = 2 (Exy - mx my) + c2,
        # b1 = mx^2 + my^2 + c1, b2 = Exx - mx^2 + Eyy - my^2 + c2
        d_b1 = np.where(on1, -s / b1, 0.0)
        d_b2 = np.where(on2, -s / b2, 0.0)
        d_a1 = a2 / (b1 * b2)
        d_a2 = a1 / (b1 * b2)
        g_mu = d_a1 * 2.0 * mu_y + d_a2 * (-2.0 * mu_y) + d_b1 * 2.0 * mu_x + d_b2 * (-2.0 * mu_x)
        g_xx = d_b2
        g_xy = d_a2 * 2.0
        scale = -cfg.ssim_weight / s.size
        k = m["k"]
        gs = (blur_transpose(g_mu * scale, k) + 2.0 * x * blur_transpose(g_xx * scale, k)
              + y * blur_transpose(g_xy * scale, k))
        g = g + np.moveaxis(gs, 0, -1)
    if cfg.gradient_weight != 0.0:
        gw = 0.5 * cfg.gradient_weight
        if W > 1:
            sx = np.sign((p[:, 1:, :] - p[:, :-1, :]) - (t[:, 1:, :] - t[:, :-1, :])) * (gw / (H * (W - 1) * 3))
            g[:, 1:, :] += sx
            g[:, :-1, :] -= sx
        if H > 1:
            sy = np.sign((p[1:, :, :] - p[:-1, :, :]) - (t[1:, :, :] - t[:-1, :, :])) * (gw / ((H - 1) * W * 3))
            g[1:, :, :] += sy
            g[:-1, :, :] -= sy
    return g
