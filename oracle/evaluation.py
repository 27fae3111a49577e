"""Oracle for the per-frame evaluation metrics (test infrastructure only: the
product path never imports this module).

Restates RustGS/src/training/evaluation/core.rs in numpy:
  compute_psnr_f32                  core.rs:238-256
  compute_gradient_sharpness_f32    core.rs:258-280
  compute_laplacian_sharpness_f32   core.rs:282-306
  luminance_at                      core.rs:308-311
  sharpness_ratio                   core.rs:313-319
  SplatEvaluationRenderer::render   core.rs:362-377 (u8 colour / 255)

Each metric is a sum over the image of per-pixel terms computed in float32
in the reference's operation order.  The reference accumulates those terms
left to right in float32 (`.sum::<f32>()` and `sum += ...`); `accumulate="f32"`
reproduces that sequence exactly (numpy's float32 cumsum is a sequential
loop).  `accumulate="f64"` sums the same float32 terms in float64 and rounds
each sum once to float32, which is what the library's reduction kernels do:
the two differ only by the float32 rounding of the reference's running sum,
bounded by (terms - 1) * 2^-24 * sum(|term|) (checked in
tests/test_oracle_evaluation.py).
"""
from __future__ import annotations

import numpy as np

F = np.float32


def _sum(terms: np.ndarray, accumulate: str) -> np.float32:
    terms = np.ascontiguousarray(terms, F).ravel()
    if terms.size == 0:
        return F(0.0)
    if accumulate == "f32":
        return F(np.cumsum(terms, dtype=F)[-1])
    return F(np.sum(terms.astype(np.float64)))


def rendered_from_u8(color_u8: np.ndarray) -> np.ndarray:
    """`value as f32 / 255.0` (core.rs:372-376)."""
    return color_u8.astype(F) / F(255.0)


def psnr(rendered: np.ndarray, target: np.ndarray, accumulate: str = "f32") -> np.float32:
    """compute_psnr_f32: 0 for empty or mismatched inputs, 100 dB for an MSE
    at or below 1e-12, else 10 * log10(1 / mse)."""
    r = np.asarray(rendered, F).ravel()
    t = np.asarray(target, F).ravel()
    if r.size != t.size or r.size == 0:
        return F(0.0)
    d = r - t
    mse = _sum(d * d, accumulate) / F(r.size)
    if mse <= F(1e-12):
        return F(100.0)
    return F(10.0) * F(np.log10(F(1.0) / mse))


def luminance(rgb: np.ndarray) -> np.ndarray:
    """0.2126 R + 0.7152 G + 0.0722 B, left to right in float32."""
    rgb = np.asarray(rgb, F)
    return (F(0.2126) * rgb[..., 0] + F(0.7152) * rgb[..., 1]) + F(0.0722) * rgb[..., 2]


def gradient_sharpness(rgb: np.ndarray, accumulate: str = "f32") -> np.float32:
    """Mean absolute luminance difference to the right and lower neighbours;
    the reference visits pixels row by row, right difference first."""
    H, W = rgb.shape[:2]
    if W < 2 or H < 2:
        return F(0.0)
    L = luminance(rgb)
    terms = np.zeros((H, W, 2), F)
    present = np.zeros((H, W, 2), bool)
    terms[:, :-1, 0] = np.abs(L[:, 1:] - L[:, :-1])
    terms[:-1, :, 1] = np.abs(L[1:, :] - L[:-1, :])
    present[:, :-1, 0] = True
    present[:-1, :, 1] = True
    count = (W - 1) * H + W * (H - 1)
    return _sum(terms[present], accumulate) / F(count)


def laplacian_sharpness(rgb: np.ndarray, accumulate: str = "f32") -> np.float32:
    """Variance of the 5-point Laplacian of luminance over the interior:
    lap = -4 c + left + right + up + down (left to right), then
    max(mean(lap^2) - mean(lap)^2, 0)."""
    H, W = rgb.shape[:2]
    if W < 3 or H < 3:
        return F(0.0)
    L = luminance(rgb)
    c = L[1:-1, 1:-1]
    lap = (((F(-4.0) * c + L[1:-1, :-2]) + L[1:-1, 2:]) + L[:-2, 1:-1]) + L[2:, 1:-1]
    count = F((W - 2) * (H - 2))
    mean = _sum(lap, accumulate) / count
    var = _sum(lap * lap, accumulate) / count - mean * mean
    return F(var) if var > F(0.0) else F(0.0)


def sharpness_ratio(rendered: float, target: float) -> np.float32:
    return F(0.0) if F(target) <= F(1e-12) else F(rendered) / F(target)


def frame_metrics(rendered: np.ndarray, target: np.ndarray, accumulate: str = "f32") -> dict:
    """The per-frame numbers of evaluate_splats (core.rs:442-449)."""
    rg, tg = gradient_sharpness(rendered, accumulate), gradient_sharpness(target, accumulate)
    rl, tl = laplacian_sharpness(rendered, accumulate), laplacian_sharpness(target, accumulate)
    return dict(psnr_db=psnr(rendered, target, accumulate), render_grad=rg, target_grad=tg, render_lap=rl,
                target_lap=tl, sharpness_grad_ratio=sharpness_ratio(rg, tg),
                sharpness_lap_ratio=sharpness_ratio(rl, tl))
