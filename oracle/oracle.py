"""ctypes wrapper of the oracle (oracle/_build/librgs_oracle.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and the CPU-baseline
leg of bench.py use this module, as the checker or the timed CPU baseline.
The C++ it wraps (rgs_oracle.cpp) restates RustGS's WGSL pipeline; see the
file:line citations there.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_int64, c_uint32, c_void_p

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "librgs_oracle.so")


class OrcCamera(ctypes.Structure):
    _fields_ = [
        ("viewmat", c_float * 16),
        ("fx", c_float),
        ("fy", c_float),
        ("cx", c_float),
        ("cy", c_float),
        ("width", c_uint32),
        ("height", c_uint32),
        ("cam_pos", c_float * 3),
    ]


class OrcStats(ctypes.Structure):
    _fields_ = [
        ("num_visible", c_int64),
        ("num_intersections", c_int64),
        ("k1_count_mismatch", c_int64),
        ("depth_tie_runs", c_int64),
        ("evaluated_fragments", c_int64),
        ("blended_fragments", c_int64),
        ("num_tiles", c_int64),
    ]


_LIB = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return LIB_PATH


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_forward.restype = c_void_p
        L.orc_forward.argtypes = [c_int64, c_int, c_void_p, c_void_p, c_void_p, POINTER(OrcCamera), c_void_p, c_float,
                                  c_int]
        L.orc_free.argtypes = [c_void_p]
        L.orc_stats.argtypes = [c_void_p, POINTER(OrcStats)]
        for name in ["orc_get_image", "orc_get_visible", "orc_get_depth_order", "orc_get_depths", "orc_get_projected",
                     "orc_get_counts", "orc_get_tile_offsets", "orc_render_bruteforce", "orc_crossing_splats"]:
            getattr(L, name).argtypes = [c_void_p, c_void_p]
        L.orc_get_isect.argtypes = [c_void_p, c_void_p, c_void_p]
        L.orc_backward.restype = c_int
        L.orc_backward.argtypes = [c_void_p, c_void_p, c_int] + [c_void_p] * 6
        L.orc_backward_f64.restype = c_int
        L.orc_backward_f64.argtypes = [c_void_p, c_void_p, c_int, c_void_p] + [c_void_p] * 4
        L.orc_l1_loss_grad.restype = c_double
        L.orc_l1_loss_grad.argtypes = [c_void_p, c_void_p, c_int64, c_int64, c_float, c_void_p]
        L.orc_adam_group.argtypes = [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float,
                                     c_float, c_float, ctypes.c_int32, c_float]
        L.orc_powi.restype = c_float
        L.orc_powi.argtypes = [c_float, ctypes.c_int32]
        L.orc_forward_loss_f64.restype = c_double
        L.orc_forward_loss_f64.argtypes = [c_int64, c_int, c_void_p, c_void_p, c_void_p, POINTER(OrcCamera),
                                           c_void_p, c_double, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                           c_int64, c_void_p, c_void_p]
        L.orc_crossing_map_f64.argtypes = [c_int64, c_int, c_void_p, c_void_p, c_void_p, POINTER(OrcCamera),
                                           c_void_p, c_double, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]
        L.orc_projection_terms_f64.argtypes = [c_int64, c_void_p, c_void_p, POINTER(OrcCamera), c_double, c_void_p,
                                               c_void_p]
        L.orc_tiled_renderer.argtypes = [c_int64, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                         POINTER(OrcCamera), c_void_p, c_float, c_int, c_void_p, c_void_p]
        L.orc_tr_sh_rgb.argtypes = [c_int64, c_int, c_void_p, c_void_p, c_void_p]
        L.orc_grad_f64.restype = c_int
        L.orc_grad_f64.argtypes = [c_int64, c_int, c_void_p, c_void_p, c_void_p, POINTER(OrcCamera), c_void_p,
                                   c_double, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p]
        L.orc_math_exp.argtypes = [c_void_p, c_void_p, c_int64]
        L.orc_expf_libm_check.restype = c_int64
        L.orc_expf_libm_check.argtypes = [c_void_p, c_int]
        L.orc_math_log.argtypes = [c_void_p, c_void_p, c_int64]
        L.orc_math_exp_core.argtypes = [c_void_p, c_void_p, c_int64]
        L.orc_math_exp_core_fast.argtypes = [c_void_p, c_void_p, c_int64]
        L.orc_sh_color.argtypes = [c_int, c_void_p, c_void_p, c_void_p]
        L.orc_sort_by_key.argtypes = [c_int64, c_void_p, c_void_p, c_void_p, c_void_p]
        L.orc_inclusive_scan.argtypes = [c_int64, c_void_p, c_void_p]
        L.orc_num_threads.restype = c_int
        _LIB = L
    return _LIB


def camera(cam) -> OrcCamera:
    c = OrcCamera()
    for i in range(16):
        c.viewmat[i] = float(cam.viewmat[i])
    c.fx, c.fy, c.cx, c.cy = float(cam.fx), float(cam.fy), float(cam.cx), float(cam.cy)
    c.width, c.height = int(cam.width), int(cam.height)
    for i in range(3):
        c.cam_pos[i] = float(cam.cam_pos[i])
    return c


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Render:
    """One FP32 forward of the oracle and its intermediates."""

    def __init__(self, splats, cam, background=(0.0, 0.0, 0.0), cov_blur: float = 0.3, tile_width: int = 16):
        L = lib()
        self.n = splats.n
        self.k = (splats.sh_degree + 1) ** 2
        self.cam = cam
        self._tf, self._sh, self._op = _f32(splats.transforms), _f32(splats.sh), _f32(splats.raw_opac)
        bg = (c_float * 3)(*background)
        self._c = camera(cam)
        self._free = L.orc_free  # held here: module globals may be gone when a late __del__ runs at exit
        self.h = L.orc_forward(self.n, splats.sh_degree, self._tf.ctypes.data, self._sh.ctypes.data,
                               self._op.ctypes.data, ctypes.byref(self._c), bg, float(cov_blur), int(tile_width))
        if not self.h:
            raise ValueError("oracle forward rejected its arguments")
        st = OrcStats()
        L.orc_stats(self.h, ctypes.byref(st))
        self.stats = {f: getattr(st, f) for f, _ in OrcStats._fields_}

    def __del__(self):
        if getattr(self, "h", None):
            self._free(self.h)
            self.h = None

    def image(self) -> np.ndarray:
        out = np.zeros((self.cam.height, self.cam.width, 4), np.float32)
        lib().orc_get_image(self.h, out.ctypes.data)
        return out

    def visible(self) -> np.ndarray:
        out = np.zeros(self.n, np.float32)
        lib().orc_get_visible(self.h, out.ctypes.data)
        return out

    def depth_order(self) -> np.ndarray:
        out = np.zeros(self.stats["num_visible"], np.uint32)
        lib().orc_get_depth_order(self.h, out.ctypes.data)
        return out

    def depths(self) -> np.ndarray:
        out = np.zeros(self.n, np.float32)
        lib().orc_get_depths(self.h, out.ctypes.data)
        return out

    def projected(self) -> np.ndarray:
        """[V,9] in depth order: x, y, a, b, c, r, g, b, opacity."""
        out = np.zeros((self.stats["num_visible"], 9), np.float32)
        lib().orc_get_projected(self.h, out.ctypes.data)
        return out

    def counts(self) -> np.ndarray:
        out = np.zeros(self.n, np.uint32)
        lib().orc_get_counts(self.h, out.ctypes.data)
        return out

    def isect(self):
        m = self.stats["num_intersections"]
        tiles = np.zeros(m, np.uint32)
        gids = np.zeros(m, np.uint32)
        lib().orc_get_isect(self.h, tiles.ctypes.data, gids.ctypes.data)
        return tiles, gids

    def tile_offsets(self) -> np.ndarray:
        out = np.zeros(2 * self.stats["num_tiles"], np.uint32)
        lib().orc_get_tile_offsets(self.h, out.ctypes.data)
        return out

    def bruteforce(self) -> np.ndarray:
        out = np.zeros((self.cam.height, self.cam.width, 4), np.float32)
        lib().orc_render_bruteforce(self.h, out.ctypes.data)
        return out

    def crossing(self) -> np.ndarray:
        out = np.zeros(self.n, np.uint8)
        lib().orc_crossing_splats(self.h, out.ctypes.data)
        return out

    def backward(self, v_out: np.ndarray, mode: int = 0):
        vt = np.zeros((self.n, 10), np.float32)
        vs = np.zeros((self.n, self.k, 3), np.float32)
        vo = np.zeros(self.n, np.float32)
        st = np.zeros((self.n, 7), np.float32)
        V = self.stats["num_visible"]
        vsp = np.zeros((V, 10), np.float32)
        ssp = np.zeros((V, 5), np.float32)
        v_out = _f32(v_out)
        rc = lib().orc_backward(self.h, v_out.ctypes.data, mode, vt.ctypes.data, vs.ctypes.data, vo.ctypes.data,
                                st.ctypes.data, vsp.ctypes.data, ssp.ctypes.data)
        if rc:
            raise ValueError("oracle backward failed")
        return dict(v_transforms=vt, v_sh=vs, v_raw_opac=vo, screen_stats=st, v_splats=vsp, screen_splats=ssp)

    def backward_f64(self, v_out: np.ndarray, mode: int = 0, v_splats=None):
        """The backward replayed in double over this FP32 forward (same lists,
        splats and FP32 skip/stop decisions).  With v_splats ([V,10] FP32
        compact splat gradients) only the projection backward runs in double,
        on those values."""
        vt = np.zeros((self.n, 10), np.float64)
        vs = np.zeros((self.n, self.k, 3), np.float64)
        vo = np.zeros(self.n, np.float64)
        V = self.stats["num_visible"]
        vsp = np.zeros((V, 10), np.float64)
        v_out = _f32(v_out)
        vin = _f32(v_splats) if v_splats is not None else None
        rc = lib().orc_backward_f64(self.h, v_out.ctypes.data, mode, vin.ctypes.data if vin is not None else None,
                                    vt.ctypes.data, vs.ctypes.data, vo.ctypes.data, vsp.ctypes.data)
        if rc:
            raise ValueError("oracle backward failed")
        return dict(v_transforms=vt, v_sh=vs, v_raw_opac=vo, v_splats=vsp)


def l1_loss_grad(pred: np.ndarray, target: np.ndarray, weight: float = 1.0):
    pred, target = _f32(pred), _f32(target)
    h, w = pred.shape[:2]
    v = np.zeros((h, w, 4), np.float32)
    loss = lib().orc_l1_loss_grad(pred.ctypes.data, target.ctypes.data, w, h, float(weight), v.ctypes.data)
    return loss, v


def adam_group(param, grad, m, v, lr_row, beta1, beta2, eps, step, weight_decay=0.0):
    """In-place Adam on float32 arrays of shape [rows, row_len]."""
    for a in (param, grad, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    rows, row_len = param.shape
    lr = _f32(lr_row)
    lib().orc_adam_group(rows, row_len, param.ctypes.data, grad.ctypes.data, m.ctypes.data, v.ctypes.data,
                         lr.ctypes.data, float(beta1), float(beta2), float(eps), int(step), float(weight_decay))


def powi(base: float, e: int) -> float:
    return lib().orc_powi(float(base), int(e))


def forward_loss_f64(splats64, sh_degree, cam, target, cov_blur=0.3, tile_width=16, flags=0, frozen_viewdir=None,
                     frozen_comp=None, background=(0.0, 0.0, 0.0), image_out=None, frozen_crossing=None):
    """FP64 forward + L1 loss (mean over 3HW) for finite differences;
    image_out ([H,W,4] float64, optional) receives the FP64 image.
    frozen_crossing = (gid, pixel mask [H,W] uint8, projection [9]): that
    splat is composited with the given projection at the masked pixels."""
    tf, sh, op = [np.ascontiguousarray(a, np.float64) for a in splats64]
    n = op.shape[0]
    c = camera(cam)
    bg = np.ascontiguousarray(background, np.float64)
    fv = np.ascontiguousarray(frozen_viewdir, np.float64) if frozen_viewdir is not None else None
    fc = np.ascontiguousarray(frozen_comp, np.float64) if frozen_comp is not None else None
    tg = _f32(target)
    if image_out is not None:
        assert image_out.dtype == np.float64 and image_out.flags.c_contiguous
        assert image_out.shape == (cam.height, cam.width, 4)
    fg, fpix, fproj = -1, None, None
    if frozen_crossing is not None:
        fg = int(frozen_crossing[0])
        fpix = np.ascontiguousarray(frozen_crossing[1], np.uint8)
        fproj = np.ascontiguousarray(frozen_crossing[2], np.float64)
        assert fpix.size == cam.height * cam.width and fproj.size == 9
    return lib().orc_forward_loss_f64(n, sh_degree, tf.ctypes.data, sh.ctypes.data, op.ctypes.data, ctypes.byref(c),
                                      bg.ctypes.data, float(cov_blur), int(tile_width), tg.ctypes.data, int(flags),
                                      fv.ctypes.data if fv is not None else None,
                                      fc.ctypes.data if fc is not None else None,
                                      image_out.ctypes.data if image_out is not None else None, fg,
                                      fpix.ctypes.data if fpix is not None else None,
                                      fproj.ctypes.data if fproj is not None else None)


def crossing_map_f64(splats64, sh_degree, cam, cov_blur=0.3, tile_width=16, flags=0, frozen_viewdir=None,
                     frozen_comp=None):
    """(pixel -> gid of the splat that takes T below 1e-4, -1 for none, as [H,W]
    int64; projections [N,9] by gid) of the unperturbed FP64 forward."""
    tf, sh, op = [np.ascontiguousarray(a, np.float64) for a in splats64]
    n = op.shape[0]
    c = camera(cam)
    fv = np.ascontiguousarray(frozen_viewdir, np.float64) if frozen_viewdir is not None else None
    fc = np.ascontiguousarray(frozen_comp, np.float64) if frozen_comp is not None else None
    pix = np.zeros((cam.height, cam.width), np.int64)
    proj = np.zeros((n, 9), np.float64)
    lib().orc_crossing_map_f64(n, sh_degree, tf.ctypes.data, sh.ctypes.data, op.ctypes.data, ctypes.byref(c), None,
                               float(cov_blur), int(tile_width), int(flags), fv.ctypes.data if fv is not None else None,
                               fc.ctypes.data if fc is not None else None, pix.ctypes.data, proj.ctypes.data)
    return pix, proj


def grad_f64(splats64, sh_degree, cam, v_out, mode=0, cov_blur=0.3, tile_width=16):
    """Gradients of the FP64 pipeline taking its own decisions (orc_grad_f64):
    dict with v_transforms [N,10], v_sh [N,K,3], v_raw_opac [N] (float64)."""
    tf, sh, op = (np.ascontiguousarray(a, np.float64) for a in splats64)
    n = tf.shape[0]
    k = sh.shape[1]
    c = camera(cam)
    vo = np.ascontiguousarray(v_out, np.float64)
    vt = np.zeros((n, 10))
    vs = np.zeros((n, k, 3))
    vp = np.zeros(n)
    rc = lib().orc_grad_f64(n, sh_degree, tf.ctypes.data, sh.ctypes.data, op.ctypes.data, ctypes.byref(c), None,
                            float(cov_blur), int(tile_width), vo.ctypes.data, int(mode), vt.ctypes.data,
                            vs.ctypes.data, vp.ctypes.data)
    assert rc == 0
    return {"v_transforms": vt, "v_sh": vs, "v_raw_opac": vp}


def projection_terms_f64(tf64, op64, cam, cov_blur=0.3):
    n = op64.shape[0]
    vd = np.zeros((n, 3), np.float64)
    comp = np.zeros(n, np.float64)
    c = camera(cam)
    tf64 = np.ascontiguousarray(tf64, np.float64)
    op64 = np.ascontiguousarray(op64, np.float64)
    lib().orc_projection_terms_f64(n, tf64.ctypes.data, op64.ctypes.data, ctypes.byref(c), float(cov_blur),
                                   vd.ctypes.data, comp.ctypes.data)
    return vd, comp


def tiled_renderer(splats, cam, background=(0.0, 0.0, 0.0), cov_blur=0.3, gpu_math=False):
    """The reference's CPU renderer (TiledRenderer::render_camera_splats,
    render/tiled_renderer.rs:174-492), single-threaded: (color [H,W,3] in
    [0, 1], depth [H,W] with f32::MAX where no splat reached alpha 0.01).
    gpu_math=True swaps libm's expf for the library's rgs_exp."""
    tf = _f32(splats.transforms)
    pos = np.ascontiguousarray(tf[:, 0:3])
    rot = np.ascontiguousarray(tf[:, 3:7])
    ls = np.ascontiguousarray(tf[:, 7:10])
    op = _f32(splats.raw_opac)
    sh = _f32(splats.sh)
    c = camera(cam)
    color = np.zeros((cam.height, cam.width, 3), np.float32)
    depth = np.zeros((cam.height, cam.width), np.float32)
    bg = (c_float * 3)(*background)
    lib().orc_tiled_renderer(splats.n, splats.sh_degree, pos.ctypes.data, ls.ctypes.data, rot.ctypes.data,
                             op.ctypes.data, sh.ctypes.data, ctypes.byref(c), bg, float(cov_blur), int(gpu_math),
                             color.ctypes.data, depth.ctypes.data)
    return color, depth


def gaussian_renderer(splats, cam, background=(0.0, 0.0, 0.0), cov_blur=0.3, gpu_math=False):
    """GaussianRenderer::render_splats (render/renderer.rs:72-98): the tiled
    render with colour quantised as `(clamp(v, 0, 1) * 255) as u8` (NaN -> 0)
    and depth f32::MAX mapped to 0."""
    color, depth = tiled_renderer(splats, cam, background, cov_blur, gpu_math)
    v = np.clip(color, np.float32(0), np.float32(1)) * np.float32(255.0)
    c8 = np.where(np.isnan(v), 0, np.trunc(v)).astype(np.uint8)
    depth = np.where(depth == np.float32(3.4028235e38), np.float32(0), depth)
    return c8, depth


def tr_sh_rgb(sh, degree, dirs):
    """sh.rs:35-112 evaluate_sh_rgb: sh [N,K,3], dirs [N,3] -> rgb [N,3]."""
    sh = _f32(sh)
    dirs = _f32(dirs)
    out = np.zeros((dirs.shape[0], 3), np.float32)
    lib().orc_tr_sh_rgb(dirs.shape[0], degree, sh.ctypes.data, dirs.ctypes.data, out.ctypes.data)
    return out


def expf_libm_check(max_out: int = 16):
    """rgs_expf_libm (the device's exp for the evaluation render) against the
    C library's expf on every float: (number of differing inputs, the first
    max_out of those inputs as float32)."""
    out = np.zeros(max_out, np.uint32)
    n = lib().orc_expf_libm_check(out.ctypes.data, max_out)
    return int(n), np.sort(out[:min(n, max_out)]).view(np.float32)


def math_exp(x: np.ndarray) -> np.ndarray:
    x = _f32(x)
    y = np.zeros_like(x)
    lib().orc_math_exp(x.ctypes.data, y.ctypes.data, x.size)
    return y


def math_exp_core(x: np.ndarray) -> np.ndarray:
    x = _f32(x)
    y = np.zeros_like(x)
    lib().orc_math_exp_core(x.ctypes.data, y.ctypes.data, x.size)
    return y


def math_exp_core_fast(x: np.ndarray) -> np.ndarray:
    x = _f32(x)
    y = np.zeros_like(x)
    lib().orc_math_exp_core_fast(x.ctypes.data, y.ctypes.data, x.size)
    return y


def math_log(x: np.ndarray) -> np.ndarray:
    x = _f32(x)
    y = np.zeros_like(x)
    lib().orc_math_log(x.ctypes.data, y.ctypes.data, x.size)
    return y


def sh_color(degree: int, coeffs: np.ndarray, viewdir) -> np.ndarray:
    co = _f32(coeffs)
    vd = _f32(viewdir)
    out = np.zeros(3, np.float32)
    lib().orc_sh_color(degree, co.ctypes.data, vd.ctypes.data, out.ctypes.data)
    return out


def sort_by_key(keys, values):
    k = np.ascontiguousarray(keys, np.uint32)
    v = np.ascontiguousarray(values, np.uint32)
    ok = np.zeros_like(k)
    ov = np.zeros_like(v)
    lib().orc_sort_by_key(k.size, k.ctypes.data, v.ctypes.data, ok.ctypes.data, ov.ctypes.data)
    return ok, ov


def inclusive_scan(x):
    a = np.ascontiguousarray(x, np.uint32)
    out = np.zeros_like(a)
    lib().orc_inclusive_scan(a.size, a.ctypes.data, out.ctypes.data)
    return out


def num_threads() -> int:
    return lib().orc_num_threads()


DENSIFY_FIELDS = ["grad_2d", "screen_grad_2d", "abs_grad_2d", "abs_pixel_grad_2d", "pixel_coverage", "camera_depth",
                  "grad_color", "num_observations", "visible_observations", "actual_visible_observations"]


class DensifyAccumulators:
    """The trainer's densification accumulators and their update, restated
    from RustGS/src/training/engine/trainer.rs:66-75 (state), :651-728
    (accumulate_gradients) and :861-871 (reset_accumulators), in float32.

    accumulate() takes one view's gradients: v_transforms [N,10] (the
    post-projection transform gradient), screen_stats [N,7], v_sh [N,K,3]
    and the forward's visible [N] (0/1)."""

    def __init__(self, n: int):
        self.n = n
        self.reset()

    def reset(self):
        self.acc = {k: np.zeros(self.n, np.float32) for k in DENSIFY_FIELDS}

    def accumulate(self, v_transforms, screen_stats, v_sh, visible, use_actual_visibility=False,
                   collect_actual_visibility=False):
        f = np.float32
        vt = np.asarray(v_transforms, f)
        st = np.asarray(screen_stats, f)
        vs = np.asarray(v_sh, f)
        vis = np.asarray(visible, f)
        a = self.acc
        # trainer.rs:663-668 slice(0..3).abs().mean_dim(1)
        a["grad_2d"] += np.abs(vt[:, 0:3]).mean(axis=1, dtype=f)
        # :669-675 and :676-682 sqrt(sum of squares) of stats 0..2 and 2..4
        a["screen_grad_2d"] += np.sqrt((st[:, 0:2] ** 2).sum(axis=1, dtype=f))
        a["abs_grad_2d"] += np.sqrt((st[:, 2:4] ** 2).sum(axis=1, dtype=f))
        # :683-694 columns 4, 5, 6 as they are
        a["abs_pixel_grad_2d"] += st[:, 4]
        a["pixel_coverage"] += st[:, 5]
        a["camera_depth"] += st[:, 6]
        # :695-701 abs().mean_dim(2).mean_dim(1): over channels, then coefficients
        a["grad_color"] += np.abs(vs).mean(axis=2, dtype=f).mean(axis=1, dtype=f)
        # :717-727
        a["num_observations"] += f(1.0)
        if collect_actual_visibility:
            a["actual_visible_observations"] += vis
        a["visible_observations"] += vis if use_actual_visibility else f(1.0)


# ---- topology step (oracle/topology.cpp) ------------------------------------
class OrcTopologyCfg(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_uint32),
        ("max_initial_gaussians", ctypes.c_uint32),
        ("frame_shuffle_seed", ctypes.c_uint64),
        ("refine_every", ctypes.c_uint32),
        ("opacity_reset_interval", ctypes.c_uint32),
        ("opacity_reset_mode", ctypes.c_uint32),
        ("topology_freeze_after_epoch", ctypes.c_int32),
        ("growth_freeze_after_epoch", ctypes.c_int32),
        ("prune_until_epoch", ctypes.c_int32),
        ("split_score_mode", ctypes.c_uint32),
        ("growth_stop_iter", ctypes.c_uint32),
        ("growth_grad_threshold", c_float),
        ("split_grad_threshold", c_float),
        ("depth_scale_gamma", c_float),
        ("growth_select_fraction", c_float),
        ("opacity_decay", c_float),
        ("scale_decay", c_float),
        ("prune_mode", ctypes.c_uint32),
        ("prune_min_age", ctypes.c_uint32),
        ("prune_invisible_epochs", ctypes.c_uint32),
        ("prune_visibility_dry_run", ctypes.c_uint32),
        ("prune_opacity_threshold", c_float),
        ("prune_visibility_threshold", c_float),
        ("prune_high_opacity_threshold", c_float),
    ]


# TrainingConfig defaults the topology step reads (config.rs:262-383, 666-686, 712-721, 790-809)
TOPOLOGY_DEFAULTS = dict(
    iterations=30000, max_initial_gaussians=100000, frame_shuffle_seed=0, refine_every=160,
    opacity_reset_interval=10, opacity_reset_mode=0, topology_freeze_after_epoch=-1, growth_freeze_after_epoch=-1,
    prune_until_epoch=-1, split_score_mode=0, growth_stop_iter=15000, growth_grad_threshold=0.00014,
    split_grad_threshold=0.00014, depth_scale_gamma=0.37, growth_select_fraction=0.25, opacity_decay=0.0,
    scale_decay=0.0, prune_mode=1, prune_min_age=5, prune_invisible_epochs=10, prune_visibility_dry_run=0,
    prune_opacity_threshold=1.0 / 255.0, prune_visibility_threshold=0.05, prune_high_opacity_threshold=0.8)


def topology_cfg(**overrides) -> OrcTopologyCfg:
    d = dict(TOPOLOGY_DEFAULTS)
    d.update(overrides)
    return OrcTopologyCfg(**d)


class OrcTopologySchedule(ctypes.Structure):
    _fields_ = [("completed_epoch", ctypes.c_int32), ("densify", ctypes.c_int32), ("prune", ctypes.c_int32),
                ("reset_opacity", ctypes.c_int32), ("allow_extra_growth", ctypes.c_int32)]


class OrcTopologyOut(ctypes.Structure):
    _fields_ = [
        ("schedule", OrcTopologySchedule),
        ("before", c_int64), ("after", c_int64), ("added", c_int64), ("pruned", c_int64),
        ("clone_candidates", c_int64), ("split_candidates", c_int64), ("growth_candidates", c_int64),
        ("prune_candidates", c_int64), ("replacement_count", c_int64), ("extra_growth_count", c_int64),
        ("selected_count", c_int64), ("mutated", ctypes.c_int32), ("adam_reset", ctypes.c_int32),
        ("center", c_float * 3), ("scene_extent", c_float), ("brush_extent", c_float),
        ("selected", c_int64 * 4096),
        ("transforms", POINTER(c_float)), ("sh", POINTER(c_float)), ("raw_opac", POINTER(c_float)),
        ("birth", POINTER(ctypes.c_uint32)), ("invisible_windows", POINTER(ctypes.c_uint32)),
        ("row_src", POINTER(c_int64)), ("row_kind", POINTER(ctypes.c_int)), ("flags", POINTER(ctypes.c_uint8)),
    ]


def _topology_lib():
    L = lib()
    if not getattr(L, "_topology_bound", False):
        L.orc_topology_schedule.argtypes = [POINTER(OrcTopologyCfg), ctypes.c_uint32, ctypes.c_uint32,
                                            POINTER(OrcTopologySchedule)]
        L.orc_topology_seed.restype = ctypes.c_uint64
        L.orc_topology_seed.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
        L.orc_topology_refine_scalar.restype = c_float
        L.orc_topology_refine_scalar.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.orc_topology_key.restype = c_float
        L.orc_topology_key.argtypes = [ctypes.c_uint64, c_int, ctypes.c_uint64, c_float]
        L.orc_topology_step.restype = c_int
        L.orc_topology_step.argtypes = [POINTER(OrcTopologyCfg), c_int64, c_int] + [c_void_p] * 6 + \
            [ctypes.c_uint32, ctypes.c_uint32, POINTER(OrcTopologyOut)]
        L.orc_topology_free.argtypes = [POINTER(OrcTopologyOut)]
        L.orc_brush_refine_offset.argtypes = [c_void_p, c_void_p, c_float, c_void_p]
        L.orc_cos_2pi.restype = c_float
        L.orc_cos_2pi.argtypes = [c_float]
        L._topology_bound = True
    return L


def topology_schedule(cfg: OrcTopologyCfg, iteration: int, frame_count: int) -> dict:
    out = OrcTopologySchedule()
    _topology_lib().orc_topology_schedule(ctypes.byref(cfg), iteration, frame_count, ctypes.byref(out))
    return {f: getattr(out, f) for f, _ in OrcTopologySchedule._fields_}


def brush_refine_offset(rot_wxyz, scale, sample_scalar) -> np.ndarray:
    r, s = _f32(rot_wxyz), _f32(scale)
    out = np.zeros(3, np.float32)
    _topology_lib().orc_brush_refine_offset(r.ctypes.data, s.ctypes.data, float(sample_scalar), out.ctypes.data)
    return out


def topology_step(cfg: OrcTopologyCfg, splats, dens, birth, invis, iteration: int, frame_count: int) -> dict:
    """One topology step of the trainer (apply_topology_mutations,
    engine/trainer.rs:792-859) on host arrays: dens [10,N] accumulators,
    birth and invisible-window state [N].  Returns the counts, flags,
    selection, row sources/kinds, the rebuilt splats and state."""
    L = _topology_lib()
    n = splats.n
    tf, sh, op = _f32(splats.transforms), _f32(splats.sh), _f32(splats.raw_opac)
    d = _f32(dens).reshape(10, n) if n else np.zeros((10, 0), np.float32)
    b = np.ascontiguousarray(birth, np.uint32)
    w = np.ascontiguousarray(invis, np.uint32)
    out = OrcTopologyOut()
    rc = L.orc_topology_step(ctypes.byref(cfg), n, splats.sh_degree, tf.ctypes.data, sh.ctypes.data, op.ctypes.data,
                             d.ctypes.data, b.ctypes.data, w.ctypes.data, iteration, frame_count, ctypes.byref(out))
    if rc:
        raise ValueError("oracle topology step rejected its arguments")
    m = out.after
    k = sh.shape[1] if sh.ndim == 3 else (splats.sh_degree + 1) ** 2
    res = {f: getattr(out, f) for f, _ in OrcTopologyOut._fields_
           if f not in ("schedule", "selected", "center", "transforms", "sh", "raw_opac", "birth",
                        "invisible_windows", "row_src", "row_kind", "flags")}
    res["schedule"] = {f: getattr(out.schedule, f) for f, _ in OrcTopologySchedule._fields_}
    res["center"] = np.array(out.center[:], np.float32)
    res["selected"] = np.array(out.selected[:min(out.selected_count, 4096)], np.int64)
    res["transforms"] = np.ctypeslib.as_array(out.transforms, (max(m, 1) * 10,))[:m * 10].reshape(m, 10).copy()
    res["sh"] = np.ctypeslib.as_array(out.sh, (max(m, 1) * k * 3,))[:m * k * 3].reshape(m, k, 3).copy()
    res["raw_opac"] = np.ctypeslib.as_array(out.raw_opac, (max(m, 1),))[:m].copy()
    res["birth"] = np.ctypeslib.as_array(out.birth, (max(m, 1),))[:m].copy()
    res["invisible_windows"] = np.ctypeslib.as_array(out.invisible_windows, (max(m, 1),))[:m].copy()
    res["row_src"] = np.ctypeslib.as_array(out.row_src, (max(m, 1),))[:m].copy()
    res["row_kind"] = np.ctypeslib.as_array(out.row_kind, (max(m, 1),))[:m].copy()
    res["flags"] = np.ctypeslib.as_array(out.flags, (max(n, 1),))[:n].copy()
    L.orc_topology_free(ctypes.byref(out))
    return res
