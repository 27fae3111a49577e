"""Scene files through the library (librgs_hip.so, host code, no device
needed): RustGS's splat PLY and .splat formats (io/scene_io.rs).

  save_splats(path, splats, metadata)     save_splats, scene_io.rs:617-627
  load_splats(path) -> (Splats, meta)      load_splats, scene_io.rs:872-878
  save_device_splats(ctx, ds, path, meta)  the device store's parameters
  load_device_splats(ctx, path)            host_splats_to_device of a file

`Splats` is the package's host container (transforms [N,10] = mean, wxyz
quaternion, log-scale; sh [N,K,3]; raw opacity logits [N]); the files hold the
HostSplats fields (positions, log_scales, rotations, opacity_logits,
sh_coeffs), which are the same numbers.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_int, c_int64, c_void_p
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .scene import Splats, sh_coeffs


@dataclass
class SplatMetadata:
    """scene_io.rs:12-19."""

    iterations: int = 0
    final_loss: float = 0.0
    gaussian_count: int = 0
    sh_degree: int = 0

    def to_c(self) -> L.RgsSplatMetadata:
        return L.RgsSplatMetadata(int(self.iterations), float(self.final_loss), int(self.gaussian_count),
                                  int(self.sh_degree))

    @staticmethod
    def from_c(m: L.RgsSplatMetadata) -> "SplatMetadata":
        return SplatMetadata(int(m.iterations), float(m.final_loss), int(m.gaussian_count), int(m.sh_degree))


class SceneIoError(RuntimeError):
    pass


def _check(rc: int, ctx=None):
    if rc != 0:
        msg = ctx.lib.rgs_last_error(ctx.h) if ctx is not None else L.lib().rgs_io_last_error()
        raise SceneIoError(msg.decode() if msg else f"scene I/O error {rc}")


def _path(path) -> bytes:
    return os.fsencode(os.fspath(path))


def save_splats(path, splats: Splats, metadata: SplatMetadata | None = None) -> None:
    tf = np.ascontiguousarray(splats.transforms, np.float32)
    pos = np.ascontiguousarray(tf[:, 0:3])
    rot = np.ascontiguousarray(tf[:, 3:7])
    ls = np.ascontiguousarray(tf[:, 7:10])
    op = np.ascontiguousarray(splats.raw_opac, np.float32)
    sh = np.ascontiguousarray(splats.sh, np.float32)
    meta = metadata or SplatMetadata(0, 0.0, splats.n, splats.sh_degree)
    m = meta.to_c()
    _check(L.lib().rgs_save_splats(None, _path(path), splats.n, splats.sh_degree, pos.ctypes.data, ls.ctypes.data,
                                   rot.ctypes.data, op.ctypes.data, sh.ctypes.data, ctypes.byref(m)))


def load_splats(path) -> tuple:
    lib = L.lib()
    h = c_void_p()
    _check(lib.rgs_scene_load(None, _path(path), ctypes.byref(h)))
    try:
        n, deg, m = c_int64(), c_int(), L.RgsSplatMetadata()
        _check(lib.rgs_scene_info(h, ctypes.byref(n), ctypes.byref(deg), ctypes.byref(m)))
        n, deg = n.value, deg.value
        k = sh_coeffs(deg)
        pos = np.zeros((n, 3), np.float32)
        ls = np.zeros((n, 3), np.float32)
        rot = np.zeros((n, 4), np.float32)
        op = np.zeros(n, np.float32)
        sh = np.zeros((n, k, 3), np.float32)
        _check(lib.rgs_scene_get(h, pos.ctypes.data, ls.ctypes.data, rot.ctypes.data, op.ctypes.data,
                                 sh.ctypes.data))
    finally:
        lib.rgs_scene_free(h)
    tf = np.concatenate([pos, rot, ls], axis=1).astype(np.float32)
    return Splats(tf, sh, op, deg), SplatMetadata.from_c(m)


def save_device_splats(ctx, splats, path, metadata: SplatMetadata | None = None) -> None:
    m = (metadata or SplatMetadata(0, 0.0, splats.n, splats.sh_degree)).to_c()
    _check(ctx.lib.rgs_splats_save(ctx.h, splats.h, _path(path), ctypes.byref(m)), ctx)


def load_device_splats(ctx, path):
    """(DeviceSplats, SplatMetadata) from a file, uploaded by the library."""
    from .rasterizer import DeviceSplats

    lib = ctx.lib
    h = c_void_p()
    _check(lib.rgs_scene_load(ctx.h, _path(path), ctypes.byref(h)), ctx)
    try:
        m = L.RgsSplatMetadata()
        _check(lib.rgs_scene_info(h, None, None, ctypes.byref(m)), ctx)
        s = c_void_p()
        _check(lib.rgs_scene_to_splats(ctx.h, h, ctypes.byref(s)), ctx)
    finally:
        lib.rgs_scene_free(h)
    return DeviceSplats.adopt(ctx, s), SplatMetadata.from_c(m)
