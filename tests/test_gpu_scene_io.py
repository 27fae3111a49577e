"""Scene files to and from a device store (rgs_scene_to_splats,
rgs_splats_save): the store holds the file's parameters bit for bit, and
saving it writes the bytes the host writer gives for those parameters (for
PLY, the original file itself; .splat quantises, so a loaded and re-saved
.splat need not equal its source)."""
import numpy as np
import pytest

from rustscan_amd import scene
from tests.gpu_util import say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ext", ["ply", "splat"])
def test_device_store_round_trip(hip_ctx, tmp_path, ext):
    from rustscan_amd.scene_io import SplatMetadata, load_device_splats, load_splats, save_device_splats, save_splats

    sp = scene.gaussian_scene(5000, seed=90, sh_degree=3 if ext == "ply" else 0)
    src = tmp_path / f"src.{ext}"
    meta = SplatMetadata(42, 0.03125, sp.n, sp.sh_degree)
    save_splats(src, sp, meta)
    say(f"device round trip .{ext}")
    ds, m = load_device_splats(hip_ctx, src)
    host, _ = load_splats(src)
    back = ds.download()
    for a, b in ((back.transforms, host.transforms), (back.sh, host.sh), (back.raw_opac, host.raw_opac)):
        assert np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))
    dst = tmp_path / f"dst.{ext}"
    save_device_splats(hip_ctx, ds, dst, m)
    ds.close()
    ref = tmp_path / f"ref.{ext}"
    save_splats(ref, host, m)
    assert dst.read_bytes() == ref.read_bytes()
    if ext == "ply":
        assert dst.read_bytes() == src.read_bytes()
