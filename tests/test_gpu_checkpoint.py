"""Render checkpoints (RenderCheckpoint per call, backward/autodiff.rs:56-71)
through the C ABI: two forwards into two checkpoints, the context's own
forward in between, and the parameters changed before the backwards; each
checkpoint's backward gives the gradients of its own forward (its camera
and the parameters it read), within 1e-6 rel-L2 of a fresh
forward+backward on those parameters (float atomics reorder sums), the
context's own forward is left alone, and a checkpoint cannot be used twice
or with a store of another size."""
import numpy as np
import pytest

from rustscan_amd import scene
from tests.gpu_util import rel_l2, say

pytestmark = pytest.mark.gpu


def _plain_grads(ctx, sp, cam, v):
    from rustscan_amd.rasterizer import DeviceSplats, render_backward, render_forward

    ds = DeviceSplats(ctx, sp)
    img = render_forward(ctx, ds, [cam])[0]
    g = render_backward(ctx, ds, [v])
    ds.close()
    return img, g


def test_checkpoints_keep_their_forward(hip_ctx):
    import torch

    from rustscan_amd.rasterizer import DeviceSplats, RenderCheckpoint, render_backward, render_forward

    sc = scene.gaussian_scene(20000, seed=300, sh_degree=3)
    sc2 = scene.gaussian_scene(20000, seed=301, sh_degree=3)
    cams = scene.ring_cameras([(320, 240), (257, 191), (200, 150)], seed=302)
    vs = [torch.from_numpy(scene.Rng(310 + i).normal(c.height * c.width * 4).reshape(c.height, c.width, 4)
                           .astype(np.float32)).cuda() for i, c in enumerate(cams)]
    ref = [_plain_grads(hip_ctx, sc, cams[0], vs[0]), _plain_grads(hip_ctx, sc, cams[1], vs[1])]
    ref2 = _plain_grads(hip_ctx, sc2, cams[2], vs[2])
    ds = DeviceSplats(hip_ctx, sc)
    ck = [RenderCheckpoint(hip_ctx), RenderCheckpoint(hip_ctx)]
    say("two checkpointed forwards, new parameters, a plain forward, backwards in reverse order")
    imgs = [render_forward(hip_ctx, ds, [cams[i]], checkpoint=ck[i])[0] for i in range(2)]
    ds.upload(sc2)  # parameters change after the checkpointed forwards
    render_forward(hip_ctx, ds, [cams[2]])  # the context's own forward state
    for i in (1, 0):
        assert torch.equal(imgs[i], ref[i][0])
        g = render_backward(hip_ctx, ds, [vs[i]], checkpoint=ck[i])
        for a, b, name in zip(g[:3], ref[i][1][:3], ("transforms", "sh", "opacity")):
            e = rel_l2(a, b)
            say(f"  view {i} {name}: rel-L2 {e:.2e}")
            assert e <= 1e-6
    # consumed: a second backward is refused; the context's forward is untouched
    with pytest.raises(RuntimeError):
        render_backward(hip_ctx, ds, [vs[0]], checkpoint=ck[0])
    g2 = render_backward(hip_ctx, ds, [vs[2]])
    assert max(rel_l2(a, b) for a, b in zip(g2[:3], ref2[1][:3])) <= 1e-6
    # a store of another size is refused
    small = DeviceSplats(hip_ctx, scene.gaussian_scene(1000, seed=303, sh_degree=3))
    render_forward(hip_ctx, ds, [cams[0]], checkpoint=ck[0])
    with pytest.raises(RuntimeError):
        render_backward(hip_ctx, small, [vs[0]], checkpoint=ck[0])
    for k in ck:
        k.close()
    small.close()
    ds.close()
