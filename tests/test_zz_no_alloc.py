"""rgs_sort_pairs_u32_noalloc, the sort entry point that allocates nothing
(its scratch is a fixed device array per GPU), against rgs_sort_pairs_u32
on the same inputs: bit-exact, on two calls in a row where the second is
smaller (the scratch keeps the first call's contents), in place, and with
sizes beyond its maximum rejected."""
import ctypes

import numpy as np
import pytest

from tests.gpu_util import from_dev_u32, say, to_dev_u32

pytestmark = pytest.mark.gpu


def noalloc_sort(ctx, keys, values, out_keys=None, out_values=None):
    import torch

    n = int(keys.numel())
    ok = torch.empty_like(keys) if out_keys is None else out_keys
    ov = torch.empty_like(values) if out_values is None else out_values
    torch.cuda.synchronize()
    rc = ctx.lib.rgs_sort_pairs_u32_noalloc(ctx.h, keys.data_ptr(), values.data_ptr(), ok.data_ptr(), ov.data_ptr(),
                                            n)
    ctx.check(rc)
    ctx.check(ctx.lib.rgs_sync(ctx.h))
    return ok, ov


def test_noalloc_sort_matches_the_allocating_sort(hip_ctx):
    import torch

    from rustscan_amd._lib import RGS_SORT_NOALLOC_MAX_PAIRS
    from rustscan_amd.rasterizer import sort_pairs_u32

    cases = [(RGS_SORT_NOALLOC_MAX_PAIRS, 1 << 32), (300_001, 4096), (1, 7)]
    for n, hi in cases:
        say(f"no-alloc sort n={n} keys below {hi}")
        rng = np.random.default_rng(n)
        keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
        vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        k, v = to_dev_u32(torch, keys), to_dev_u32(torch, vals)
        ref_k, ref_v = sort_pairs_u32(hip_ctx, k, v)
        ok, ov = noalloc_sort(hip_ctx, k, v)
        np.testing.assert_array_equal(from_dev_u32(ok), from_dev_u32(ref_k))
        np.testing.assert_array_equal(from_dev_u32(ov), from_dev_u32(ref_v))
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(from_dev_u32(ov), vals[order])
        # the inputs are left unchanged
        np.testing.assert_array_equal(from_dev_u32(k), keys)


def test_noalloc_sort_in_place_and_limits(hip_ctx):
    import torch

    from rustscan_amd._lib import RGS_ERR_INVALID, RGS_SORT_NOALLOC_MAX_PAIRS

    say("no-alloc sort in place, n=70000")
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 1 << 20, 70_000, dtype=np.uint64).astype(np.uint32)
    k, v = to_dev_u32(torch, keys), to_dev_u32(torch, np.arange(70_000))
    noalloc_sort(hip_ctx, k, v, out_keys=k, out_values=v)
    order = np.argsort(keys, kind="stable").astype(np.uint32)
    np.testing.assert_array_equal(from_dev_u32(v), order)
    np.testing.assert_array_equal(from_dev_u32(k), keys[order])
    # beyond the maximum: rejected on the host before any device work
    too_many = RGS_SORT_NOALLOC_MAX_PAIRS + 1
    p = ctypes.c_void_p(k.data_ptr())
    rc = hip_ctx.lib.rgs_sort_pairs_u32_noalloc(hip_ctx.h, p, p, p, p, too_many)
    assert rc == RGS_ERR_INVALID
    assert "exceeds" in hip_ctx.lib.rgs_last_error(hip_ctx.h).decode()
    assert hip_ctx.lib.rgs_sort_pairs_u32_noalloc(hip_ctx.h, None, None, None, None, 0) == 0
