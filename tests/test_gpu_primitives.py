"""Device radix sort and scan through the C ABI (rgs_sort_pairs_u32,
rgs_inclusive_scan_u32) against the reference's known-answer vectors
(RustGS/src/training/gpu_primitives/radix_sort/tests.rs:89-121,
prefix_sum/tests.rs:41-59) and, at sizes beyond one workgroup, against
numpy's stable argsort / cumsum: bit-exact."""
import numpy as np
import pytest

from tests.gpu_util import from_dev_u32, say, to_dev_u32

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    return t


ARGSORT_CASES = [
    ([42], [0]),
    ([9, 4, 1, 7, 4, 3, 2], [2, 6, 5, 1, 4, 3, 0]),
    ([10, 2, 8, 6, 4, 0, 12, 14], [5, 1, 4, 3, 2, 0, 6, 7]),
]


def test_sort_empty(hip_ctx, torch):
    from rustscan_amd.rasterizer import sort_pairs_u32

    k = torch.zeros(0, dtype=torch.int32, device="cuda")
    ok, ov = sort_pairs_u32(hip_ctx, k, k.clone())
    assert ok.numel() == 0 and ov.numel() == 0


@pytest.mark.parametrize("keys,expected", ARGSORT_CASES)
def test_sort_golden(hip_ctx, torch, keys, expected):
    from rustscan_amd.rasterizer import sort_pairs_u32

    say(f"sort golden n={len(keys)}")
    k = to_dev_u32(torch, keys)
    v = to_dev_u32(torch, np.arange(len(keys)))
    ok, ov = sort_pairs_u32(hip_ctx, k, v)
    assert from_dev_u32(ov).tolist() == expected
    assert from_dev_u32(ok).tolist() == sorted(keys)


def test_sort_513(hip_ctx, torch):
    from rustscan_amd.rasterizer import sort_pairs_u32

    keys = np.array([(i * 37) % 101 for i in range(513)], np.uint32)
    ok, ov = sort_pairs_u32(hip_ctx, to_dev_u32(torch, keys), to_dev_u32(torch, np.arange(513)))
    assert from_dev_u32(ov).tolist() == sorted(range(513), key=lambda i: (int(keys[i]), i))


def test_sort_by_key_golden(hip_ctx, torch):
    from rustscan_amd.rasterizer import sort_pairs_u32

    ok, ov = sort_pairs_u32(hip_ctx, to_dev_u32(torch, [9, 4, 1, 7, 4, 3, 2]),
                            to_dev_u32(torch, [90, 40, 10, 70, 41, 30, 20]))
    assert from_dev_u32(ok).tolist() == [1, 2, 3, 4, 4, 7, 9]
    assert from_dev_u32(ov).tolist() == [10, 20, 30, 40, 41, 70, 90]


@pytest.mark.parametrize("n,hi", [(1023, 7), (1025, 1 << 32), (300_000, 4096), (2_000_003, 1 << 32)])
def test_sort_random_stable(hip_ctx, torch, n, hi):
    from rustscan_amd.rasterizer import sort_pairs_u32

    say(f"sort random n={n} hi={hi}")
    rng = np.random.default_rng(n)
    keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
    ok, ov = sort_pairs_u32(hip_ctx, to_dev_u32(torch, keys), to_dev_u32(torch, np.arange(n)))
    order = np.argsort(keys, kind="stable").astype(np.uint32)
    np.testing.assert_array_equal(from_dev_u32(ov), order)
    np.testing.assert_array_equal(from_dev_u32(ok), keys[order])


SCAN_CASES = [([7], [7]), ([1, 2, 3, 4, 5, 6, 7, 8], [1, 3, 6, 10, 15, 21, 28, 36]),
              ([4, 1, 0, 9, 2, 8, 3], [4, 5, 5, 14, 16, 24, 27])]


@pytest.mark.parametrize("inp,expected", SCAN_CASES)
def test_scan_golden(hip_ctx, torch, inp, expected):
    from rustscan_amd.rasterizer import inclusive_scan_u32

    out = inclusive_scan_u32(hip_ctx, to_dev_u32(torch, inp))
    assert from_dev_u32(out).tolist() == expected


@pytest.mark.parametrize("n", [1000, 1024, 1025, 999_999, 3_000_000])
def test_scan_random(hip_ctx, torch, n):
    from rustscan_amd.rasterizer import inclusive_scan_u32

    say(f"scan random n={n}")
    x = np.random.default_rng(n).integers(0, 100, n).astype(np.uint32)
    out = inclusive_scan_u32(hip_ctx, to_dev_u32(torch, x))
    np.testing.assert_array_equal(from_dev_u32(out), np.cumsum(x, dtype=np.uint64).astype(np.uint32))
