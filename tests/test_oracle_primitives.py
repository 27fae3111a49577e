"""Oracle sort and scan against the reference's own known-answer tests.

Pins: orc_sort_by_key (the sort contract of radix_sort.wgsl:21-26, stable,
ties on input index) with the vectors of
RustGS/src/training/gpu_primitives/radix_sort/tests.rs:89-121, and
orc_inclusive_scan (prefix_sum.wgsl) with the vectors of
RustGS/src/training/gpu_primitives/prefix_sum/tests.rs:41-59.  Expected
outputs are written out literally (they are what the reference's tests
compute with Rust's sort_by_key / scan).
"""
import numpy as np
import pytest


ARGSORT_CASES = [
    ([], []),
    ([42], [0]),
    ([9, 4, 1, 7, 4, 3, 2], [2, 6, 5, 1, 4, 3, 0]),
    ([10, 2, 8, 6, 4, 0, 12, 14], [5, 1, 4, 3, 2, 0, 6, 7]),
]


@pytest.mark.parametrize("keys,expected", ARGSORT_CASES)
def test_argsort_golden(orc, keys, expected):
    k = np.array(keys, np.uint32)
    _, v = orc.sort_by_key(k, np.arange(len(keys), dtype=np.uint32))
    assert v.tolist() == expected


def test_argsort_large(orc):
    # radix_sort/tests.rs:110-113: 513 keys (i*37) % 101, expected = stable sort by (key, index)
    keys = np.array([(i * 37) % 101 for i in range(513)], np.uint32)
    _, v = orc.sort_by_key(keys, np.arange(513, dtype=np.uint32))
    expected = sorted(range(513), key=lambda i: (int(keys[i]), i))
    assert v.tolist() == expected
    # the first run of equal keys is key 0: indices 0, 101, 202, 303, 404, 505
    assert v[:6].tolist() == [0, 101, 202, 303, 404, 505]


def test_sort_by_key_golden(orc):
    # radix_sort/tests.rs:115-121
    k, v = orc.sort_by_key(np.array([9, 4, 1, 7, 4, 3, 2], np.uint32),
                           np.array([90, 40, 10, 70, 41, 30, 20], np.uint32))
    assert k.tolist() == [1, 2, 3, 4, 4, 7, 9]
    assert v.tolist() == [10, 20, 30, 40, 41, 70, 90]


SCAN_CASES = [
    ([], []),
    ([7], [7]),
    ([1, 2, 3, 4, 5, 6, 7, 8], [1, 3, 6, 10, 15, 21, 28, 36]),
    ([4, 1, 0, 9, 2, 8, 3], [4, 5, 5, 14, 16, 24, 27]),
]


@pytest.mark.parametrize("inp,expected", SCAN_CASES)
def test_scan_golden(orc, inp, expected):
    assert orc.inclusive_scan(np.array(inp, np.uint32)).tolist() == expected
