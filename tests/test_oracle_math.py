"""The deterministic scalar exp/log (rustscan_amd/csrc/rgs_math.h) that the
oracle and the kernels share, checked against float64 numpy.

Pins: rgs_exp within 2 ulp of the correctly rounded exp over [-87, 88.7]
(and 0 below, +inf above, NaN through); rgs_log within 2 ulp over positive
normal and subnormal floats; both through the oracle's exported
orc_math_exp / orc_math_log.  A dropped polynomial term, a wrong Cody-Waite
constant or a wrong exponent bias fails these bounds by orders of magnitude.
"""
import numpy as np


def ulp_err(y32: np.ndarray, ref64: np.ndarray) -> np.ndarray:
    ref32 = ref64.astype(np.float32)
    spacing = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(y32.astype(np.float64) - ref64) / spacing


def test_exp_accuracy(orc):
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.uniform(-87.0, 88.7, 2_000_000),
        rng.uniform(-10.0, 0.0, 1_000_000),  # the compositing range exp(-sigma)
        np.linspace(-1e-3, 1e-3, 20001),
        [0.0, -0.0, 1.0, -1.0, 88.0, -86.9],
    ]).astype(np.float32)
    y = orc.math_exp(x)
    err = ulp_err(y, np.exp(x.astype(np.float64)))
    assert err.max() <= 2.0, f"max ulp error {err.max()} at x={x[err.argmax()]}"
    assert orc.math_exp(np.array([0.0], np.float32))[0] == 1.0


def test_expf_libm_equals_the_c_library_on_every_float(orc):
    """The evaluation render's exp (rgs_expf_libm, shared by the device and the
    oracle) against this C library's expf, which Rust's f32::exp calls, over
    all 2^32 inputs: equal bits except on two inputs (one float ulp each),
    both outside |x| < 32."""
    n, xs = orc.expf_libm_check()
    assert n <= 2, (n, xs)
    assert all(abs(float(x)) > 32.0 for x in xs), xs


def test_exp_special(orc):
    x = np.array([-88.0, -1000.0, 89.0, 1000.0, np.nan], np.float32)
    y = orc.math_exp(x)
    assert y[0] == 0.0 and y[1] == 0.0
    assert np.isinf(y[2]) and np.isinf(y[3])
    assert np.isnan(y[4])


def test_log_accuracy(orc):
    rng = np.random.default_rng(12)
    bits = rng.integers(0x00000001, 0x7F800000, 3_000_000, dtype=np.uint32)
    x = bits.view(np.float32)
    x = np.concatenate([x, np.array([1.0, 2.0, 0.5, 255.0, 1.0000001, 0.9999999], np.float32)])
    y = orc.math_log(x)
    ref = np.log(x.astype(np.float64))
    err = ulp_err(y, ref)
    ok = np.abs(ref) > 0
    assert err[ok].max() <= 2.0, f"max ulp error {err[ok].max()} at x={x[ok][err[ok].argmax()]}"
    assert orc.math_log(np.array([1.0], np.float32))[0] == 0.0


def test_log_special(orc):
    y = orc.math_log(np.array([0.0, -1.0, np.inf, np.nan], np.float32))
    assert np.isneginf(y[0]) and np.isneginf(y[1])
    assert np.isinf(y[2]) and y[2] > 0
    assert np.isnan(y[3])


def test_exp_core_equals_exp(orc):
    """The branch-free exp used by the compositing kernels equals rgs_exp bit
    for bit on (-87, 88], in particular on the compositing range [-5.6, 0]."""
    lo = np.float32(-5.6)
    # every 53rd float in [-5.6, -0] (negative floats are ordered by their bit
    # patterns), about 20 million arguments
    b0 = np.float32(-0.0).view(np.uint32)
    b1 = lo.view(np.uint32)
    x = np.arange(b0, b1 + 1, 53, dtype=np.uint64).astype(np.uint32).view(np.float32)
    np.testing.assert_array_equal(orc.math_exp_core(x), orc.math_exp(x))
    rng = np.random.default_rng(13)
    y = rng.uniform(-86.99, 88.0, 2_000_000).astype(np.float32)
    np.testing.assert_array_equal(orc.math_exp_core(y), orc.math_exp(y))


def test_exp_core_fast_equals_exp(orc):
    """The clamp-free, ldexp-scaled exp of the compositing kernels equals
    rgs_exp bit for bit on (-87, 88]; the kernels discard its value elsewhere."""
    b0 = np.float32(-0.0).view(np.uint32)
    b1 = np.float32(-5.6).view(np.uint32)
    x = np.arange(b0, b1 + 1, 53, dtype=np.uint64).astype(np.uint32).view(np.float32)
    np.testing.assert_array_equal(orc.math_exp_core_fast(x), orc.math_exp(x))
    rng = np.random.default_rng(14)
    y = rng.uniform(-86.99, 88.0, 2_000_000).astype(np.float32)
    np.testing.assert_array_equal(orc.math_exp_core_fast(y), orc.math_exp(y))
    # edges of the range: the largest argument and arguments just above -87
    e = np.array([88.0, np.nextafter(np.float32(-87.0), np.float32(0.0)), -86.5, 0.0, -0.0], np.float32)
    np.testing.assert_array_equal(orc.math_exp_core_fast(e), orc.math_exp(e))
