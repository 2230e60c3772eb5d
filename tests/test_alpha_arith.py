"""The raster's rewrites of the reference's alpha arithmetic, checked in f32
over ordinary, boundary, denormal and non-finite inputs (k_raster.hip,
alpha_fast / clamp_alpha; the reference's form is backward.rs:338-352):

* the cull test `power > 0` with power = -0.5 qf is the test qf < -2^-149;
* the hardware exp2 argument qf * (-0.5 log2 e) is bit-identical to
  power * log2 e wherever |qf| >= 2^-125;
* `!(raw <= 0.99) ? 0.99 : raw` is f32::min(raw, 0.99), NaN included.
"""
import numpy as np

f = np.float32


def sample_qf():
    rng = np.random.default_rng(7)
    tiny = np.float32(2.0 ** -149)
    vals = [0.0, -0.0, tiny, -tiny, 2 * tiny, -2 * tiny, 3 * tiny, -3 * tiny,
            np.float32(2.0 ** -126), -np.float32(2.0 ** -126), np.float32(2.0 ** -125), 1.0, -1.0,
            np.inf, -np.inf, np.nan, np.finfo(np.float32).max, -np.finfo(np.float32).max]
    normal = rng.normal(0.0, 10.0, 200_000).astype(f)
    wide = (rng.uniform(-1, 1, 200_000) * 2.0 ** rng.uniform(-149, 127, 200_000)).astype(f)
    den = (rng.integers(-2 ** 23, 2 ** 23, 50_000) * tiny).astype(f)
    return np.concatenate([np.array(vals, f), normal, wide, den])


def test_cull_test_is_power_positive():
    qf = sample_qf()
    with np.errstate(invalid="ignore", over="ignore"):
        power = f(-0.5) * qf
        assert np.array_equal(power > 0, qf < f(-(2.0 ** -149)))


def test_exp_argument_bit_identical():
    qf = sample_qf()
    qf = qf[np.isfinite(qf) & (np.abs(qf) >= f(2.0 ** -125))]
    log2e = f(1.44269504088896341)
    with np.errstate(over="ignore"):
        a = (f(-0.5) * qf) * log2e
        b = qf * (f(-0.5) * log2e)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_clamp_is_f32_min():
    raw = np.concatenate([sample_qf(), np.array([0.99, np.nextafter(f(0.99), f(1)), np.nextafter(f(0.99), f(0))], f)])
    with np.errstate(invalid="ignore"):
        got = np.where(~(raw <= f(0.99)), f(0.99), raw)
    ref = np.fmin(raw, f(0.99))  # NaN -> the other operand, as Rust's f32::min
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
