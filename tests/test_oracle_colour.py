"""The oracle's colour conversions (io/images.rs:63-82 load_target's loop and
srgb_to_linear; cpu/rasterizer.rs:38-49 to_u8_srgb and :271-279
linear_to_srgb), pinned by the reference's tests (io/images.rs:89-101) and
by values that follow from the definitions."""
import numpy as np

import oracle

f = np.float32


def test_srgb_linear_endpoints():
    """io/images.rs:89-95."""
    L = oracle.lib()
    assert abs(L.orc_srgb_to_linear(0.0)) < 1e-6
    assert abs(L.orc_srgb_to_linear(1.0) - 1.0) < 1e-5
    assert L.orc_srgb_to_linear(0.5) < 0.5


def test_rgba8_to_linear_definition():
    img = np.zeros((2, 256, 4), np.uint8)
    img[0, :, 0] = np.arange(256)
    img[0, :, 3] = np.arange(256)
    img[1, :, 1] = np.arange(256)
    out = oracle.rgba8_to_linear(img)
    c = np.arange(256).astype(f) / f(255)
    assert np.array_equal(out[0, :, 3], c)  # alpha is linear: c / 255
    lin = out[0, :, 0]
    assert np.array_equal(lin[c <= f(0.04045)], (c / f(12.92))[c <= f(0.04045)])  # the linear segment
    assert lin[255] == 1.0 and lin[0] == 0.0
    assert np.all(np.diff(lin) > 0)  # monotone
    assert np.array_equal(out[1, :, 1], lin)  # the same table for every colour channel
    # within float rounding of the closed form (c/255 + 0.055)/1.055)^2.4
    hi = c > f(0.04045)
    ref = ((c.astype(np.float64) + 0.055) / 1.055) ** 2.4
    assert np.max(np.abs(lin[hi] - ref[hi]) / ref[hi]) < 1e-6


def test_to_u8_srgb_definition():
    """Truncation, not rounding (Rust `as u8`): linear 0.5 -> sRGB 0.7353569
    -> 187.5 -> 187; alpha 0.5 -> 127.5 -> 127; and linear 1 -> 254, because
    1.055 * 1 - 0.055 is 0.99999994 in f32. Out-of-range values clamp first;
    NaN clamps to NaN and `as u8` makes it 0."""
    v = np.array([[0.0, 1.0, 0.5, 0.5], [1.0, 0.0031308, -1.0, 2.0], [np.nan, 0.0, 0.0, np.nan]], f)
    assert f(1.055) * f(1.0) - f(0.055) == f(0.99999994)
    got = oracle.to_u8_srgb(v)
    assert got.tolist() == [[0, 254, 187, 127], [254, 10, 0, 255], [0, 0, 0, 0]]
