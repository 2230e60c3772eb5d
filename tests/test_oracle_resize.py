"""The oracle's restatement of load_target's resize (io/images.rs:57-61:
image 0.25.10's resize_exact with FilterType::Triangle; the crate is not
vendored in the reference, and the reference holds no resized fixtures)
against what the filter fixes in closed form: the 2x reduction weights
(1, 3, 3, 1) / 8 with the crate's edge windows, 2x enlargement as linear
interpolation at quarter offsets, constant images, the unchanged size."""
import numpy as np
import pytest

import oracle


def weights_down2(src):
    """Rows of the 2x reduction: taps 2o-1 .. 2o+2 weighted (1, 3, 3, 1) / 8;
    at o = 0 the window starts at 0 (weights 3, 3, 1 over 7), at the last
    output it ends at src (1, 3, 3 over 7)."""
    dst = src // 2
    W = np.zeros((dst, src))
    for o in range(dst):
        for i, wv in zip(range(2 * o - 1, 2 * o + 3), (1, 3, 3, 1)):
            if 0 <= i < src:
                W[o, i] = wv
    return W / W.sum(1, keepdims=True)


def weights_up2(src):
    """Rows of the 2x enlargement: output 2k is 0.25 x[k-1] + 0.75 x[k],
    output 2k+1 is 0.75 x[k] + 0.25 x[k+1]; the outermost outputs copy the
    edge pixel (their windows are clamped to the image)."""
    dst = 2 * src
    W = np.zeros((dst, src))
    for o in range(dst):
        k = o // 2
        taps = ((k - 1, 0.25), (k, 0.75)) if o % 2 == 0 else ((k, 0.75), (k + 1, 0.25))
        for i, wv in taps:
            if 0 <= i < src:
                W[o, i] = wv
    return W / W.sum(1, keepdims=True)


def apply(Wv, Wh, img):
    out = np.einsum("ys,sxc->yxc", Wv, img.astype(np.float64))
    out = np.einsum("xs,ysc->yxc", Wh, out)
    return np.floor(np.clip(out, 0, 255) + 0.5)


@pytest.mark.parametrize("shape", [(16, 24), (10, 6), (64, 40)])
def test_down2_closed_form(shape):
    h, w = shape
    img = np.random.default_rng(h * w).integers(0, 256, (h, w, 4), dtype=np.uint8)
    got = oracle.resize_rgba8(img, w // 2, h // 2).astype(np.int64)
    want = apply(weights_down2(h), weights_down2(w), img).astype(np.int64)
    # rows and columns away from the edges use dyadic weights: exact in f32
    assert np.array_equal(got[1:-1, 1:-1], want[1:-1, 1:-1])
    assert np.abs(got - want).max() <= 1


@pytest.mark.parametrize("shape", [(5, 7), (12, 3)])
def test_up2_closed_form(shape):
    h, w = shape
    img = np.random.default_rng(h + w).integers(0, 256, (h, w, 4), dtype=np.uint8)
    got = oracle.resize_rgba8(img, 2 * w, 2 * h).astype(np.int64)
    want = apply(weights_up2(h), weights_up2(w), img).astype(np.int64)
    assert np.array_equal(got, want)


def test_constant_and_identity():
    rng = np.random.default_rng(1)
    for (h, w), (dh, dw) in [((23, 37), (7, 100)), ((480, 640), (270, 480)), ((3, 3), (1000, 2)), ((9, 1), (1, 9))]:
        img = np.empty((h, w, 4), np.uint8)
        img[:] = rng.integers(0, 256, 4, dtype=np.uint8)
        out = oracle.resize_rgba8(img, dw, dh)
        assert np.array_equal(out, np.broadcast_to(img[0, 0], out.shape))
    img = rng.integers(0, 256, (11, 13, 4), dtype=np.uint8)
    assert np.array_equal(oracle.resize_rgba8(img, 13, 11), img)


def test_ramp_stays_monotone():
    ramp = np.tile(np.arange(0, 250, 5, dtype=np.uint8)[None, :, None], (3, 1, 4))
    for dw in (17, 50, 73, 180):
        out = oracle.resize_rgba8(ramp, dw, 5).astype(np.int64)
        assert np.all(np.diff(out[:, :, 0], axis=1) >= 0)
        assert out[:, 0, 0].min() >= 0 and out[:, -1, 0].max() <= 245
