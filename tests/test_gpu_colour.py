"""Colour conversions through the C ABI against the oracle: load_target's
RGBA8 -> linear (bit-identical: the table is computed with the same libm
powf) and Framebuffer::to_u8_srgb (an f64 pow on the device: checked on
every byte boundary of linear_to_srgb, where an ulp could change a byte)."""
import numpy as np
import pytest

import oracle
from rgsplat_amd import synthetic as S

pytestmark = pytest.mark.gpu
f = np.float32


@pytest.fixture(scope="module")
def ctx():
    from rgsplat_amd import runtime
    c = runtime.Context(0)
    yield c
    c.close()


def test_rgba8_to_linear_bitexact(ctx):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (67, 129, 4), dtype=np.uint8)
    img[0, :256 // 2, :] = np.arange(128, dtype=np.uint8)[:, None]
    img[1, :128, :] = (128 + np.arange(128)).astype(np.uint8)[:, None]
    got = ctx.rgba8_to_linear(img)
    assert np.array_equal(got.view(np.uint32), oracle.rgba8_to_linear(img).view(np.uint32))


def boundary_inputs():
    """For every byte k, the f32 inputs around the x where
    linear_to_srgb(x) * 255 reaches k (found by bisection on the oracle's
    linear_to_srgb), +-4 ulp."""
    L = oracle.lib()
    val = lambda x: f(L.orc_linear_to_srgb(float(x))) * f(255)
    xs = []
    for k in range(1, 256):
        lo, hi = f(0.0), f(1.0)
        if val(hi) < k:
            continue
        for _ in range(60):
            mid = f((np.float64(lo) + np.float64(hi)) / 2)
            if mid in (lo, hi):
                break
            if val(mid) < k:
                lo = mid
            else:
                hi = mid
        x = hi
        for d in range(-4, 5):
            xs.append(np.nextafter(x, f(2) if d > 0 else f(-1), dtype=f) if d else x)
            for _ in range(abs(d) - 1):
                xs[-1] = np.nextafter(xs[-1], f(2) if d > 0 else f(-1), dtype=f)
    return np.array(xs, f)


def test_to_u8_srgb(ctx):
    xs = boundary_inputs()
    rng = np.random.default_rng(4)
    extra = np.concatenate([rng.uniform(-0.1, 1.1, 200_000).astype(f),
                            np.array([0, 1, 0.5, 0.0031308, 0.00313081, np.nan, np.inf, -np.inf, 1e-40], f)])
    v = np.concatenate([xs, extra])
    v = np.concatenate([v, np.zeros((-len(v)) % 4, f)]).reshape(-1, 4)
    assert np.array_equal(ctx.to_u8_srgb(v), oracle.to_u8_srgb(v))


def test_views_from_rgba8(ctx):
    """Targets registered from RGBA8 train exactly as the same targets
    converted by the oracle and registered as floats."""
    from rgsplat_amd import runtime
    wl = S.scaled(S.WORKLOADS["c0_10k_256"], n=2000)
    g, cams = wl.scene(), wl.cameras()
    rng = np.random.default_rng(5)
    ims = [rng.integers(0, 256, (c.height, c.width, 4), dtype=np.uint8) for c in cams]
    cfg = S.benchmark_train_config()
    losses = []
    for mode in ("rgba8", "float"):
        c = runtime.Context(0)
        c.scene_upload(g)
        if mode == "rgba8":
            c.set_views_rgba8(cams, ims)
        else:
            c.set_views(cams, [oracle.rgba8_to_linear(im) for im in ims])
        losses.append([c.train_step([0], cfg, use_graph=False)[0] for _ in range(2)])
        c.close()
    assert losses[0] == losses[1]
