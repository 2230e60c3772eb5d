"""load_target's resize (io/images.rs:57-61, image 0.25's resize_exact with
the Triangle filter) on the device against the oracle: bit-identical bytes,
and views registered from images of other sizes train exactly as the
oracle's resized and converted targets."""
import numpy as np
import pytest

import oracle
from rgsplat_amd import synthetic as S

pytestmark = pytest.mark.gpu

SIZES = [((16, 24), (8, 12)), ((5, 7), (14, 10)), ((480, 640), (270, 480)), ((23, 37), (100, 7)),
         ((1, 9), (9, 1)), ((1500, 2000), (540, 960)), ((300, 200), (301, 199)), ((11, 13), (11, 13))]


@pytest.fixture(scope="module")
def ctx():
    from rgsplat_amd import runtime
    c = runtime.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("k", range(len(SIZES)))
def test_resize_matches_oracle(ctx, k):
    (h, w), (dh, dw) = SIZES[k]
    img = np.random.default_rng(k).integers(0, 256, (h, w, 4), dtype=np.uint8)
    assert np.array_equal(ctx.resize_rgba8(img, dw, dh), oracle.resize_rgba8(img, dw, dh))


def test_views_from_resized_rgba8(ctx):
    from rgsplat_amd import runtime
    wl = S.scaled(S.WORKLOADS["c0_10k_256"], n=2000)
    g, cams = wl.scene(), wl.cameras()
    rng = np.random.default_rng(6)
    src = [(c.height * 3 // 2 + 1, c.width * 2 // 3) for c in cams]
    ims = [rng.integers(0, 256, (h, w, 4), dtype=np.uint8) for h, w in src]
    cfg = S.benchmark_train_config()
    losses = []
    for mode in ("sized", "float"):
        c = runtime.Context(0)
        c.scene_upload(g)
        if mode == "sized":
            c.set_views_rgba8_sized(cams, ims)
        else:
            c.set_views(cams, [oracle.rgba8_to_linear(oracle.resize_rgba8(im, cm.width, cm.height))
                               for im, cm in zip(ims, cams)])
        losses.append([c.train_step([0], cfg, use_graph=False)[0] for _ in range(2)])
        c.close()
    assert losses[0] == losses[1]
