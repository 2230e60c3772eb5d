"""The committed golden vectors (tests/golden/c0_small.npz): the CPU oracle
reproduces them exactly (CPU), and the HIP path matches them to the parity
bars (GPU)."""
import numpy as np
import pytest

import oracle
from tests.golden.make_golden import load


def test_oracle_reproduces_golden():
    d, cam = load()
    r = oracle.Render(d["scene"], cam, d["target"], lam=0.2)
    rgb, ft, last = r.image()
    assert np.array_equal(rgb, d["rgb"]) and np.array_equal(ft, d["final_t"])
    assert np.array_equal(last, d["last_contrib"])
    off, gis = r.tiles()
    assert np.array_equal(off, d["tile_offsets"]) and np.array_equal(gis, d["tile_gis"])
    assert r.loss == d["loss"]
    assert np.array_equal(r.grads(), d["grads"])


@pytest.mark.gpu
def test_hip_matches_golden():
    from rgsplat_amd import runtime
    from rgsplat_amd import synthetic as S
    d, cam = load()
    ctx = runtime.Context(0)
    try:
        rgb = ctx.render_color(d["scene"], cam)
        assert np.array_equal(rgb, d["rgb"])  # bar 1e-4; bit-exact expected
        off, gis = ctx.debug_tiles(len(d["tile_offsets"]) - 1, len(d["tile_gis"]) + 8)
        assert np.array_equal(off, d["tile_offsets"]) and np.array_equal(gis, d["tile_gis"])
        loss, grads = ctx.render_and_backward(d["scene"], cam, d["target"])
        assert abs(loss - d["loss"]) <= 1e-5 * d["loss"]
        ref = d["grads"]
        err = np.linalg.norm(grads - ref) / np.linalg.norm(ref)
        assert err <= 1e-3
        # Adam on the golden gradients is bit-exact
        g1 = d["scene"].copy()
        st = runtime.AdamState(len(g1))
        ctx.adam_step(g1, d["grads"], st, S.train_config())
        assert np.array_equal(g1, d["adam_params"])
        assert np.array_equal(st.m1[: len(g1)], d["adam_m1"]) and np.array_equal(st.m2[: len(g1)], d["adam_m2"])
    finally:
        ctx.close()
