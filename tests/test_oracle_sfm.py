"""The oracle's gaussians_from_sfm (pipeline/sfm.rs:365-455) against closed
forms, brute-force nearest neighbours and an independent numpy restatement of
the window search."""
import numpy as np
import pytest

import oracle
from tests import ply_files as F
from tests import sfm_points as S

POS, OPL, ROT, LS, SH = F.POS, F.OPL, F.ROT, F.LS, F.SH


def test_line_closed_form():
    for n, h in [(4, 1.0), (7, 0.5), (51, 0.25), (300, 2.0), (1000, 0.125)]:
        xyz, rgb, perm = S.line(n, h, seed=n)
        g, sc = oracle.gaussians_from_sfm(xyz, rgb)
        np.testing.assert_allclose(sc, S.line_scales(n, h, perm), rtol=1e-7)
        np.testing.assert_allclose(np.exp(g[:, LS:LS + 3].astype(np.float64)),
                                   np.repeat(S.line_scales(n, h, perm)[:, None], 3, 1), rtol=2e-7)


def test_brute_force_first_half():
    """Up to 50 points the window is the whole cloud for every point in the
    first half of the x order: its scale is the plain 3-NN mean."""
    for n, seed in [(5, 1), (20, 2), (50, 3)]:
        xyz, rgb = S.cloud(n, seed)
        _, sc = oracle.gaussians_from_sfm(xyz, rgb)
        rank = np.argsort(np.argsort(xyz[:, 0], kind="stable"), kind="stable")
        d = np.sqrt(((xyz[:, None] - xyz[None]) ** 2).sum(-1))
        np.fill_diagonal(d, np.inf)
        knn = np.sort(d, 1)[:, :3].mean(1)
        sel = rank < n // 2
        np.testing.assert_allclose(sc[sel], knn[sel], rtol=1e-6)


@pytest.mark.parametrize("case", ["plain", "dups", "grid"])
def test_window_restatement(case):
    xyz, rgb = {"plain": lambda: S.cloud(3000, 4), "dups": lambda: S.cloud(2000, 5, dup_frac=0.3),
                "grid": lambda: S.cloud(2500, 6, int_grid=True)}[case]()
    _, sc = oracle.gaussians_from_sfm(xyz, rgb)
    assert np.array_equal(sc.view(np.uint32), S.window_scales(xyz).view(np.uint32))


def test_record_fields():
    xyz, rgb = S.cloud(600, 7)
    rgb[:256, 0] = np.arange(256)
    g, sc = oracle.gaussians_from_sfm(xyz, rgb)
    assert np.array_equal(g[:, POS:POS + 3], xyz.astype(np.float32))
    assert np.all(g[:, ROT] == 1) and np.all(g[:, ROT + 1:ROT + 4] == 0)
    # colour: 0.5 + SH_C0 * dc reproduces (c / 255)^2.2
    for ch in range(3):
        lin = (rgb[:, ch] / 255.0) ** 2.2
        np.testing.assert_allclose(0.5 + S.SH_C0 * g[:, SH + 16 * ch].astype(np.float64), lin, atol=3e-7)
    assert g[rgb[:, 0] == 255, SH][0] == np.float32(0.5 / np.float32(S.SH_C0))
    # everything else in the SH rows is zero
    rest = np.delete(g[:, SH:SH + 48], [0, 16, 32], axis=1)
    assert not rest.any()
    # opacity logit(0.1): sigmoid gives 0.1
    np.testing.assert_allclose(1 / (1 + np.exp(-g[:, OPL].astype(np.float64))), 0.1, rtol=1e-6)
    assert np.all(g[:, OPL] == g[0, OPL])
    assert np.all(g[:, LS] == g[:, LS + 1]) and np.all(g[:, LS] == g[:, LS + 2])
    np.testing.assert_allclose(np.exp(g[:, LS].astype(np.float64)), np.maximum(sc, 1e-4), rtol=2e-7)


def test_small_and_degenerate():
    # one point: no neighbours, scale 0.01
    g, sc = oracle.gaussians_from_sfm(np.array([[1.0, 2.0, 3.0]]), np.array([[10, 20, 30]], np.uint8))
    assert sc[0] == np.float32(0.01) and g[0, LS] == np.float32(np.log(np.float32(0.01)))
    # coincident points: distance 0, clamped to 1e-4 before the log
    g, sc = oracle.gaussians_from_sfm(np.ones((10, 3)), np.zeros((10, 3), np.uint8))
    assert np.all(sc == 0) and np.all(g[:, LS] == np.float32(np.log(np.float32(1e-4))))
    # two points: the single distance
    g, sc = oracle.gaussians_from_sfm(np.array([[0.0, 0, 0], [3.0, 4.0, 0]]), np.zeros((2, 3), np.uint8))
    assert np.all(sc == 5.0)
    # empty input
    g, sc = oracle.gaussians_from_sfm(np.zeros((0, 3)), np.zeros((0, 3), np.uint8))
    assert g.shape == (0, 59)


def test_nan_rules():
    """A NaN x panics in the reference's sort or search; a NaN distance panics
    in the distance sort once there are two distances, and a lone NaN
    distance is the scale (then clamped by f32::max to 1e-4)."""
    xyz, rgb = S.cloud(100, 8)
    bad = xyz.copy()
    bad[17, 0] = np.nan
    with pytest.raises(ValueError):
        oracle.gaussians_from_sfm(bad, rgb)
    bad = xyz.copy()
    bad[17, 1] = np.nan
    with pytest.raises(ValueError):
        oracle.gaussians_from_sfm(bad, rgb)
    with pytest.raises(ValueError):
        oracle.gaussians_from_sfm(np.array([[np.nan, 0, 0]]), np.zeros((1, 3), np.uint8))
    g, sc = oracle.gaussians_from_sfm(np.array([[0.0, np.nan, 0], [1.0, 0, 0]]), np.zeros((2, 3), np.uint8))
    assert np.all(np.isnan(sc)) and np.all(g[:, LS] == np.float32(np.log(np.float32(1e-4))))
