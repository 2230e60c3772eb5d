"""The culling of the training step, restated in numpy f32 and checked
against the alpha test itself on the oracle's splats: the raster's per-group
test (k_raster.hip wave_mask, with the reach bound R from k_project.hip) and
the tight tile rects of the training step's binning (k_project.hip, tight).
No (pixel, splat) pair whose alpha reaches 1/255 may be culled by either:
skipping one would change the image, so this is the culling's correctness
condition (the GPU parity tests cover the device code itself)."""
import numpy as np
import pytest

import oracle
from rgsplat_amd import synthetic as S

f = np.float32


def reach_bound(sp):
    """k_project.hip: R = 2 ln(255 op) (1.02 + 20 q_err) + 1e-3, +inf when
    kappa = l1 / l2 of the 2-D covariance exceeds 2e4."""
    cov, op = sp["cov2d"], sp["opacity"]
    a, b, c = cov[:, 0], cov[:, 1], cov[:, 2]
    mid = (a + c) * f(0.5)
    disc = np.sqrt((a - c) * (a - c) * f(0.25) + b * b)
    l1 = np.maximum(mid + disc, f(0))
    l2 = np.maximum(mid - disc, f(0))
    q_err = f(1e-6) * (l1 / np.maximum(l2, f(1e-30)))
    R = f(2) * np.log(f(255) * op).astype(f) * (f(1.02) + f(20) * q_err) + f(1e-3)
    R = np.maximum(R, f(1e-3))
    R[q_err > f(0.02)] = np.inf
    return R


def tight_rect(sp, R):
    """k_project.hip with tight != 0: the reference's 3-sigma tile rect cut to
    the tiles whose pixel centres meet the box |dx| <= sqrt(R a) + 0.5,
    |dy| <= sqrt(R c) + 0.5 of the alpha-reach ellipse."""
    rect = sp["rect"].astype(np.int64).copy()
    cov = sp["cov2d"]
    fin = np.isfinite(R)
    with np.errstate(invalid="ignore"):
        hx = np.sqrt(R * cov[:, 0]) + f(0.5)
        hy = np.sqrt(R * cov[:, 2]) + f(0.5)
        fx0 = np.ceil((sp["sx"] - hx - f(15.5)) * f(1 / 16))
        fx1 = np.floor((sp["sx"] + hx - f(0.5)) * f(1 / 16)) + f(1)
        fy0 = np.ceil((sp["sy"] - hy - f(15.5)) * f(1 / 16))
        fy1 = np.floor((sp["sy"] + hy - f(0.5)) * f(1 / 16)) + f(1)
    out = rect.copy()
    out[fin, 0] = np.maximum(rect[fin, 0], np.maximum(fx0[fin], rect[fin, 0]).astype(np.int64))
    out[fin, 1] = np.maximum(rect[fin, 1], np.maximum(fy0[fin], rect[fin, 1]).astype(np.int64))
    out[fin, 2] = np.minimum(rect[fin, 2], np.minimum(fx1[fin], rect[fin, 2]).astype(np.int64))
    out[fin, 3] = np.minimum(rect[fin, 3], np.minimum(fy1[fin], rect[fin, 3]).astype(np.int64))
    return out


def group_masks(sx, sy, ca, cb, cc, R, tx, ty):
    """k_raster.hip wave_mask: one boolean per 16x4 pixel group."""
    inv_a = f(1) / ca
    slope = -cb * inv_a
    k = cc - cb * cb * inv_a
    X0 = (f(tx * 16) + f(0.5)) - sx
    X1 = (f(tx * 16) + f(15.5)) - sx
    out = []
    for q in range(4):
        Y0 = (f(ty * 16 + 4 * q) + f(0.5)) - sy
        Y1 = (f(ty * 16 + 4 * q) + f(3.5)) - sy
        dmin = np.where((Y0 <= 0) & (Y1 >= 0), f(0), np.minimum(np.abs(Y0), np.abs(Y1)))
        dmin = np.maximum(dmin - f(0.01), f(0))
        rem = R - k * dmin * dmin
        w = np.sqrt(np.maximum(rem, f(0)) * inv_a) * f(1.0001) + f(0.01)
        c0, c1 = slope * Y0, slope * Y1
        culled = (rem < 0) | (np.maximum(c0, c1) + w < X0) | (np.minimum(c0, c1) - w > X1)
        out.append(~(culled & np.isfinite(R)))
    return out


@pytest.mark.parametrize("name,n,size", [("c0_10k_256", None, None),
                                         ("c1_100k_800", 20000, (400, 300)),
                                         ("c4_3m_clustered_4views", 3000, (320, 180))])
def test_culling_keeps_every_contributing_pair(name, n, size):
    wl = S.scaled(S.WORKLOADS[name], n=n, sizes=[size] if size else None)
    g, cam = wl.scene(), wl.cameras()[0]
    r = oracle.Render(g, cam, None, backward=False)
    sp = r.splats()
    off, gis = r.tiles()
    R = reach_bound(sp)
    tight = tight_rect(sp, R)
    conic, op = sp["conic"], sp["opacity"]
    row = {int(gi): k for k, gi in enumerate(sp["gi"])}
    ntx = (cam.width + 15) // 16
    kept = passing = 0
    for t in range(len(off) - 1):
        lst = gis[off[t]:off[t + 1]]
        if len(lst) == 0:
            continue
        ty, tx = divmod(t, ntx)
        ks = np.array([row[int(x)] for x in lst])
        m = group_masks(sp["sx"][ks], sp["sy"][ks], conic[ks, 0], conic[ks, 1], conic[ks, 2], R[ks], tx, ty)
        ys, xs = np.mgrid[ty * 16:ty * 16 + 16, tx * 16:tx * 16 + 16]
        inside = (xs < cam.width) & (ys < cam.height)
        fx, fy = xs.astype(f) + f(0.5), ys.astype(f) + f(0.5)
        for j, k in enumerate(ks):
            dx, dy = fx - sp["sx"][k], fy - sp["sy"][k]
            A, B, C = conic[k]
            power = f(-0.5) * (A * dx * dx + f(2) * B * dx * dy + C * dy * dy)
            alpha = np.minimum(op[k] * np.exp(power).astype(f), f(0.99))
            hit = inside & (power <= 0) & (alpha >= f(1 / 255))
            if hit.any():
                # the training step's binning keeps this (tile, splat) pair
                x0, y0, x1, y1 = tight[k]
                assert x0 <= tx < x1 and y0 <= ty < y1, (t, int(lst[j]))
            for q in range(4):
                kept += int(m[q][j])
                if hit[4 * q:4 * q + 4].any():
                    passing += 1
                    assert m[q][j], (t, int(lst[j]), q)
    # the test culls something, and keeps little beyond the groups that pass
    assert passing <= kept < 1.25 * passing
    # the tight rects drop a good share of the reference's pairs
    area = lambda r: np.maximum(r[:, 2] - r[:, 0], 0) * np.maximum(r[:, 3] - r[:, 1], 0)
    assert area(tight).sum() < 0.95 * area(sp["rect"].astype(np.int64)).sum()


def faint_sparse_scene():
    """A few faint splats (alpha reach well inside 3 sigma) in a small view:
    many tiles of their reference rects see no pixel reach alpha 1/255."""
    n = 60
    g = S.make_scene(n, 31, log_scale_mean=-2.5)
    g[:, 3] = S.Stream(5).normal((n,), mean=-3.0, std=0.5)
    cam = S.make_cameras(3005, [(128, 96)], focal=150.0)[0]
    return g, cam


def test_tight_binning_can_empty_a_tile():
    """Precondition of the GPU test with a background colour: some tile has a
    non-empty reference list while every splat's tight rect drops it (the
    reference composites it to the background, rgb = bg)."""
    g, cam = faint_sparse_scene()
    r = oracle.Render(g, cam, None, backward=False)
    sp = r.splats()
    off, gis = r.tiles()
    tight = tight_rect(sp, reach_bound(sp))
    row = {int(gi): k for k, gi in enumerate(sp["gi"])}
    ntx = (cam.width + 15) // 16
    emptied = 0
    for t in range(len(off) - 1):
        lst = gis[off[t]:off[t + 1]]
        if len(lst) == 0:
            continue
        ty, tx = divmod(t, ntx)
        ks = [row[int(x)] for x in lst]
        if not any(tight[k][0] <= tx < tight[k][2] and tight[k][1] <= ty < tight[k][3] for k in ks):
            emptied += 1
    assert emptied > 0
