"""The preview's tile cut (k_preview.hip k_preview_rects): a splat is left
out of the tiles its alpha-reach box misses. The cut is only correct if no
pixel outside that box reaches alpha 1/255 in composite's own f32 test.
This restates the box in numpy f32 and checks it against composite's alpha
test, pixel by pixel, on splats from the oracle's projection with strongly
anisotropic covariances, all opacities and a range of sizes."""
import numpy as np

import oracle
from rgsplat_amd import synthetic as S
from tests import preview_cases as C

f = np.float32


def reach_box(sp):
    """k_preview_rects: half-widths sqrt(R Sxx) + 0.5, sqrt(R Syy) + 0.5 with
    R = max(2 ln(255 alpha) (1.02 + 20 q_err) + 1e-3, 1e-3), q_err = 1e-6 l1/l2;
    None where the kernel does not cut."""
    a, b, c, alpha = f(sp[2]), f(sp[3]), f(sp[4]), f(sp[8])
    mid = f(0.5) * (a + c)
    disc = np.sqrt(f(0.25) * (a - c) * (a - c) + b * b)
    l1, l2 = mid + disc, mid - disc
    q_err = f(1e-6) * (max(l1, f(0)) / max(l2, f(1e-30)))
    det = a * c - b * b
    if not (q_err <= f(0.02) and l2 > 0 and det > 0):
        return None
    R = max(f(2) * np.log(f(255) * alpha) * (f(1.02) + f(20) * q_err) + f(1e-3), f(1e-3))
    return np.sqrt(R * (c / det)) + f(0.5), np.sqrt(R * (a / det)) + f(0.5)


def test_reach_box_contains_every_contributing_pixel():
    checked = 0
    for seed, (ls_mean, ls_std) in enumerate([(-3.0, 1.0), (-2.0, 1.5), (-4.0, 0.3)]):
        g = S.make_scene(400, 100 + seed, log_scale_mean=ls_mean, log_scale_std=ls_std)
        p = np.random.default_rng(seed).uniform(0.004, 0.999, len(g))
        g[:, 3] = np.log(p / (1 - p)).astype(np.float32)
        cam = S.look_at_origin(np.array([0.3, 0.4, 1.0]), 3.0, 160, 120, 150.0)
        sp = oracle.preview_project(g, cam)
        for s in sp:
            if not (s[8] >= 1 / 255 and s[9] > 0):
                continue
            box = reach_box(s)
            if box is None:
                continue
            hx, hy = box
            # composite's per-pixel test over a generous window around the splat
            cx, cy = float(s[0]), float(s[1])
            if not (np.isfinite(cx) and np.isfinite(cy)) or abs(cx) > 1e4 or abs(cy) > 1e4:
                continue
            r = int(min(max(hx, hy) * 3 + 4, 400))
            px = np.arange(int(cx) - r, int(cx) + r + 1)
            py = np.arange(int(cy) - r, int(cy) + r + 1)
            X, Y = np.meshgrid(px, py)
            dx = (X.astype(f) + f(0.5)) - f(s[0])
            dy = (Y.astype(f) + f(0.5)) - f(s[1])
            power = f(-0.5) * (f(s[2]) * dx * dx + f(2) * f(s[3]) * dx * dy + f(s[4]) * dy * dy)
            alpha = np.minimum(f(s[8]) * np.exp(power.astype(np.float64)), 0.99)
            hit = (power <= 0) & (alpha >= (1 / 255) * (1 - 1e-6))
            ccx = X.astype(f) + f(0.5)
            ccy = Y.astype(f) + f(0.5)
            inside = (np.abs(ccx - f(s[0])) <= hx) & (np.abs(ccy - f(s[1])) <= hy)
            assert not np.any(hit & ~inside), (s, hx, hy)
            checked += int(hit.sum())
    assert checked > 10000
