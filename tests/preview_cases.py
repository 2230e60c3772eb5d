"""Scenes and independent float64 restatements for the preview renderer
tests (gpu/compute.rs, gpu/splat.wgsl)."""
import numpy as np

from rgsplat_amd import synthetic as S

POS, OPL, ROT, LS, SH = 0, 3, 4, 8, 11
SH_C0 = 0.28209479177387814


def smoke_gaussian():
    """tests/gpu_smoke.rs:40-44."""
    g = np.zeros((1, 59), np.float32)
    g[0, POS:POS + 3] = [0.0, 0.0, 2.0]
    g[0, OPL] = np.log(np.float32(0.9) / (np.float32(1) - np.float32(0.9)))
    g[0, ROT] = 1.0
    g[0, LS:LS + 3] = -1.5
    g[0, SH] = 1.0
    return g


def scene(n, seed, opacity=None, iso=False, spread=1.0):
    g = S.make_scene(n, seed, log_scale_mean=-3.0)
    g[:, POS:POS + 3] *= spread
    if iso:
        g[:, LS + 1] = g[:, LS]
        g[:, LS + 2] = g[:, LS]
    if opacity is not None:
        p = np.random.default_rng(seed).uniform(opacity[0], opacity[1], len(g))
        g[:, OPL] = np.log(p / (1 - p)).astype(np.float32)
    return g.astype(np.float32)


def view64(cam):
    return np.linalg.inv(np.asarray(cam.c2w, np.float64))


def project64(g, cam):
    """splat.wgsl `project` in float64 with numpy matrices: the Jacobian as
    the shader builds it (mat3x3 columns, i.e. the transposed perspective
    Jacobian), unnormalised quaternion, 0.3 low-pass. Returns (n, 10) with
    NaN rows for culled Gaussians."""
    V = view64(cam)
    g = np.asarray(g, np.float64)
    out = np.full((len(g), 10), np.nan)
    for i, r in enumerate(g):
        pc = V @ np.r_[r[POS:POS + 3], 1.0]
        tx, ty, tz = pc[:3]
        alpha = 1 / (1 + np.exp(-r[OPL]))
        if tz <= 0.001 or alpha < 1 / 255:
            continue
        w, x, y, z = r[ROT:ROT + 4]
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        RS = R @ np.diag(np.exp(r[LS:LS + 3]))
        sig = RS @ RS.T
        J = np.array([[cam.fx / tz, 0, 0], [0, cam.fy / tz, 0],
                      [-cam.fx * tx / tz ** 2, -cam.fy * ty / tz ** 2, 0]])
        T = J @ V[:3, :3]
        s2 = T @ sig @ T.T
        a, b, c = s2[0, 0] + 0.3, s2[1, 0], s2[1, 1] + 0.3
        det = a * c - b * b
        if abs(det) < 1e-8:
            continue
        col = np.clip(0.5 + r[[SH, SH + 16, SH + 32]] * SH_C0, 0, 1)
        out[i] = [cam.fx * tx / tz + cam.cx, cam.fy * ty / tz + cam.cy, c / det, -b / det, a / det, *col, alpha, tz]
    return out


def composite_all_pixels(sp, w, h, bg):
    """Front-to-back compositing of every visible splat over every pixel in
    float64 (no 3-sigma rect): equal to composite (gpu/compute.rs:243-297)
    wherever no splat reaches alpha 1/255 outside its rect, i.e. whenever
    every splat's alpha is below 0.35 (exp(-4.5) * 0.35 < 1/255)."""
    vis = [i for i in range(len(sp)) if not np.isnan(sp[i, 0]) and sp[i, 8] >= 1 / 255 and sp[i, 9] > 0]
    vis.sort(key=lambda i: (sp[i, 9], i))
    px, py = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
    col = np.zeros((h, w, 3))
    T = np.ones((h, w))
    for i in vis:
        s = sp[i]
        dx, dy = px - s[0], py - s[1]
        power = -0.5 * (s[2] * dx * dx + 2 * s[3] * dx * dy + s[4] * dy * dy)
        a = np.minimum(s[8] * np.exp(np.minimum(power, 0)), 0.99)
        use = (power <= 0) & (a >= 1 / 255) & (T >= 1e-4)
        wgt = np.where(use, a * T, 0.0)
        col += wgt[..., None] * s[5:8]
        T = np.where(use, T * (1 - a), T)
    return col + T[..., None] * np.asarray(bg, np.float64)
