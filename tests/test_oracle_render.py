"""Pins for the oracle's renderer, loss, backward and optimizer
(cpu/backward.rs, cpu/rasterizer.rs, pipeline/optimizer.rs).

Ports of the reference's own tests keep their inputs and tolerances; the other
checks are independent of the restatement: fp64 central finite differences,
closed-form images, brute-force binning, a direct 2-D convolution SSIM, a
textbook fp64 Adam and a line-by-line Python densify.
"""
import math

import numpy as np
import pytest

from rgsplat_amd import synthetic as S


def gaussian(pos, opacity=None, logit=None, log_scale=(-3.0, -3.0, -3.0), rot=(1.0, 0, 0, 0)):
    g = np.zeros(59, np.float32)
    g[0:3] = pos
    g[3] = -2.0 if logit is None else logit
    if opacity is not None:
        g[3] = math.log(opacity / (1 - opacity))
    g[4:8] = rot
    g[8:11] = log_scale
    return g


def solid_target(w, h, color):
    t = np.zeros((h, w, 4), np.float32)
    t[..., :3] = color
    t[..., 3] = 1.0
    return t


# ------------------------------------------------------------------ loss ---
def test_ssim_of_identical_images_is_one(orc):
    # backward.rs:849-856
    w = h = 16
    img = np.array([(i % 7) / 7.0 for i in range(w * h)], np.float32)
    s = orc.lib().orc_ssim_channel(orc._p(img), orc._p(img), w, h, None)
    assert abs(s / (w * h) - 1.0) < 1e-4


def _ssim_direct(x, y):
    """SSIM sum by direct 2-D convolution with the outer-product window."""
    k1 = np.exp(-0.5 * (np.arange(11) - 5.0) ** 2 / 1.5 ** 2)
    k1 /= k1.sum()
    K = np.outer(k1, k1)
    h, w = x.shape

    def conv(a):
        p = np.zeros((h + 10, w + 10))
        p[5:-5, 5:-5] = a
        out = np.zeros((h, w))
        for dy in range(11):
            for dx in range(11):
                out += K[dy, dx] * p[dy:dy + h, dx:dx + w]
        return out
    mx, my = conv(x), conv(y)
    sx2 = conv(x * x) - mx * mx
    sy2 = conv(y * y) - my * my
    sxy = conv(x * y) - mx * my
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    return np.sum((2 * mx * my + C1) * (2 * sxy + C2) / ((mx * mx + my * my + C1) * (sx2 + sy2 + C2)))


def test_ssim_matches_direct_convolution(orc):
    r = S.Stream(5)
    x = (r.normal((13, 17)) * 0.2 + 0.5).astype(np.float32)
    y = (r.normal((13, 17)) * 0.2 + 0.5).astype(np.float32)
    s = orc.lib().orc_ssim_channel(orc._p(x), orc._p(y), 17, 13, None)
    assert abs(s - _ssim_direct(x.astype(np.float64), y.astype(np.float64))) < 2e-5 * x.size


def test_gaussian_kernel_closed_form(orc):
    k = np.zeros(11, np.float32)
    orc.lib().orc_gaussian_kernel(orc._p(k))
    ref = np.exp(-0.5 * (np.arange(11) - 5.0) ** 2 / 2.25)
    np.testing.assert_allclose(k, ref / ref.sum(), rtol=1e-6)
    assert abs(k.sum() - 1) < 1e-6


def _loss_inputs(w, h):
    rendered = np.zeros(w * h * 3, np.float32)
    target = np.zeros(w * h * 4, np.float32)
    for p in range(w * h):
        for ch in range(3):
            rendered[p * 3 + ch] = np.float32(math.sin(np.float32((p * 3 + ch) * np.float32(0.123)))) * 0.4 + 0.5
            target[p * 4 + ch] = np.float32(math.cos(np.float32((p * 5 + ch) * np.float32(0.077)))) * 0.4 + 0.5
        target[p * 4 + 3] = 1.0
    return rendered, target


def test_image_loss_gradient_fd_reference(orc):
    # backward.rs:859-892 (f32, eps 1e-3, |analytic - numeric| < 2e-3)
    w = h = 12
    rendered, target = _loss_inputs(w, h)
    _, grad = orc.loss_grad(rendered, target, w, h, 0.2)
    eps = 1e-3
    for p, ch in [(0, 0), (40, 1), (77, 2), (143, 0)]:
        idx = p * 3 + ch
        rp, rm = rendered.copy(), rendered.copy()
        rp[idx] += eps
        rm[idx] -= eps
        num = (orc.loss_grad(rp, target, w, h, 0.2)[0] - orc.loss_grad(rm, target, w, h, 0.2)[0]) / (2 * eps)
        assert abs(grad[idx] - num) < 2e-3


def test_image_loss_gradient_fd_f64_all_pixels(orc):
    w, h = 9, 7
    rendered, target = _loss_inputs(w, h)
    rendered = rendered.astype(np.float64)
    target = target.astype(np.float64)
    _, grad = orc.loss_grad(rendered, target, w, h, 0.2, f64=True)
    eps = 1e-6
    for idx in range(w * h * 3):
        rp, rm = rendered.copy(), rendered.copy()
        rp[idx] += eps
        rm[idx] -= eps
        num = (orc.loss_grad(rp, target, w, h, 0.2, True)[0] - orc.loss_grad(rm, target, w, h, 0.2, True)[0]) / (2 * eps)
        assert abs(grad[idx] - num) < 1e-8 + 1e-6 * abs(num)
    # L1 sign convention: Rust signum(+0.0) = +1 (a zero residual gets +(1-lambda)/M)
    r2 = target.reshape(-1, 4)[:, :3].reshape(-1).copy()
    _, g0 = orc.loss_grad(r2, target, w, h, 0.0, f64=True)
    np.testing.assert_allclose(g0, 1.0 / (w * h * 3))


# ---------------------------------------------------------------- render ---
def test_render_lights_center_pixel(orc):
    # backward.rs:895-904
    cam = S.identity_camera(32, 32)
    g = gaussian([0, 0, 2.0], opacity=0.9, log_scale=(-1.5, -1.5, -1.5))
    g[11] = 2.0
    img = orc.render_color(g[None], cam)
    assert img[16, 16, 0] > 0.01


def test_on_axis_isotropic_closed_form(orc):
    # A single isotropic Gaussian on the optical axis renders as
    # op * exp(-r^2 / (2 var)) * colour with var = (f s / z)^2 + 0.3.
    w = h = 64
    cam = S.identity_camera(w, h)
    s, z, op = math.exp(-3.0), 2.0, 0.7
    g = gaussian([0, 0, z], opacity=op, log_scale=(-3.0, -3.0, -3.0))
    img = orc.render_color(g[None], cam)
    var = (w * s / z) ** 2 + 0.3
    ys, xs = np.mgrid[0:h, 0:w]
    r2 = (xs + 0.5 - w / 2) ** 2 + (ys + 0.5 - h / 2) ** 2
    alpha = np.minimum(op * np.exp(-r2 / (2 * var)), 0.99)
    # pixels with alpha < 1/255 are skipped; outside the 3-sigma tile rect nothing is drawn
    expect = np.where(alpha >= 1 / 255, alpha * 0.5, 0.0)
    rad = 3 * math.sqrt(var)
    x0, x1 = int(w / 2 - rad) // 16, (int(w / 2 + rad) + 1 + 15) // 16
    inside = (xs // 16 >= x0) & (xs // 16 < x1) & (ys // 16 >= x0) & (ys // 16 < x1)
    expect = np.where(inside, expect, 0.0)
    np.testing.assert_allclose(img[..., 0], expect, atol=2e-6)
    np.testing.assert_allclose(img[..., 1], expect, atol=2e-6)


def test_binning_brute_force(orc):
    wl = S.scaled(S.WORKLOADS["c0_10k_256"], n=3000, sizes=[(200, 136)])
    g = wl.scene()
    cam = wl.cameras()[0]
    r = orc.Render(g, cam, None, backward=False)
    sp = r.splats()
    off, gis = r.tiles()
    ntx, nty = (200 + 15) // 16, (136 + 15) // 16
    # depth order with index tie-break
    order = np.lexsort((sp["gi"], sp["depth"]))
    assert np.all(order == np.arange(len(order)))
    lists = [[] for _ in range(ntx * nty)]
    for gi, (x0, y0, x1, y1) in zip(sp["gi"], sp["rect"]):
        for ty in range(max(y0, 0), min(y1, nty)):
            for tx in range(max(x0, 0), min(x1, ntx)):
                lists[ty * ntx + tx].append(int(gi))
    flat = np.array([x for L in lists for x in L], np.uint32)
    assert np.array_equal(flat, gis)
    assert np.array_equal(np.cumsum([0] + [len(L) for L in lists]), off)


def test_rasterize_rgba_semantics(orc):
    wl = S.scaled(S.WORKLOADS["c0_10k_256"], n=500, sizes=[(96, 80)])
    g = wl.scene()
    cam = wl.cameras()[0]
    rgba = orc.rasterize_rgba(g, cam, (0.0, 0.0, 0.0))
    r = orc.Render(g, cam, None, backward=False)
    rgb, t, last = r.image()
    np.testing.assert_allclose(rgba[..., :3], rgb, atol=1e-6)
    drawn = rgba[..., 3] != 0
    np.testing.assert_allclose(rgba[..., 3][drawn], 1 - t[drawn], atol=1e-6)
    # with a background, RGB = bg + sum(w c) where a tile is drawn, 0 elsewhere
    bg = (0.2, 0.3, 0.4)
    rgba_bg = orc.rasterize_rgba(g, cam, bg)
    off, _ = r.tiles()
    ntx = (96 + 15) // 16
    for ti in range(len(off) - 1):
        ty, tx = divmod(ti, ntx)
        blk = rgba_bg[ty * 16:(ty + 1) * 16, tx * 16:(tx + 1) * 16]
        if off[ti + 1] == off[ti]:
            assert np.all(blk == 0)
        else:
            ref = rgb[ty * 16:(ty + 1) * 16, tx * 16:(tx + 1) * 16] + np.array(bg, np.float32)
            np.testing.assert_allclose(blk[..., :3], ref, atol=2e-6)


# --------------------------------------------------------------- backward ---
def test_parameter_gradients_fd_reference(orc):
    # backward.rs:907-982 (f32, eps 1e-3; rel < 0.05 or abs < 3e-4; rotation
    # compared in the tangent space with rel < 0.1)
    cam = S.identity_camera(24, 24)
    target = solid_target(24, 24, [0, 0, 0])
    g = gaussian([0.1, -0.05, 2.0], opacity=0.6, log_scale=(-2.2, -2.0, -2.1),
                 rot=orc.call3("orc_q_normalize", [0.9, 0.1, -0.2, 0.15], out_shape=(4,)))
    for ch in range(3):
        g[11 + ch * 16] = 0.4
        for k in range(1, 16):
            g[11 + ch * 16 + k] = math.sin((ch * 16 + k) * 0.05) * 0.04
    gs = g[None].copy()
    _, grads = orc.render_and_backward(gs, cam, target, lam=0.0)
    grad = grads[0]

    def num(idx, eps=1e-3):
        gp, gm = gs.copy(), gs.copy()
        gp[0, idx] += eps
        gm[0, idx] -= eps
        return (orc.render_and_backward(gp, cam, target, lam=0.0)[0] -
                orc.render_and_backward(gm, cam, target, lam=0.0)[0]) / (2 * eps)

    def check(a, n):
        rel = abs(a - n) / max(abs(a), abs(n), 1e-4)
        assert rel < 0.05 or abs(a - n) < 3e-4, (a, n)
    for k in range(3):
        check(grad[k], num(k))
    check(grad[10], num(3))
    for k in range(3):
        check(grad[7 + k], num(8 + k))
    for k in [0, 1, 16, 33]:
        check(grad[11 + k], num(11 + k))
    q = gs[0, 4:8].astype(np.float64)
    nr = np.array([num(4 + k) for k in range(4)])
    tang = lambda v: v - np.dot(v, q) * q
    a, n = tang(grad[3:7].astype(np.float64)), tang(nr)
    assert np.linalg.norm(a - n) / max(np.linalg.norm(n), 1e-3) < 0.1


def _fd_scene():
    """Three overlapping Gaussians with view-dependent colour well inside (0,1)."""
    cam = S.identity_camera(40, 32)
    r = S.Stream(99)
    gs = []
    for i, (x, y, z) in enumerate([(0.05, -0.02, 2.0), (-0.08, 0.04, 2.3), (0.02, 0.1, 2.6)]):
        g = gaussian([x, y, z], opacity=0.5 + 0.1 * i, log_scale=(-2.3 + 0.1 * i, -2.0, -2.2))
        q = r.normal((4,))
        g[4:8] = q / np.linalg.norm(q)
        for ch in range(3):
            g[11 + ch * 16] = 0.3 + 0.1 * ch
            g[11 + ch * 16 + 1:11 + ch * 16 + 16] = r.normal((15,), std=0.05)
        gs.append(g)
    tgt = np.zeros((32, 40, 4))
    tgt[..., :3] = 0.3 + 0.2 * np.sin(np.arange(32 * 40 * 3).reshape(32, 40, 3) * 0.37)
    tgt[..., 3] = 1
    return np.array(gs, np.float64), cam, tgt


def test_parameter_gradients_fd_f64_all_params(orc):
    """Every one of the 59 parameters of every Gaussian against fp64 central
    differences, with D-SSIM on and a non-zero background."""
    gs, cam, tgt = _fd_scene()
    bg = (0.1, 0.05, 0.2)
    r = orc.Render(gs, cam, tgt, bg=bg, lam=0.2, f64=True)
    grads = r.grads()

    def loss(x):
        return orc.Render(x, cam, tgt, bg=bg, lam=0.2, f64=True).loss
    # Gaussian3D offset -> GaussianGrad offset
    gmap = {0: 0, 1: 1, 2: 2, 3: 10, 8: 7, 9: 8, 10: 9}
    gmap.update({11 + k: 11 + k for k in range(48)})
    eps = 1e-6
    for i in range(len(gs)):
        for off, goff in gmap.items():
            gp, gm = gs.copy(), gs.copy()
            gp[i, off] += eps
            gm[i, off] -= eps
            n = (loss(gp) - loss(gm)) / (2 * eps)
            a = grads[i, goff]
            assert abs(a - n) <= 1e-6 * max(abs(n), 1e-3), (i, off, a, n)
        q = gs[i, 4:8]
        nr = np.zeros(4)
        for k in range(4):
            gp, gm = gs.copy(), gs.copy()
            gp[i, 4 + k] += eps
            gm[i, 4 + k] -= eps
            nr[k] = (loss(gp) - loss(gm)) / (2 * eps)
        tang = lambda v: v - np.dot(v, q) * q
        np.testing.assert_allclose(tang(grads[i, 3:7]), tang(nr), atol=1e-6 * max(np.abs(nr).max(), 1e-3))


def test_loss_exact_sum(orc):
    """The loss with the reference's per-pixel f32 terms added in double.
    With lambda = 0 it is the mean of |x - y| of the f32 differences, which
    numpy sums in double as well; with D-SSIM it stays within f32 rounding of
    the all-f64 loss, while the reference's serial f32 sum of 3 w h terms
    drifts at 600x600 (its rounding, not the terms')."""
    r = S.Stream(31)
    w, h = 600, 600
    x = (r.normal((h, w, 3)) * 0.2 + 0.5).astype(np.float32)
    y = np.concatenate([(r.normal((h, w, 3)) * 0.2 + 0.5), np.ones((h, w, 1))], -1).astype(np.float32)
    loss0, exact0, _ = orc.loss_grad_exact(x, y, w, h, 0.0)
    want = np.mean(np.abs(x - y[..., :3]), dtype=np.float64)
    assert abs(exact0 - want) <= 1e-13 * want
    loss, exact, dl = orc.loss_grad_exact(x, y, w, h, 0.2)
    l64, _ = orc.loss_grad(x, y, w, h, 0.2, f64=True)
    assert abs(exact - l64) <= 1e-5 * l64
    assert abs(loss - exact) > 1e-5 * exact  # the serial f32 sum has drifted (2.3e-5 here)
    _, dl32 = orc.loss_grad(x, y, w, h, 0.2)
    assert np.array_equal(dl, dl32)


def test_backward_with_given_loss_gradient(orc):
    """Render.backward_with_dl drives the same backward_pass and
    geometry_backprop (backward.rs:131-146) with a given dL/dpixel: with the
    handle's own loss gradient it gives its gradients bit for bit, the
    backward is linear in dL/dpixel (2 dl gives exactly 2 g: scaling by a
    power of two commutes with every rounding), and a dL/dpixel that is zero
    except on one pixel reaches only the Gaussians that pixel's blend used."""
    gs, cam, tgt = _fd_scene()
    bg = (0.1, 0.05, 0.2)
    g32 = gs.astype(np.float32)
    r = orc.Render(g32, cam, tgt.astype(np.float32), bg=bg, lam=0.2)
    dl = r.dl()
    assert np.array_equal(r.backward_with_dl(dl), r.grads())
    assert np.array_equal(r.backward_with_dl(2 * dl), 2 * r.grads())
    one = np.zeros_like(dl)
    one[16, 20, 1] = 1.0
    g1 = r.backward_with_dl(one)
    assert np.all(g1[:, 11:27] == 0) and np.all(g1[:, 43:59] == 0)  # red and blue SH untouched
    assert np.any(g1[:, 27:43] != 0)
    assert np.all(r.backward_with_dl(np.zeros_like(dl)) == 0)


def test_screen_grad_norm_is_pixel_mean_gradient(orc):
    """screen_grad_norm = |dL/d(sx, sy)|. On the optical axis with an axis-
    aligned covariance and DC-only colour, moving the centre by (dx, dy) in
    camera space moves the splat by f/z (dx, dy) pixels and changes nothing
    else to first order, so the norm equals |dL/d(x, y)| z / f."""
    gs, cam, tgt = _fd_scene()
    gs = gs[:1].copy()
    gs[0, 0:2] = 0.0
    gs[0, 4:8] = [1, 0, 0, 0]
    for ch in range(3):
        gs[0, 12 + ch * 16:11 + ch * 16 + 16] = 0.0
    r = orc.Render(gs, cam, tgt, lam=0.2, f64=True)
    gr = r.grads()[0]
    z = gs[0, 2]
    f = cam.fx

    def loss_at(dx, dy):
        x = gs.copy()
        x[0, 0] += dx
        x[0, 1] += dy
        return orc.Render(x, cam, tgt, lam=0.2, f64=True).loss
    e = 1e-7
    gx = (loss_at(e, 0) - loss_at(-e, 0)) / (2 * e)
    gy = (loss_at(0, e) - loss_at(0, -e)) / (2 * e)
    assert abs(gr[59] - math.hypot(gx, gy) * z / f) < 1e-5 * gr[59]


# -------------------------------------------------------------- optimizer ---
def test_adam_matches_textbook_f64(orc):
    r = S.Stream(3)
    n = 50
    g = S.make_scene(n, 5)
    cfg = S.train_config()
    st = orc.AdamState(n)
    ref_p = g.astype(np.float64).copy()
    m = np.zeros((n, 59))
    v = np.zeros((n, 59))
    flat_of = [0, 1, 2, 4, 5, 6, 7, 8, 9, 10, 3] + list(range(11, 59))
    lr = np.array([cfg["lr_position"]] * 3 + [cfg["lr_rotation"]] * 4 + [cfg["lr_scale"]] * 3 +
                  [cfg["lr_opacity"]] + [cfg["lr_sh_dc"]] * 3 + [cfg["lr_sh_rest"]] * 45)
    for t in range(1, 6):
        grads = np.zeros((n, 60), np.float32)
        grads[:, :59] = r.normal((n, 59), std=1e-2)
        grads[:, 59] = np.abs(r.normal((n,)))
        orc.adam_step(g, grads, st, cfg)
        gd = grads[:, :59].astype(np.float64)
        m = 0.9 * m + 0.1 * gd
        v = 0.999 * v + 0.001 * gd * gd
        upd = lr * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
        ref_p[:, flat_of] -= upd
        q = ref_p[:, 4:8]
        ref_p[:, 4:8] = q / np.linalg.norm(q, axis=1, keepdims=True)
    assert st.t == 5
    np.testing.assert_allclose(g, ref_p, atol=2e-6, rtol=1e-5)
    assert np.all(st.denom[:n] == 5)


def _densify_python(g, n, st, cfg, extent, split_div=0.0):
    """Line-by-line Python restatement of optimizer.rs:206-311 in fp64 math,
    rounded to f32 at the same points; split_div > 0: every child log-scale is
    ln(max(s_j / split_div, 1e-6)) instead of the reference's axis-0 split."""
    f = np.float32
    sig = lambda x: f(1) / (f(1) + f(math.exp(-float(x))))
    to_add, remove = [], [False] * n
    avg = [f(st.accum[i] / f(st.denom[i])) if st.denom[i] > 0 else f(0) for i in range(n)]
    for i in range(n):
        gi = g[i]
        a = sig(gi[3])
        s = [f(math.exp(float(gi[8 + k]))) for k in range(3)]
        ms = max(s)
        if a < f(cfg["prune_opacity_thresh"]) or ms > f(0.1) * f(extent):
            remove[i] = True
            continue
        if avg[i] >= f(cfg["densify_grad_threshold"]) and n + len(to_add) < cfg["max_gaussians"]:
            if ms > f(0.01) * f(extent):
                remove[i] = True
                q = gi[4:8].astype(np.float64)
                q = q / np.linalg.norm(q)
                w, x, y, z = q
                R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                              [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                              [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
                ax = R[:, 0]
                off = ax / np.linalg.norm(ax) * s[0]
                c1, c2 = gi.copy(), gi.copy()
                c1[0:3] += off
                c2[0:3] -= off
                ns = math.log(max(s[0] / math.sqrt(2), 1e-6))
                op = min(max(a * 0.8, 1e-6), 0.999)
                lg = math.log(op / (1 - op))
                for c in (c1, c2):
                    c[8] = ns
                    if split_div > 0:
                        c[8:11] = [math.log(max(s[k] / split_div, 1e-6)) for k in range(3)]
                    c[3] = lg
                to_add += [c1, c2]
            else:
                c = gi.copy()
                c[0] += avg[i] * 0.01
                c[1] += avg[i] * 0.01
                to_add.append(c)
    kept = [g[i] for i in range(n) if not remove[i]]
    return np.array(kept + to_add, np.float32).reshape(-1, 59)


def test_densify_matches_python_restatement(orc):
    n = 400
    g = S.make_scene(n, 17, log_scale_mean=-3.0, log_scale_std=1.0)
    g[:, 3] = S.Stream(1).normal((n,), std=3.0)  # some below the prune threshold
    extent = orc.scene_extent(g)
    cfg = S.train_config(max_gaussians=n + 120)  # the cap binds part-way
    st = orc.AdamState(n, cap=2 * n)
    rr = S.Stream(2)
    st.accum[:n] = np.abs(rr.normal((n,), std=4e-4))
    st.denom[:n] = 1
    st.m1[:n] = 1.0
    ref = _densify_python(g, n, st, cfg, extent)
    buf = np.zeros((2 * n, 59), np.float32)
    buf[:n] = g
    nn = orc.densify_and_prune(buf, n, st, cfg, extent)
    assert nn == len(ref)
    got = buf[:nn]
    # positions/logits/log-scales may differ in the last bits (libm vs numpy)
    np.testing.assert_allclose(got, ref, rtol=2e-6, atol=2e-6)
    # moments are not compacted, only grown (optimizer.rs:97-105)
    assert st.rows == max(n, nn)
    assert np.all(st.m1[:n] == 1.0) and np.all(st.m1[n:st.rows] == 0.0)
    assert np.all(st.accum[:st.rows] == 0) and np.all(st.denom[:st.rows] == 0)
    # the max_gaussians cap is reached: exactly one addition sequence fits
    assert nn <= n + 120 + 2


def test_densify_split_scale_divisor(orc):
    """The split-factor knob (not the reference's): children of a split get
    every scale divided by the divisor; everything else, and the decisions,
    are the reference's."""
    n = 400
    g = S.make_scene(n, 17, log_scale_mean=-3.0, log_scale_std=1.0)
    extent = orc.scene_extent(g)
    cfg = S.train_config()
    acc = np.abs(S.Stream(3).normal((n,), std=4e-4)).astype(np.float32)
    out = {}
    for d in (0.0, 1.6):
        st = orc.AdamState(n, cap=3 * n)
        st.accum[:n] = acc
        st.denom[:n] = 1
        buf = np.zeros((3 * n, 59), np.float32)
        buf[:n] = g
        nn = orc.densify_and_prune(buf, n, st, cfg, extent, split_div=d)
        st2 = orc.AdamState(n, cap=3 * n)
        st2.accum[:n] = acc
        st2.denom[:n] = 1
        ref = _densify_python(g, n, st2, cfg, extent, split_div=d)
        assert nn == len(ref)
        np.testing.assert_allclose(buf[:nn], ref, rtol=2e-6, atol=2e-6)
        out[d] = buf[:nn]
    a, b = out[0.0], out[1.6]
    assert a.shape == b.shape
    differ = np.any(a != b, axis=1)
    assert differ.sum() > 0 and np.array_equal(np.delete(a[differ], [8, 9, 10], 1), np.delete(b[differ], [8, 9, 10], 1))


def test_reset_opacities(orc):
    g = S.make_scene(64, 3)
    before = 1 / (1 + np.exp(-g[:, 3].astype(np.float64)))
    orc.reset_opacities(g)
    after = 1 / (1 + np.exp(-g[:, 3].astype(np.float64)))
    np.testing.assert_allclose(after, np.minimum(before, 0.01), rtol=1e-5)


def test_training_reduces_loss(orc):
    # tests/training_reduces_loss.rs:38-84
    w = h = 32
    cam = S.identity_camera(w, h)
    target = solid_target(w, h, [0.2, 0.55, 0.5])
    gs = []
    for x, y in [(-0.3, -0.3), (0.3, -0.3), (-0.3, 0.3), (0.3, 0.3), (0.0, 0.0)]:
        gs.append(gaussian([x, y, 2.0], opacity=0.5, log_scale=(-1.3, -1.3, -1.3)))
    gs = np.array(gs, np.float32)
    cfg = S.train_config()
    st = orc.AdamState(len(gs))
    initial = orc.render_and_backward(gs, cam, target)[0]
    last = initial
    for _ in range(200):
        last, grads = orc.render_and_backward(gs, cam, target)
        orc.adam_step(gs, grads, st, cfg)
    assert last < initial * 0.5
    img = orc.render_color(gs, cam)
    assert np.all(np.abs(img[h // 2, w // 2] - target[0, 0, :3]) < 0.15)


def test_errors_are_reported(orc):
    cam = S.identity_camera(16, 16)
    cam.c2w = np.zeros((4, 4), np.float32)
    with pytest.raises(ValueError):
        orc.Render(S.make_scene(4, 1), cam, None, backward=False)
    cam = S.identity_camera(16, 16)
    g = gaussian([0, 0, 2.0], opacity=0.9)
    g2 = g.copy()
    g2[0] = np.nan  # NaN depth would panic in the reference's sort
    g2[2] = np.nan
    with pytest.raises(ValueError):
        orc.Render(np.stack([g, g2]), cam, None, backward=False)
