"""Checks of one device training step against the CPU oracle.

A step of ``Context.train_step`` is (a) the forward and backward of every
listed view -- compared with ``oracle.Render`` per view (render_and_backward,
cpu/backward.rs:107-149), the device gradients summed over the views -- and
(b) one Adam step on those summed gradients -- compared with
``oracle.adam_step`` (pipeline/optimizer.rs:116-163) applied to the device's
own captured gradients from the device's state before the step, so that the
Adam arithmetic is checked apart from the gradient summation order.

Bars:
* losses: 1e-5 relative to the reference's per-pixel f32 terms added in
  double (the device sums per tile in f32, then in f64). The reference adds
  the 3 w h terms serially in f32 (backward.rs:670-676, and per channel in
  ssim_channel), which at 800x800 and above is off by 1e-3 to a few 1e-2
  relative; that deviation is reported, not asserted (no gradient depends on
  the scalar);
* images: 1e-4 absolute;
* gradients: 1e-3 relative per parameter group, ||g - g_ref|| / ||g_ref||
  (wave reductions and FP32 atomics sum in another order than the reference's
  serial per-tile merge); and 1e-4 against the reference backward driven by
  the loss gradient of the device's own image (the signum of the L1 term
  flips where rounding moves a pixel across its target value);
* Adam: |p - p_ref| <= 1e-5 lr_j + 4 ulp(p) per parameter (the step's Adam
  uses the hardware reciprocal and square root: a few ulp of the update,
  whose size is at most a few lr_j), |m1 - m1_ref| <= 4 ulp of
  |beta1 m1_prev| + |(1 - beta1) g| (fused multiply-add against two
  roundings), |m2 - m2_ref| <= 4 ulp(m2_ref); grad_accum bit-exact (the same
  f32 addition of the captured norm sum), grad_denom + one per view.
"""
from __future__ import annotations

import numpy as np

import oracle

GROUPS = {"position": slice(0, 3), "rotation": slice(3, 7), "log_scale": slice(7, 10),
          "opacity": slice(10, 11), "sh_dc": [11, 27, 43], "sh": slice(11, 59), "screen_grad_norm": slice(59, 60)}

# Gaussian3D column -> flat Adam index (optimizer.rs:165-173)
AOS_TO_FLAT = np.array([0, 1, 2, 10, 3, 4, 5, 6, 7, 8, 9] + list(range(11, 59)))


def with_alpha(rgb):
    return np.concatenate([rgb, np.ones(rgb.shape[:2] + (1,), np.float32)], -1)


def group_rel_err(a, b) -> dict:
    out = {}
    for k, sl in GROUPS.items():
        x, y = a[:, sl].astype(np.float64), b[:, sl].astype(np.float64)
        den = np.linalg.norm(y)
        out[k] = float(np.linalg.norm(x - y) / den if den > 0 else np.linalg.norm(x - y))
    return out


def component_rel_err(a, b) -> dict:
    """Per-component |g - g_ref| / max(|g_ref|, 1e-6 max|g_ref|) over the 59
    parameter gradients: the maximum and the 99.9th percentile."""
    x, y = a[:, :59].astype(np.float64), b[:, :59].astype(np.float64)
    floor = 1e-6 * np.abs(y).max()
    r = np.abs(x - y) / np.maximum(np.abs(y), floor)
    return {"max": float(r.max()), "p999": float(np.quantile(r, 0.999))}


def oracle_views(g, cams, targets, bg, lam, dev_img=None):
    """Per-view losses (the reference's f32 sums, and its f32 per-pixel terms
    added in double), images (flat, views in order) and the summed f64
    gradients of the oracle's render_and_backward. With dev_img (the device's
    images, flat, views in order) also the summed gradients of the oracle's
    backward driven by the loss gradient of the device's image."""
    losses, losses64, imgs = [], [], []
    tot = np.zeros((len(g), 60), np.float64)
    tot_dev = np.zeros((len(g), 60), np.float64) if dev_img is not None else None
    base = 0
    for c, t in zip(cams, targets):
        r = oracle.Render(g, c, t, bg=bg, lam=lam)
        losses.append(r.loss)
        losses64.append(r.loss_exact)
        rgb = r.image()[0]
        imgs.append(rgb.reshape(-1, 3))
        tot += r.grads()
        if dev_img is not None:
            npx = c.width * c.height
            dimg = dev_img[base:base + npx].reshape(c.height, c.width, 3)
            base += npx
            _, dl = oracle.loss_grad(dimg, t, c.width, c.height, lam)
            tot_dev += r.backward_with_dl(dl)
        del r
    return np.array(losses), np.array(losses64), np.concatenate(imgs), tot, tot_dev


def lr_per_flat(cfg) -> np.ndarray:
    lr = np.empty(59, np.float64)
    lr[0:3] = cfg["lr_position"]
    lr[3:7] = cfg["lr_rotation"]
    lr[7:10] = cfg["lr_scale"]
    lr[10] = cfg["lr_opacity"]
    lr[11:14] = cfg["lr_sh_dc"]
    lr[14:] = cfg["lr_sh_rest"]
    return lr


def check_adam(g_before, st_before, grads, g_after, st_after, cfg, n_views) -> dict:
    """The device's Adam step against oracle.adam_step on the device's own
    gradients from the same state; returns the worst ratios to the bars."""
    n = len(g_before)
    st_ref = oracle.AdamState(n)
    st_ref.m1[:] = st_before.m1[:n]
    st_ref.m2[:] = st_before.m2[:n]
    st_ref.accum[:] = st_before.grad_accum[:n]
    st_ref.denom[:] = st_before.grad_denom[:n]
    st_ref.t = st_before.t
    g_ref = np.ascontiguousarray(g_before.copy())
    gr = np.ascontiguousarray(grads, np.float32)
    oracle.adam_step(g_ref, gr, st_ref, cfg)
    assert st_after.t == st_ref.t
    lr = lr_per_flat(cfg)[AOS_TO_FLAT]
    p_bar = 1e-5 * lr[None, :] + 4 * np.spacing(np.maximum(np.abs(g_ref), np.abs(g_before)).astype(np.float32))
    dp = np.abs(g_after.astype(np.float64) - g_ref)
    b1, b2 = cfg["beta1"], cfg["beta2"]
    m_prev = st_before.m1[:n].astype(np.float32)
    m_scale = (np.abs(np.float32(b1) * m_prev) + np.abs(np.float32(1 - b1) * gr[:, :59])).astype(np.float32)
    m_bar = 4 * np.spacing(m_scale)
    dm = np.abs(st_after.m1[:n].astype(np.float64) - st_ref.m1)
    v_bar = 4 * np.spacing(st_ref.m2.astype(np.float32))
    dv = np.abs(st_after.m2[:n].astype(np.float64) - st_ref.m2)
    out = {"params": float((dp / p_bar).max()), "m1": float((dm / m_bar).max()), "m2": float((dv / v_bar).max())}
    assert out["params"] <= 1.0, f"parameters off the Adam bar: {out}"
    assert out["m1"] <= 1.0, f"m1 off its bar: {out}"
    assert out["m2"] <= 1.0, f"m2 off its bar: {out}"
    assert np.array_equal(st_after.grad_accum[:n], st_ref.accum), "grad_accum differs"
    assert np.array_equal(st_after.grad_denom[:n], st_before.grad_denom[:n] + n_views), "grad_denom"
    return out


def run_checked_step(ctx, ids, cams, targets, cfg, bg=(0.0, 0.0, 0.0), lam=0.2, use_graph=True,
                     check_grads=True) -> dict:
    """One train_step on views ids (gradient capture on), checked against the
    oracle; cams and targets are the listed views' in order."""
    g0 = ctx.scene_download()
    st0 = ctx.adam_download()
    loss = ctx.train_step(ids, cfg, bg=bg, lambda_dssim=lam, use_graph=use_graph)
    grads = ctx.step_grads()
    g1 = ctx.scene_download()
    st1 = ctx.adam_download()
    rep = {"loss": loss.tolist()}
    if check_grads:
        npix = sum(c.width * c.height for c in cams)
        img = ctx.step_image(npix)
        ref_loss, ref_loss64, ref_img, ref_g, ref_g_dev = oracle_views(g0, cams, targets, bg, lam, img)
        rep["loss_rel_err"] = float(np.max(np.abs(loss - ref_loss64) / np.abs(ref_loss64)))
        rep["reference_f32_loss_rel_dev"] = float(np.max(np.abs(ref_loss - ref_loss64) / np.abs(ref_loss64)))
        rep["image_max_abs_err"] = float(np.abs(img - ref_img).max())
        rep["grad_group_rel_err"] = group_rel_err(grads, ref_g)
        rep["grad_component_rel_err"] = component_rel_err(grads, ref_g)
        # The same backward driven by the device image's loss gradient: the
        # L1 term's signum(x - y) jumps where the rendered value meets the
        # target, so rounding-level image differences (the step's forward
        # evaluates alpha with the hardware exp) flip it on a few pixels of a
        # trained scene; against this the step's backward is compared alone.
        rep["grad_group_rel_err_same_dl"] = group_rel_err(grads, ref_g_dev)
        assert rep["loss_rel_err"] <= 1e-5, rep
        assert rep["image_max_abs_err"] <= 1e-4, rep
        bad = {k: v for k, v in rep["grad_group_rel_err"].items() if not v <= 1e-3}
        assert not bad, f"gradient groups above 1e-3 relative: {bad} ({rep})"
        bad = {k: v for k, v in rep["grad_group_rel_err_same_dl"].items() if not v <= 1e-4}
        assert not bad, f"gradient groups above 1e-4 relative for the same dL/dpixel: {bad} ({rep})"
        culled = np.all(ref_g == 0, axis=1)
        assert np.all(grads[culled] == 0), "a Gaussian no view sees got a gradient"
    rep["adam_bar_ratio"] = check_adam(g0, st0, grads, g1, st1, cfg, len(ids))
    return rep
