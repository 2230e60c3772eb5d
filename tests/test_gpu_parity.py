"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.

Bars (written at each assert): tile binning and integer outputs bit-exact;
images within 1e-4 absolute (the forward keeps the reference's rounding, so
they are expected bit-exact and that is asserted too); gradients within 1e-3
relative per parameter group (||g_gpu - g_cpu|| / ||g_cpu||), since the GPU
sums per-pixel contributions in a different order (wave reduction + FP32
atomics) than the reference's per-tile serial merge and uses fused
multiply-adds in gradient arithmetic.
"""
import math

import numpy as np
import pytest

import oracle
from rgsplat_amd import synthetic as S

pytestmark = pytest.mark.gpu

GROUPS = {"position": slice(0, 3), "rotation": slice(3, 7), "log_scale": slice(7, 10),
          "opacity": slice(10, 11), "sh_dc": [11, 27, 43], "sh": slice(11, 59), "screen_grad_norm": slice(59, 60)}


@pytest.fixture(scope="module")
def ctx():
    from rgsplat_amd import runtime
    c = runtime.Context(0)
    yield c
    c.close()


def with_alpha(rgb):
    return np.concatenate([rgb, np.ones(rgb.shape[:2] + (1,), np.float32)], -1)


def workload(name="c0_10k_256", n=None, sizes=None):
    wl = S.scaled(S.WORKLOADS[name], n=n, sizes=sizes)
    g = wl.scene()
    cams = wl.cameras()
    tg = wl.target_scene()
    targets = [with_alpha(oracle.render_color(tg, c)) for c in cams]
    return g, cams, targets


def group_rel_err(a, b):
    out = {}
    for k, sl in GROUPS.items():
        x, y = a[:, sl].astype(np.float64), b[:, sl].astype(np.float64)
        den = np.linalg.norm(y)
        out[k] = np.linalg.norm(x - y) / den if den > 0 else np.linalg.norm(x - y)
    return out


def assert_grads_close(got, ref, tol=1e-3):
    errs = group_rel_err(got, ref)
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, f"gradient groups above {tol} relative: {bad} (all: {errs})"
    # zero where the oracle culls (no contribution at all)
    culled = np.all(ref == 0, axis=1)
    assert np.all(got[culled] == 0)


# ------------------------------------------------------------------ expf ---
def test_device_expf_matches_glibc(ctx):
    """Bit-exact against glibc expf: every 251st bit pattern of all 2^32 and
    a dense sweep of the raster's domain (power in [-30, 0])."""
    bits = np.arange(0, 2 ** 32, 251, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[~np.isnan(x)]
    dense = np.linspace(-30.0, 0.0, 2_000_001, dtype=np.float32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, 88.72283, 88.72284, -103.97208, -103.97209, -87.33655,
                        1e-30, -1e-30, 1.4e-45], np.float32)
    for arr in (x, dense, special):
        got = ctx.eval_expf(arr)
        ref = oracle.expf_batch(arr)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
            f"{np.count_nonzero(got.view(np.uint32) != ref.view(np.uint32))} mismatches"
    nan = ctx.eval_expf(np.array([np.nan], np.float32))
    assert np.isnan(nan[0])


# --------------------------------------------------------------- forward ---
# (4128, 4096): 66048 tiles, beyond the 16-bit bin keys of the tile sort
@pytest.mark.parametrize("name,n,size", [("c0_10k_256", None, None), ("c1_100k_800", 30_000, (480, 360)),
                                         ("c0_10k_256", 3000, (100, 75)), ("c0_10k_256", 4000, (4128, 4096))])
def test_binning_and_image_bitexact(ctx, name, n, size):
    g, cams, _ = workload(name, n, [size] if size else None)
    cam = cams[0]
    rgb = ctx.render_color(g, cam)
    r = oracle.Render(g, cam, None, backward=False)
    ref_rgb, ref_t, ref_last = r.image()
    off_ref, gis_ref = r.tiles()
    n_tiles = len(off_ref) - 1
    off, gis = ctx.debug_tiles(n_tiles, len(gis_ref) + 16)
    # binning: bit-exact (same lists, same order)
    assert np.array_equal(off.astype(np.int64), off_ref.astype(np.int64))
    assert np.array_equal(gis, gis_ref)
    # image: bar 1e-4 absolute; expected bit-exact
    assert np.max(np.abs(rgb - ref_rgb)) <= 1e-4
    assert np.array_equal(rgb, ref_rgb)


def test_rasterize_rgba_parity(ctx):
    g, cams, _ = workload("c0_10k_256", 4000, [(200, 120)])
    for bg in [(0.0, 0.0, 0.0), (0.2, 0.5, 0.9)]:
        got = ctx.rasterize(g, cams[0], bg)
        ref = oracle.rasterize_rgba(g, cams[0], bg)
        assert np.max(np.abs(got - ref)) <= 1e-4
        assert np.array_equal(got, ref)


def test_render_with_background(ctx):
    g, cams, _ = workload("c0_10k_256", 2000, [(96, 64)])
    bg = (0.3, 0.1, 0.7)
    got = ctx.render_color(g, cams[0], bg)
    ref = oracle.render_color(g, cams[0], bg)
    assert np.array_equal(got, ref)


def high_opacity(g, frac=0.3, seed=3):
    """A share of the splats given opacity logits in [3.5, 9] (opacity 0.971
    to 0.9999): splats below and above 0.98 and ones that clamp at 0.99 mixed
    in the same tile lists."""
    r = np.random.default_rng(seed)
    g = g.copy()
    sel = r.random(len(g)) < frac
    g[sel, 3] = r.uniform(3.5, 9.0, int(sel.sum())).astype(np.float32)
    return g


def test_high_opacity_splats(ctx):
    """Scenes where many splats reach alpha 0.99 (clamped) or come within the
    exact-exp window of it: the API renders stay bit-exact and the backward
    and the training step match the oracle (same bars as above)."""
    from rgsplat_amd import runtime
    g, cams, targets = workload("c0_10k_256", 4000, [(200, 120)])
    g = high_opacity(g)
    cam = cams[0]
    assert np.array_equal(ctx.render_color(g, cam), oracle.render_color(g, cam))
    bg = (0.2, 0.5, 0.9)
    assert np.array_equal(ctx.rasterize(g, cam, bg), oracle.rasterize_rgba(g, cam, bg))
    loss, grads = ctx.render_and_backward(g, cam, targets[0], (0.0, 0.0, 0.0), 0.2)
    r = oracle.Render(g, cam, targets[0])
    assert abs(loss - r.loss) <= 1e-5 * abs(r.loss)
    assert_grads_close(grads, r.grads(), 1e-3)
    # the training step (hardware exp2, 32x32 bins): image within 1e-5 of
    # the oracle's, loss and gradients to the bars above
    cfg = S.benchmark_train_config()
    c = runtime.Context(0)
    try:
        c.scene_upload(g, capacity=len(g))
        c.set_views(cams, targets)
        step_loss = c.train_step([0], cfg, use_graph=False)
        img = c.step_image(cam.width * cam.height)
        assert np.abs(img - oracle.render_color(g, cam).reshape(-1, 3)).max() <= 1e-5
        ref_loss, ref_g = _oracle_step_grads(g, cams, targets, (0, 0, 0), 0.2)
        assert abs(step_loss[0] - ref_loss[0]) <= 1e-5 * ref_loss[0]
        got = c.adam_download().m1[: len(g)] / np.float32(0.1)
        assert_grads_close(np.concatenate([got, np.zeros((len(g), 1))], 1)[:, :60],
                           np.concatenate([ref_g[:, :59], np.zeros((len(g), 1))], 1), 1e-3)
    finally:
        c.close()


# ------------------------------------------------------------------ loss ---
def test_loss_grad_parity(ctx):
    r = S.Stream(21)
    w, h = 77, 53
    rgb = (r.normal((h, w, 3)) * 0.2 + 0.5).astype(np.float32)
    tg = with_alpha((r.normal((h, w, 3)) * 0.2 + 0.5).astype(np.float32))
    rgb[3, 4, 1] = tg[3, 4, 1]  # a zero residual: Rust signum(+0.0) = +1
    loss, dl = ctx.photometric_loss_grad(rgb, tg, w, h, 0.2)
    loss_ref, dl_ref = oracle.loss_grad(rgb, tg, w, h, 0.2)
    # per-pixel gradient: blur taps in the reference's order with fused
    # multiply-adds and the five SSIM coefficient maps folded into three, so
    # equal to float rounding (bar: gradients within 1e-3 relative)
    d = (dl - dl_ref).astype(np.float64)
    assert np.linalg.norm(d) <= 1e-5 * np.linalg.norm(dl_ref)
    assert np.abs(d).max() <= 1e-4 * np.abs(dl_ref).max()
    # at the zero residual the L1 term is +(1 - lambda) / M (Rust signum(+0.0)
    # = +1), which dominates the SSIM term there
    zi = (3 * w + 4) * 3 + 1
    assert np.sign(dl.reshape(-1)[zi]) == np.sign(dl_ref.reshape(-1)[zi])
    # the loss scalar is a sum over w*h*3 terms: the reference sums serially
    # in f32, the GPU per tile and then in f64
    assert abs(loss - loss_ref) <= 1e-5 * abs(loss_ref)


@pytest.mark.parametrize("w,h", [(333, 400), (1, 170), (170, 1), (54, 155), (55, 156)])
def test_loss_grad_parity_unit_edges(ctx, w, h):
    """The loss sweeps 54-column x 155-row units eight rows per step: images
    spanning several units, one pixel wide or tall, and on and just past the
    unit size, against the oracle (same bars as test_loss_grad_parity)."""
    r = S.Stream(23)
    rgb = np.clip(r.normal((h, w, 3)) * 0.2 + 0.5, 0, 1).astype(np.float32)
    tg = with_alpha(np.clip(rgb + r.normal((h, w, 3)) * 0.05, 0, 1).astype(np.float32))
    loss, dl = ctx.photometric_loss_grad(rgb, tg, w, h, 0.2)
    # the reference's loss is a serial f32 sum (drifting by ~1e-5 relative
    # at 400k terms); exact: the reference's own f32 terms added in double
    _, loss_exact, dl_ref = oracle.loss_grad_exact(rgb, tg, w, h, 0.2)
    d = (dl - dl_ref).astype(np.float64)
    assert np.linalg.norm(d) <= 1e-5 * np.linalg.norm(dl_ref)
    assert np.abs(d).max() <= 1e-4 * np.abs(dl_ref).max()
    assert abs(loss - loss_exact) <= 1e-6 * abs(loss_exact)


def test_step_loss_uses_target_statistics():
    """The training step reads the registered targets' blurred statistics
    (mu_y and E[y^2], made once per set_views) instead of blurring the targets
    every step; the values are the same fused multiply-add chains, so each
    view's loss equals, bit for bit, rgs_loss_grad's (which blurs the target
    itself) on the step's own image. Registering other targets replaces the
    statistics."""
    from rgsplat_amd import runtime
    g, cams, targets = workload("c0_10k_256", 20_000, [(333, 400), (256, 192), (55, 156)])
    cfg = S.benchmark_train_config()
    frozen = dict(cfg, lr_position=0.0, lr_rotation=0.0, lr_scale=0.0, lr_opacity=0.0, lr_sh_dc=0.0,
                  lr_sh_rest=0.0)
    npx = sum(c.width * c.height for c in cams)
    c = runtime.Context(0)
    try:
        c.scene_upload(g)
        for round_ in range(2):
            if round_ == 1:  # other targets: the statistics must follow
                r = S.Stream(5)
                targets = [with_alpha(np.clip(t[..., :3] + r.normal(t[..., :3].shape) * 0.1, 0, 1)
                                      .astype(np.float32)) for t in targets]
            c.set_views(cams, targets)
            loss = c.train_step([0, 1, 2], frozen, use_graph=False)
            img = c.step_image(npx)
            off = 0
            for v, cam in enumerate(cams):
                n = cam.width * cam.height
                api_loss, _ = c.photometric_loss_grad(img[off:off + n], targets[v], cam.width, cam.height, 0.2)
                assert np.float32(api_loss) == loss[v], (round_, v, api_loss, loss[v])
                off += n
    finally:
        c.close()


# -------------------------------------------------------------- backward ---
@pytest.mark.parametrize("bg,lam", [((0.0, 0.0, 0.0), 0.2), ((0.1, 0.4, 0.2), 0.5)])
def test_render_and_backward_parity(ctx, bg, lam):
    g, cams, targets = workload()
    loss, grads = ctx.render_and_backward(g, cams[0], targets[0], bg, lam)
    r = oracle.Render(g, cams[0], targets[0], bg=bg, lam=lam)
    assert abs(loss - r.loss) <= 1e-5 * abs(r.loss)
    assert_grads_close(grads, r.grads(), 1e-3)


def test_backward_against_fp64_finite_differences(ctx):
    """GPU analytic gradients of a 3-Gaussian scene against fp64 central
    differences of the fp64 oracle (1e-3 relative per group)."""
    from tests.test_oracle_render import _fd_scene
    gs, cam, tgt = _fd_scene()
    bg = (0.1, 0.05, 0.2)
    _, grads = ctx.render_and_backward(gs.astype(np.float32), cam, tgt.astype(np.float32), bg, 0.2)

    def loss(x):
        return oracle.Render(x, cam, tgt, bg=bg, lam=0.2, f64=True).loss
    eps = 1e-6
    num = np.zeros((len(gs), 60))
    gmap = {0: 0, 1: 1, 2: 2, 3: 10, 8: 7, 9: 8, 10: 9}
    gmap.update({11 + k: 11 + k for k in range(48)})
    for i in range(len(gs)):
        for off, goff in gmap.items():
            gp, gm = gs.copy(), gs.copy()
            gp[i, off] += eps
            gm[i, off] -= eps
            num[i, goff] = (loss(gp) - loss(gm)) / (2 * eps)
    for k in ["position", "log_scale", "opacity", "sh"]:
        sl = GROUPS[k]
        err = np.linalg.norm(grads[:, sl] - num[:, sl]) / np.linalg.norm(num[:, sl])
        assert err <= 1e-3, (k, err)


# ------------------------------------------------------------- optimizer ---
def test_adam_step_bitexact(ctx):
    from rgsplat_amd.runtime import AdamState
    r = S.Stream(9)
    n = 5000
    g = S.make_scene(n, 4)
    g_ref = g.copy()
    cfg = S.benchmark_train_config()
    st = AdamState(n)
    st_ref = oracle.AdamState(n)
    for _ in range(3):
        grads = np.zeros((n, 60), np.float32)
        grads[:, :59] = r.normal((n, 59), std=1e-3)
        grads[:, 59] = np.abs(r.normal((n,)))
        grads[::7] = 0.0
        ctx.adam_step(g, grads, st, cfg)
        oracle.adam_step(g_ref, grads, st_ref, cfg)
    assert st.t == st_ref.t == 3
    assert np.array_equal(g, g_ref)
    assert np.array_equal(st.m1[:n], st_ref.m1[:n])
    assert np.array_equal(st.m2[:n], st_ref.m2[:n])
    assert np.array_equal(st.grad_accum[:n], st_ref.accum[:n])
    assert np.array_equal(st.grad_denom[:n], st_ref.denom[:n])


def ulp_diff(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    return np.abs(ai - bi)


def test_densify_and_prune_parity(ctx):
    from rgsplat_amd.runtime import AdamState
    n = 3000
    g = S.make_scene(n, 17, log_scale_mean=-3.0, log_scale_std=1.0)
    g[:, 3] = S.Stream(1).normal((n,), std=3.0)
    extent = oracle.scene_extent(g)
    cfg = S.train_config(max_gaussians=n + 700)
    rr = S.Stream(2)
    acc = np.abs(rr.normal((n,), std=4e-4)).astype(np.float32)
    st = AdamState(n, capacity=3 * n)
    st.grad_accum[:n] = acc
    st.grad_denom[:n] = 2
    st.m1[:n] = 1.0
    st_ref = oracle.AdamState(n, cap=3 * n)
    st_ref.accum[:n] = acc
    st_ref.denom[:n] = 2
    st_ref.m1[:n] = 1.0
    buf = np.zeros((3 * n, 59), np.float32)
    buf[:n] = g
    buf_ref = buf.copy()
    nn = ctx.densify_and_prune(buf, n, st, cfg, extent)
    nn_ref = oracle.densify_and_prune(buf_ref, n, st_ref, cfg, extent)
    assert nn == nn_ref
    got, ref = buf[:nn], buf_ref[:nn]
    # log-scale / logit of split children come from logf: the device's
    # f64-log-rounded-once may differ from glibc logf by 1 ulp; all else exact
    d = ulp_diff(got, ref)
    assert d.max() <= 1
    exact_cols = [c for c in range(59) if c not in (3, 8)]
    assert np.array_equal(got[:, exact_cols], ref[:, exact_cols])
    assert st.rows == st_ref.rows
    assert np.array_equal(st.m1[:st.rows], st_ref.m1[:st.rows])
    assert np.all(st.grad_accum[:st.rows] == 0) and np.all(st.grad_denom[:st.rows] == 0)
    # capacity is respected
    small = np.zeros((n + 1, 59), np.float32)
    small[:n] = g
    st2 = AdamState(n, capacity=n + 1)
    st2.grad_accum[:n] = acc
    st2.grad_denom[:n] = 2
    from rgsplat_amd.runtime import RgsError
    if nn_ref > n + 1:
        with pytest.raises(RgsError):
            ctx.densify_and_prune(small, n, st2, cfg, extent)


def test_reset_opacities_parity(ctx):
    g = S.make_scene(20000, 3)
    ref = g.copy()
    ctx.reset_opacities(g)
    oracle.reset_opacities(ref)
    assert ulp_diff(g[:, 3], ref[:, 3]).max() <= 1
    assert np.array_equal(np.delete(g, 3, axis=1), np.delete(ref, 3, axis=1))


# ----------------------------------------------------- device-resident step ---
def _oracle_step_grads(g, cams, targets, bg, lam):
    tot = np.zeros((len(g), 60), np.float64)
    losses = []
    for c, t in zip(cams, targets):
        l, gr = oracle.render_and_backward(g, c, t, bg, lam)
        tot += gr
        losses.append(l)
    return np.array(losses), tot


@pytest.mark.parametrize("use_graph", [False, True])
def test_train_step_single_view(ctx, use_graph):
    from rgsplat_amd import runtime
    g, cams, targets = workload()
    cfg = S.benchmark_train_config()
    c = runtime.Context(0)
    c.scene_upload(g, capacity=len(g) + 100)
    c.set_views(cams, targets)
    loss = c.train_step([0], cfg, use_graph=use_graph)
    ref_loss, ref_g = _oracle_step_grads(g, cams, targets, (0, 0, 0), 0.2)
    assert abs(loss[0] - ref_loss[0]) <= 1e-5 * ref_loss[0]
    st = c.adam_download()
    # after one step m1 = (1 - beta1) g
    got = st.m1[: len(g)] / np.float32(0.1)
    assert_grads_close(np.concatenate([got, np.zeros((len(g), 1))], 1)[:, :60],
                       np.concatenate([ref_g[:, :59], np.zeros((len(g), 1))], 1), 1e-3)
    assert np.all(st.grad_denom[: len(g)] == 1)
    # the screen-gradient norms: the same bar as the gradients (relative norm
    # error; single small entries are sums with cancellation whose rounding
    # depends on the summation order)
    ga = st.grad_accum[: len(g)].astype(np.float64)
    assert np.linalg.norm(ga - ref_g[:, 59]) <= 1e-3 * np.linalg.norm(ref_g[:, 59])
    # the parameters moved by Adam's first step: |dp| <= lr
    g1 = c.scene_download()
    lr = np.array([cfg["lr_position"]] * 3 + [cfg["lr_opacity"]])
    dp = np.abs(g1[:, [0, 1, 2, 3]] - g[:, [0, 1, 2, 3]])
    # |update| <= lr, up to the f32 rounding of p - update (2 ulp of p)
    assert np.all(dp <= lr * 1.0001 + 2 * np.spacing(np.abs(g[:, [0, 1, 2, 3]])))
    # further steps: graph replay and eager agree on the loss trajectory
    for _ in range(3):
        c.train_step([0], cfg, use_graph=use_graph)
    c.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_train_step_background_with_emptied_tiles(ctx, use_graph):
    """A background colour on a scene whose tight binning empties tiles that
    the reference composites (rgb = bg there, 0 for tiles it never bins):
    loss and gradients of the step against the oracle."""
    from rgsplat_amd import runtime
    from tests.test_reach_cull import faint_sparse_scene
    g, cam = faint_sparse_scene()
    tg = [with_alpha(oracle.render_color(S.make_scene(len(g), 32), cam))]
    bg = (0.3, 0.1, 0.7)
    cfg = S.benchmark_train_config()
    c = runtime.Context(0)
    c.scene_upload(g)
    c.set_views([cam], tg)
    loss = c.train_step([0], cfg, bg=bg, use_graph=use_graph)
    ref_loss, ref_g = _oracle_step_grads(g, [cam], tg, bg, 0.2)
    # The loss is a sum of 36,864 similar terms (flat background regions),
    # which the reference adds serially in f32 (4.7e-5 relative off here);
    # the device sums per tile and then in f64. Against the f64 loss of the
    # oracle's image the device loss is exact to float rounding.
    rgb_ref = oracle.Render(g, cam, None, bg=bg, backward=False).image()[0]
    loss64, _ = oracle.loss_grad(rgb_ref, tg[0], cam.width, cam.height, 0.2, f64=True)
    assert abs(loss[0] - loss64) <= 1e-6 * loss64
    assert abs(loss[0] - ref_loss[0]) <= 1e-4 * ref_loss[0]
    got = c.adam_download().m1[: len(g)] / np.float32(0.1)
    assert_grads_close(np.concatenate([got, np.zeros((len(g), 1))], 1)[:, :60],
                       np.concatenate([ref_g[:, :59], np.zeros((len(g), 1))], 1), 1e-3)
    c.close()


def test_train_step_images(ctx):
    """The training step's forward against the oracle's image: both binnings
    (16x16 tiles, 32x32 bins), with and without a depth cut, on a dense scene
    with a background colour. The step evaluates alpha with the hardware exp2
    (exact expf near the alpha thresholds), so images agree to float rounding
    (bar: 1e-4 absolute); the four variants agree bit for bit."""
    from rgsplat_amd import runtime
    g, cams, targets = workload("c3_1m_1080p_4views", 60_000, [(256, 192), (224, 160)])
    cfg = S.benchmark_train_config()
    bg = (0.2, 0.3, 0.1)
    ref = np.concatenate([oracle.render_color(g, c, bg).reshape(-1, 3) for c in cams])
    # zero learning rates: the second step renders the scene of the first
    frozen = dict(cfg, lr_position=0.0, lr_rotation=0.0, lr_scale=0.0, lr_opacity=0.0, lr_sh_dc=0.0,
                  lr_sh_rest=0.0)
    # Fixed cuts: 3.0 leaves most tiles unsaturated in front, 4.0 splits bins
    # into saturated and unsaturated tiles. The adaptive cut acts from the
    # second step on; that step renders the scene after the first step's
    # quaternion renormalisation, so its images are compared with whole-list
    # two-step runs.
    imgs = {1: [], 2: []}
    for shift, per_tile in [(0, 1), (1, 1), (1, 0)]:
        for mode, depth, steps in [(1, 0.0, 1), (2, 3.0, 1), (2, 4.0, 1), (1, 0.0, 2), (0, 0.0, 2)]:
            c = runtime.Context(0)
            c.set_train_binning(shift, per_tile)
            c.set_depth_cut(mode, depth)
            c.scene_upload(g)
            c.set_views(cams, targets)
            for _ in range(steps):
                c.train_step([0, 1], frozen, bg=bg, use_graph=False)
            imgs[steps].append(c.step_image(len(ref)))
            st = c.cut_stats()
            if mode != 1:
                assert st["front_pairs"] < c.step_pairs() or st["flagged_tiles"] > 0
            c.close()
    for group in imgs.values():
        for im in group:
            assert np.abs(im - ref).max() <= 1e-5
            assert np.array_equal(im, group[0])


@pytest.mark.parametrize("use_graph", [False, True])
def test_depth_cut_binning(ctx, use_graph):
    """The training step's depth-cut binning against whole lists: a fixed
    shallow cut (most tiles need their back lists), the adaptive cut and no
    cut give the same loss bit for bit and gradients within the atomics'
    summation order; the gradients match the oracle."""
    from rgsplat_amd import runtime
    # dense enough that whole tiles saturate (the adaptive cut needs them)
    g, cams, targets = workload("c3_1m_1080p_4views", 60_000, [(256, 192), (224, 160)])
    cfg = S.benchmark_train_config()
    bg = (0.2, 0.3, 0.1)

    def run(mode, depth, steps, step_cfg):
        c = runtime.Context(0)
        c.set_depth_cut(mode, depth)
        c.scene_upload(g)
        c.set_views(cams, targets)
        out = []
        for _ in range(steps):
            loss = c.train_step([0, 1], step_cfg, bg=bg, use_graph=use_graph)
            out.append((loss, c.adam_download().m1[: len(g)].astype(np.float64), c.cut_stats(), c.step_pairs()))
        c.close()
        return out

    whole = run(1, 0.0, 1, cfg)[0]
    fixed = run(2, 3.0, 1, cfg)[0]
    assert whole[2]["back_pairs"] == 0 and whole[2]["flagged_tiles"] == 0
    assert fixed[2]["flagged_tiles"] > 0 and fixed[2]["back_pairs"] > 0
    assert fixed[2]["front_pairs"] + fixed[2]["back_pairs"] <= whole[3]
    assert np.array_equal(fixed[0], whole[0])
    gw = np.linalg.norm(whole[1])
    assert np.linalg.norm(fixed[1] - whole[1]) <= 1e-5 * gw
    # adaptive: no cut on the first step, the first step's saturation depth on
    # the second. Zero learning rates make the scene after the first step
    # independent of its gradients (only the quaternion renormalisation
    # moves it), so two-step runs with and without the cut see the same scene.
    frozen = dict(cfg, lr_position=0.0, lr_rotation=0.0, lr_scale=0.0, lr_opacity=0.0, lr_sh_dc=0.0,
                  lr_sh_rest=0.0)
    s1, s2 = run(0, 0.0, 2, frozen)
    w1, w2 = run(1, 0.0, 2, frozen)
    assert np.isfinite(s1[2]["next_cut"]) and s1[2]["back_pairs"] == 0
    assert s2[2]["front_pairs"] + s2[2]["back_pairs"] < w2[3]
    assert np.array_equal(s1[0], whole[0]) and np.array_equal(s2[0], w2[0])
    assert np.linalg.norm(s2[1] - w2[1]) <= 1e-5 * np.linalg.norm(w2[1])
    ref_loss, ref_g = _oracle_step_grads(g, cams, targets, bg, 0.2)
    np.testing.assert_allclose(fixed[0], ref_loss, rtol=1e-4)
    got = (fixed[1] / 0.1).astype(np.float32)
    assert_grads_close(np.concatenate([got, np.zeros((len(g), 1))], 1)[:, :60],
                       np.concatenate([ref_g[:, :59], np.zeros((len(g), 1))], 1), 1e-3)


@pytest.mark.parametrize("use_graph", [False, True])
def test_depth_cut_rerun(ctx, use_graph):
    """A step run without the back-list stage (the previous step needed no
    back list) that leaves a bin unsaturated is re-run with the stage. On a
    scene where every bin saturates, the adaptive cut needs no back list;
    a shallow fixed cut on the next step then flags bins. The re-run step's
    losses, gradients and images equal a whole-list run's."""
    from rgsplat_amd import runtime
    g, cams, targets = workload("c3_1m_1080p_4views", 200_000, [(128, 96), (96, 64)])
    cfg = S.benchmark_train_config()
    frozen = dict(cfg, lr_position=0.0, lr_rotation=0.0, lr_scale=0.0, lr_opacity=0.0, lr_sh_dc=0.0,
                  lr_sh_rest=0.0)
    npx = sum(c.width * c.height for c in cams)
    out = []
    for whole in (False, True):
        c = runtime.Context(0)
        c.set_depth_cut(1 if whole else 0)
        c.scene_upload(g)
        c.set_views(cams, targets)
        c.train_step([0, 1], frozen, use_graph=use_graph)
        c.train_step([0, 1], frozen, use_graph=use_graph)
        st2 = c.cut_stats()
        if not whole:
            c.set_depth_cut(2, 2.0)
        loss = c.train_step([0, 1], frozen, use_graph=use_graph)
        out.append((loss, c.adam_download().m1[: len(g)].astype(np.float64), c.step_image(npx), st2,
                    c.cut_stats()))
        c.close()
    (la, ma, ia, s2a, s3a), (lw, mw, iw, s2w, s3w) = out
    # the adaptive cut's step needed no back list and ran as captured without
    # the stage; the shallow cut's step was re-run once with it
    assert s2a["flagged_tiles"] == 0 and np.isfinite(s2a["next_cut"])
    assert s3a["reruns"] == s2a["reruns"] + 1 and s3a["flagged_tiles"] > 0 and s3a["back_pairs"] > 0
    assert s3w["reruns"] == 0
    assert np.array_equal(la, lw) and np.array_equal(ia, iw)
    assert np.linalg.norm(ma - mw) <= 1e-5 * np.linalg.norm(mw)


def test_adaptive_cut_pause():
    """The single-view 100k-Gaussian workload at 800x800 leaves tiles that
    never saturate, so the adaptive cut needs its back-list stage on every
    step while leaving few entries behind it: after the first such step the
    cut pauses (whole lists, no back lists, nothing flagged) for 256 steps,
    then resumes with the back-list stage on its first steps, so no step is
    re-run. Losses equal a whole-list run's throughout (frozen scene)."""
    from rgsplat_amd import runtime
    g, cams, targets = workload("c1_100k_800")
    cfg = S.benchmark_train_config()
    frozen = dict(cfg, lr_position=0.0, lr_rotation=0.0, lr_scale=0.0, lr_opacity=0.0, lr_sh_dc=0.0,
                  lr_sh_rest=0.0)
    runs = {}
    for mode in (0, 1):
        c = runtime.Context(0)
        c.set_depth_cut(mode)
        c.scene_upload(g)
        c.set_views(cams, targets)
        rec = []
        for k in range(261):
            loss = c.train_step([0], frozen, use_graph=True)
            if k in (0, 1, 2, 3, 257, 258, 259, 260):
                rec.append((k, loss, c.cut_stats()))
        runs[mode] = rec
        c.close()
    st = {k: s for k, _, s in runs[0]}
    assert st[1]["flagged_tiles"] > 0  # the cut's first step needed back lists
    for k in (2, 3, 257, 260):  # paused (257: the pause ends after it)
        assert st[k]["flagged_tiles"] == 0 and st[k]["back_pairs"] == 0 and not np.isfinite(st[k]["next_cut"]), k
    # resumed: no cut on step 258, its cut on 259, which ran with the back
    # lists (no re-run) and paused the cut again
    assert st[258]["flagged_tiles"] == 0
    assert st[259]["flagged_tiles"] > 0 and st[259]["reruns"] == st[1]["reruns"]
    for (k, la, _), (_, lw, _) in zip(runs[0], runs[1]):
        assert np.array_equal(la, lw), k


def test_graph_replay_matches_eager(ctx):
    from rgsplat_amd import runtime
    g, cams, targets = workload("c0_10k_256", 5000, [(128, 96), (160, 128)])
    cfg = S.benchmark_train_config()
    out = []
    for use_graph in (False, True):
        c = runtime.Context(0)
        c.scene_upload(g)
        c.set_views(cams, targets)
        out.append([c.train_step([0, 1], cfg, use_graph=use_graph) for _ in range(4)])
        c.close()
    np.testing.assert_allclose(np.array(out[0]), np.array(out[1]), rtol=1e-4)


def test_batched_uneven_views(ctx):
    """Several views of different sizes in one launch sequence: gradients are
    the sum of the per-view CPU gradients."""
    from rgsplat_amd import runtime
    g, cams, targets = workload("c2_500k_8views", 20_000, [(160, 120), (240, 136), (320, 180)])
    cfg = S.benchmark_train_config()
    c = runtime.Context(0)
    c.scene_upload(g)
    c.set_views(cams, targets)
    loss = c.train_step([0, 1, 2], cfg, use_graph=True)
    ref_loss, ref_g = _oracle_step_grads(g, cams, targets, (0, 0, 0), 0.2)
    np.testing.assert_allclose(loss, ref_loss, rtol=1e-5)
    st = c.adam_download()
    got = st.m1[: len(g)] / np.float32(0.1)
    errs = group_rel_err(np.concatenate([got, np.zeros((len(g), 1))], 1),
                         np.concatenate([ref_g[:, :59], np.zeros((len(g), 1))], 1))
    for k in ["position", "rotation", "log_scale", "opacity", "sh"]:
        assert errs[k] <= 1e-3, (k, errs)
    assert np.all(st.grad_denom[: len(g)] == 3)
    # a different subset and order of views replays correctly too
    c.train_step([2, 0], cfg, use_graph=True)
    c.close()


def test_scene_densify_matches_api(ctx):
    from rgsplat_amd import runtime
    n = 4000
    g = S.make_scene(n, 23, log_scale_mean=-3.0, log_scale_std=1.0)
    g[:, 3] = S.Stream(4).normal((n,), std=3.0)
    cams = S.make_cameras(3003, [(96, 96)], focal=100.0)
    tg = [with_alpha(oracle.render_color(g, cams[0]))]
    cfg = S.benchmark_train_config()
    cfg["densify_grad_threshold"] = 1e-7
    extent = runtime.scene_extent(g)
    c = runtime.Context(0)
    c.scene_upload(g, capacity=3 * n)
    c.set_views(cams, tg)
    c.train_step([0], cfg, use_graph=False)
    before = c.scene_download()
    st = c.adam_download()
    c.scene_densify(cfg, extent)
    after = c.scene_download()
    # same operation through the reference-layout entry point
    buf = np.zeros((3 * n, 59), np.float32)
    buf[:n] = before
    st_api = runtime.AdamState(n, capacity=3 * n)
    st_api.grad_accum[:n] = st.grad_accum[:n]
    st_api.grad_denom[:n] = st.grad_denom[:n]
    nn = ctx.densify_and_prune(buf, n, st_api, cfg, extent)
    assert len(after) == nn
    assert np.array_equal(after, buf[:nn])
    c.close()


# ------------------------------------------------------------- edge cases ---
def test_edge_cases(ctx):
    from rgsplat_amd.runtime import RgsError
    cam = S.identity_camera(40, 24)
    # empty scene
    img = ctx.render_color(np.zeros((0, 59), np.float32), cam)
    assert img.shape == (24, 40, 3) and np.all(img == 0)
    # everything behind the camera: all culled, zero gradients
    g = S.make_scene(50, 1)
    g[:, 2] = -5.0
    loss, gr = ctx.render_and_backward(g, cam, np.ones((24, 40, 4), np.float32))
    assert np.all(gr == 0)
    ref_loss, _ = oracle.render_and_backward(g, cam, np.ones((24, 40, 4), np.float32))
    assert abs(loss - ref_loss) <= 1e-6
    # one Gaussian covering the whole image (a long tile list per tile)
    big = np.zeros((1, 59), np.float32)
    big[0, 2] = 1.0
    big[0, 3] = 3.0
    big[0, 4] = 1.0
    big[0, 8:11] = 1.0
    img = ctx.render_color(big, cam)
    assert np.array_equal(img, oracle.render_color(big, cam))
    # NaN depth and a singular camera are errors, not crashes
    bad = S.make_scene(10, 2)
    bad[3, 0:3] = np.nan
    bad[3, 2] = np.nan
    with pytest.raises(RgsError):
        ctx.render_color(bad, S.identity_camera(16, 16))
    sing = S.identity_camera(16, 16)
    sing.c2w = np.zeros((4, 4), np.float32)
    with pytest.raises(RgsError):
        ctx.render_color(S.make_scene(4, 1), sing)
    # the context still works afterwards
    g, cams, _ = workload("c0_10k_256", 500, [(33, 17)])
    assert np.array_equal(ctx.render_color(g, cams[0]), oracle.render_color(g, cams[0]))


def test_training_reduces_loss_on_gpu(ctx):
    """tests/training_reduces_loss.rs:38-84 through the HIP entry points."""
    from rgsplat_amd.runtime import AdamState
    w = h = 32
    cam = S.identity_camera(w, h)
    target = np.zeros((h, w, 4), np.float32)
    target[..., :3] = [0.2, 0.55, 0.5]
    target[..., 3] = 1
    gs = []
    for x, y in [(-0.3, -0.3), (0.3, -0.3), (-0.3, 0.3), (0.3, 0.3), (0.0, 0.0)]:
        g = np.zeros(59, np.float32)
        g[0:3] = [x, y, 2.0]
        g[4] = 1.0
        g[8:11] = -1.3
        gs.append(g)
    gs = np.array(gs, np.float32)
    cfg = S.train_config()
    st = AdamState(len(gs))
    initial = ctx.render_and_backward(gs, cam, target)[0]
    last = initial
    for _ in range(200):
        last, grads = ctx.render_and_backward(gs, cam, target)
        ctx.adam_step(gs, grads, st, cfg)
    assert last < initial * 0.5
    img = ctx.render_color(gs, cam)
    assert np.all(np.abs(img[h // 2, w // 2] - target[0, 0, :3]) < 0.15)


def test_full_size_properties(ctx):
    """1M Gaussians at 1080p through the reference-binning API render: the
    whole-image tile lists (132M pairs) are checked by size-independent
    properties (every list in depth order, ties by index, pair counts that
    agree); the images, losses and gradients of the training step at this
    size are compared with the oracle in tests/test_gpu_fullsize.py."""
    from rgsplat_amd import runtime
    wl = S.WORKLOADS["c3_1m_1080p_4views"]
    g = wl.scene()
    cam = wl.cameras()[0]
    rgb = ctx.render_color(g, cam)
    assert np.all(np.isfinite(rgb)) and rgb.max() <= 1.0 + 1e-6 and rgb.min() >= 0.0
    K = ctx.step_pairs()
    off, gis = ctx.debug_tiles(120 * 68, K + 16)
    assert len(gis) == K and off[-1] == K
    # every tile list is in depth order (ties by index): depth is the f32
    # p_cam.z in the reference's operation order ((v0 x + v1 y) + v2 z) + v3
    view = oracle.view_matrix(cam)
    v = view[2]
    depth = ((v[0] * g[:, 0] + v[1] * g[:, 1]) + v[2] * g[:, 2]) + v[3] * np.float32(1.0)
    assert depth.dtype == np.float32
    for t in range(0, 120 * 68, 97):
        L = gis[off[t]:off[t + 1]].astype(np.int64)
        if len(L) > 1:
            d = depth[L]
            assert np.all((d[1:] > d[:-1]) | ((d[1:] == d[:-1]) & (L[1:] > L[:-1])))
    # spot-check one region against the oracle on the Gaussians touching it
    sub = np.unique(gis[off[4000]:off[4001]])
    assert len(sub) > 0


def test_raster_params_and_autotune(ctx):
    """The raster work decomposition changes no image and gradients only by
    summation order; the autotuner leaves the scene untouched and picks one
    of its candidates."""
    from rgsplat_amd import runtime
    g, cams, targets = workload("c2_500k_8views", 20_000, [(240, 136), (320, 180)])
    cfg = S.benchmark_train_config()
    ref_loss, ref_g = _oracle_step_grads(g, cams, targets, (0, 0, 0), 0.2)
    c = runtime.Context(0)
    assert c.raster_params() == (8, 2)
    for st, sl in [(1, 0), (1024, 64), (3, 5), (0, 2)]:
        c.set_raster_params(st, sl)
        assert c.raster_params() == (st, sl)
        img = c.render_color(g, cams[1])
        assert np.array_equal(img, oracle.render_color(g, cams[1]))  # bar 1e-4 abs; expected bit-exact
        _, gr = c.render_and_backward(g, cams[0], targets[0])
        _, ref0 = oracle.render_and_backward(g, cams[0], targets[0])
        assert_grads_close(gr, ref0, 1e-3)
    for bad in [(-1, 2), (8, 65), (2000, 2)]:
        with pytest.raises(runtime.RgsError):
            c.set_raster_params(*bad)
    assert c.raster_batch() == 64
    for bad in [0, 16, 96, 256]:
        with pytest.raises(runtime.RgsError):
            c.set_raster_batch(bad)
    c.scene_upload(g)
    c.set_views(cams, targets)
    best = c.autotune_raster([0, 1])
    # a candidate, or the setting it started from (3, 5) when nothing beat it
    assert best["super_tile"] in (0, 2, 4, 8, 16, 1024, 3) and best["sparse_lanes"] in (0, 2, 4, 8, 16, 32, 5)
    assert best["batch"] in (32, 64, 128) and c.raster_batch() == best["batch"]
    assert best["queue"] in (0, 1, 2, 3) and c.raster_queue() == best["queue"]
    assert best["raster_ms"] > 0
    assert c.raster_params() == (best["super_tile"], best["sparse_lanes"])
    assert np.array_equal(c.scene_download(), g)
    loss = c.train_step([0, 1], cfg, use_graph=True)
    np.testing.assert_allclose(loss, ref_loss, rtol=1e-5)
    st = c.adam_download()
    got = st.m1[: len(g)] / np.float32(0.1)
    errs = group_rel_err(np.concatenate([got, np.zeros((len(g), 1))], 1),
                         np.concatenate([ref_g[:, :59], np.zeros((len(g), 1))], 1))
    for k in ["position", "rotation", "log_scale", "opacity", "sh"]:
        assert errs[k] <= 1e-3, (k, errs)
    c.close()


def test_hbm_copy_probe(ctx):
    """The streaming-copy probe behind the memory-bound stages' frac_of_copy:
    a plausible rate (below the 8 TB/s peak, with margin for the clock) and
    its argument checks."""
    from rgsplat_amd import runtime
    gbs = ctx.hbm_copy_probe(64 << 20, 3)
    assert 100.0 < gbs < 9000.0, gbs
    for nbytes, reps in [(8, 3), (1 << 20, 0)]:
        with pytest.raises(runtime.RgsError):
            ctx.hbm_copy_probe(nbytes, reps)


def test_create_on_a_missing_device():
    """rgs_create on a device id the machine does not have: RGS_E_ARG and a
    message, no context; the current device (-1) and device 0 still work."""
    import ctypes
    from rgsplat_amd import runtime
    L = runtime.lib()
    h = ctypes.c_void_p()
    assert L.rgs_create(4096, ctypes.byref(h)) == 1 and not h.value
    assert b"out of range" in L.rgs_last_error()
    with pytest.raises(runtime.RgsError):
        runtime.Context(4096)
    for dev in (-1, 0):
        c = runtime.Context(dev)
        g, cams, _ = workload("c0_10k_256", 200, [(40, 24)])
        assert np.array_equal(c.render_color(g, cams[0]), oracle.render_color(g, cams[0]))
        c.close()


def test_raster_work_counts():
    """The training step's raster kernels counting their own work: the
    backward revisits exactly the (pixel, splat) pairs the forward
    composited (the same alpha decisions), lane evaluations lie between the
    composites and the reference's evaluated pairs on its own lists, and the
    counting pass changes no image."""
    from rgsplat_amd import runtime
    g, cams, targets = workload("c3_1m_1080p_4views", 60_000, [(256, 192), (224, 160)])
    cfg = S.benchmark_train_config()
    c = runtime.Context(0)
    c.set_depth_cut(1)
    c.scene_upload(g)
    c.set_views(cams, targets)
    g0 = c.scene_download()
    c.train_step([0, 1], cfg, use_graph=True)
    npx = sum(cm.width * cm.height for cm in cams)
    img = c.step_image(npx)
    k = c.raster_work_counts()
    assert np.array_equal(c.step_image(npx), img)
    assert k["fwd_composites"] == k["bwd_contributions"] > 0
    assert k["fwd_composites"] <= k["fwd_lane_evals"] <= 64 * k["fwd_wave_visits"]
    assert k["bwd_contributions"] <= k["bwd_lane_evals"] <= 64 * k["bwd_wave_visits"]
    ref_pairs = sum(oracle.Render(g0, cm, None, backward=False).pairs_fwd for cm in cams)
    assert k["fwd_lane_evals"] <= ref_pairs
    # a step that resumed tiles on back lists is refused, not miscounted
    c.set_depth_cut(2, 2.0)
    c.train_step([0, 1], cfg, use_graph=False)
    if c.cut_stats()["back_pairs"]:
        with pytest.raises(runtime.RgsError):
            c.raster_work_counts()
    c.close()


@pytest.mark.parametrize("fwd_tile_blocks,cut", [(1, None), (0, None), (1, 3.0)])
def test_raster_batch_sizes(fwd_tile_blocks, cut):
    """The binned training raster stages 32, 64 or 128 splats per batch
    (128 as two halves of 64): the splats are composited in the same order,
    so images and losses are bit-identical, and gradients differ only by the
    order of float summation; with a fixed depth cut the resumed back lists
    take the same path. Against the oracle as well."""
    from rgsplat_amd import runtime
    g, cams, targets = workload("c2_500k_8views", 30_000, [(160, 120), (240, 136), (320, 180)])
    cfg = S.benchmark_train_config()
    npx = sum(c.width * c.height for c in cams)
    out = {}
    for fb in (64, 32, 128):
        c = runtime.Context(0)
        c.set_train_binning(1, fwd_tile_blocks)
        if cut is not None:
            c.set_depth_cut(2, cut)
        c.set_raster_batch(fb)
        c.scene_upload(g)
        c.set_views(cams, targets)
        c.set_grad_capture(True)
        loss = c.train_step([0, 1, 2], cfg, use_graph=True)
        out[fb] = (loss, c.step_image(npx), c.step_grads(), c.cut_stats())
        c.close()
    for fb in (32, 128):
        assert np.array_equal(out[fb][0], out[64][0]) and np.array_equal(out[fb][1], out[64][1]), fb
        gs, gl = out[64][2].astype(np.float64), out[fb][2].astype(np.float64)
        assert np.linalg.norm(gl - gs) <= 1e-5 * np.linalg.norm(gs), fb
    ref_loss, ref_g = _oracle_step_grads(g, cams, targets, (0, 0, 0), 0.2)
    assert_grads_close(out[128][2], ref_g.astype(np.float32), 1e-3)


@pytest.mark.parametrize("fwd_tile_blocks,super_tile,cut", [(1, 8, None), (0, 4, None), (1, 0, None), (0, 0, 3.0)])
def test_raster_queue_modes(fwd_tile_blocks, super_tile, cut):
    """The persistent training raster (bins from per-XCD queues; bit 0 the
    forward, bit 1 the backward) against one workgroup per bin over two
    graph-replayed steps of three uneven views, under the static schedule
    and the longest-first one (strided queue shares), with and without
    resumed back lists: every bin is rasterised exactly once, so images and
    losses are bit-identical and gradients differ only by the atomics'
    summation order; the gradients against the oracle as well."""
    from rgsplat_amd import runtime
    g, cams, targets = workload("c2_500k_8views", 30_000, [(160, 120), (240, 136), (320, 180)])
    cfg = S.benchmark_train_config()
    npx = sum(c.width * c.height for c in cams)
    out = {}
    for q in (0, 1, 2, 3):
        c = runtime.Context(0)
        c.set_train_binning(1, fwd_tile_blocks)
        c.set_raster_params(super_tile, 2)
        if cut is not None:
            c.set_depth_cut(2, cut)
        c.set_raster_queue(q)
        assert c.raster_queue() == q
        c.scene_upload(g)
        c.set_views(cams, targets)
        c.set_grad_capture(True)
        rec = []
        for _ in range(2):
            loss = c.train_step([0, 1, 2], cfg, use_graph=True)
            rec.append((loss, c.step_image(npx), c.step_grads()))
        out[q] = rec
        c.close()
    for q in (1, 2, 3):
        l0, i0, g0 = out[0][0]
        l1, i1, g1 = out[q][0]
        assert np.array_equal(l1, l0) and np.array_equal(i1, i0), q
        gs, gq = g0.astype(np.float64), g1.astype(np.float64)
        assert np.linalg.norm(gq - gs) <= 1e-5 * np.linalg.norm(gs), q
        # the second step starts from scenes that differ by the first
        # step's summation order only
        np.testing.assert_allclose(out[q][1][0], out[0][1][0], rtol=1e-5)
    c = runtime.Context(0)
    for bad in (-1, 4):
        with pytest.raises(runtime.RgsError):
            c.set_raster_queue(bad)
    c.close()
    ref_loss, ref_g = _oracle_step_grads(g, cams, targets, (0, 0, 0), 0.2)
    assert_grads_close(out[3][0][2], ref_g.astype(np.float32), 1e-3)


@pytest.mark.parametrize("fwd_tile_blocks", [1, 0])
def test_longest_first_schedule(fwd_tile_blocks):
    """The backward's longest-first schedule (super_tile 0: bins in
    descending order of the step's forward walk lengths) against the static
    one over four graph-replayed steps of three uneven views: the same images
    and losses bit for bit on the first step, gradients to the atomics'
    summation order, and the gradients against the oracle."""
    from rgsplat_amd import runtime
    g, cams, targets = workload("c2_500k_8views", 30_000, [(160, 120), (240, 136), (320, 180)])
    cfg = S.benchmark_train_config()
    npx = sum(c.width * c.height for c in cams)
    out = {}
    for st in (8, 0):
        c = runtime.Context(0)
        c.set_train_binning(1, fwd_tile_blocks)
        c.set_raster_params(st, 2)
        c.scene_upload(g)
        c.set_views(cams, targets)
        c.set_grad_capture(True)
        rec = []
        for _ in range(4):
            g_before = c.scene_download()
            loss = c.train_step([0, 1, 2], cfg, use_graph=True)
            rec.append((g_before, loss, c.step_image(npx), c.step_grads()))
        out[st] = rec
        c.close()
    for (g0, l0, i0, gr0), (g1, l1, i1, gr1) in zip(out[8], out[0]):
        assert np.abs(g0 - g1).max() < 1e-5
    # the first step's scenes are equal: bit-identical forward
    assert np.array_equal(out[8][0][1], out[0][0][1]) and np.array_equal(out[8][0][2], out[0][0][2])
    for k in range(4):
        gs, gl = out[8][k][3].astype(np.float64), out[0][k][3].astype(np.float64)
        assert np.linalg.norm(gl - gs) <= 1e-5 * np.linalg.norm(gs), k
    ref_loss, ref_g = _oracle_step_grads(out[0][3][0], cams, targets, (0, 0, 0), 0.2)
    assert_grads_close(out[0][3][3], ref_g.astype(np.float32), 1e-3)
