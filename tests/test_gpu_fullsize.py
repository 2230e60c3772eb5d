"""Full-size parity of the device training step (the timed path) against the
CPU oracle, on the workloads' own sizes:

* C1: 100k Gaussians, SH3, one 800x800 view (fx = fy = 1100);
* C2: 500k Gaussians, 8 views cycling 640x480 / 960x540 / 1280x720 in one
  graph-replayed launch sequence;
* C3: 1M Gaussians, 4 views of 1920x1080 -- the benchmark's own step
  (autotuned raster decomposition, 32x32 bins, the adaptive depth cut from the
  second step on, hipGraph replay).

Every step runs with gradient capture, so the step's summed gradients are
compared with the oracle's per-view render_and_backward (cpu/backward.rs:
107-149) and its fused Adam with oracle.adam_step on those same gradients
(pipeline/optimizer.rs:116-163). Bars are in tests/step_parity.py.
"""
import json

import numpy as np
import pytest

import oracle
from rgsplat_amd import synthetic as S
from tests.step_parity import run_checked_step, with_alpha

pytestmark = pytest.mark.gpu


def setup(name, n=None, sizes=None, autotune=False, capacity=None):
    from rgsplat_amd import runtime
    wl = S.scaled(S.WORKLOADS[name], n=n, sizes=sizes)
    g = wl.scene()
    cams = wl.cameras()
    ctx = runtime.Context(0)
    tscene = wl.target_scene()
    # targets: the target scene through the API render, bit-identical to
    # oracle.render_color (test_binning_and_image_bitexact)
    targets = [with_alpha(ctx.render_color(tscene, c)) for c in cams]
    del tscene
    ctx.scene_upload(g, capacity=capacity or len(g))
    ctx.set_views(cams, targets)
    if autotune:
        ctx.autotune_raster(list(range(len(cams))))
    ctx.set_grad_capture(True)
    return ctx, cams, targets


def report(name, reps):
    print(f"\n[fullsize] {name} " + json.dumps(reps), flush=True)


@pytest.mark.timeout(300)
def test_c1_train_step_full_size():
    """C1: 100k Gaussians at 800x800, three steps; gradients checked on the
    first and third, Adam on all three."""
    cfg = S.benchmark_train_config()
    ctx, cams, targets = setup("c1_100k_800")
    try:
        reps = [run_checked_step(ctx, [0], cams, targets, cfg, check_grads=k != 1) for k in range(3)]
    finally:
        ctx.close()
    report("c1", reps)


@pytest.mark.timeout(300)
def test_c2_batched_uneven_views_full_size():
    """C2: 500k Gaussians, 8 views of three sizes in one step (graph replay),
    per-view losses and summed gradients against 8 oracle runs; two steps."""
    cfg = S.benchmark_train_config()
    ctx, cams, targets = setup("c2_500k_8views", autotune=True)
    ids = list(range(len(cams)))
    try:
        reps = [run_checked_step(ctx, ids, cams, targets, cfg) for _ in range(2)]
        # a different subset and order of the views replays correctly too
        order = [5, 0, 7, 2]
        reps.append(run_checked_step(ctx, order, [cams[i] for i in order], [targets[i] for i in order], cfg))
    finally:
        ctx.close()
    report("c2", reps)


@pytest.mark.timeout(500)
def test_c3_benchmark_step_full_size():
    """C3: the benchmark's step at 1M Gaussians x 4 views of 1920x1080. Step 1
    runs whole lists (no cut yet); step 2 runs the adaptive depth cut on the
    graph variant without the back-list stage, as every timed step does;
    step 3 is checked for Adam only; after 30 more steps (the loss falls
    about tenfold, as in the benchmark's timed steps) one more step is
    checked in full on the trained scene."""
    cfg = S.benchmark_train_config()
    ctx, cams, targets = setup("c3_1m_1080p_4views", autotune=True)
    ids = list(range(len(cams)))
    try:
        r1 = run_checked_step(ctx, ids, cams, targets, cfg)
        c1 = ctx.cut_stats()
        r2 = run_checked_step(ctx, ids, cams, targets, cfg)
        c2 = ctx.cut_stats()
        r3 = run_checked_step(ctx, ids, cams, targets, cfg, check_grads=False)
        for _ in range(30):
            ctx.train_step(ids, cfg, use_graph=True)
        r4 = run_checked_step(ctx, ids, cams, targets, cfg)
        r4["cut"] = ctx.cut_stats()
    finally:
        ctx.close()
    assert np.isfinite(c1["next_cut"])
    # the second step sorted only the front pairs of the cut
    assert c2["front_pairs"] > 0
    r2["cut"] = c2
    assert r4["loss"][0] < 0.5 * r1["loss"][0]
    report("c3", [r1, r2, r3, r4])


def _ulps(a, b):
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))


@pytest.mark.timeout(400)
def test_c3_densify_full_size():
    """One densify_and_prune (optimizer.rs:206-272) of the 1M-Gaussian scene
    after three 4-view steps: the device scene's densify, the reference-API
    entry point and the oracle from the same pre-densify state. Decisions,
    order, counts and every field but the logf ones are bit-exact; the split
    children's log-scale and opacity logit come from logf (within 1 ulp). The
    same with the split-factor knob (every scale / 1.6)."""
    from rgsplat_amd import runtime
    cfg = S.benchmark_train_config()
    wl = S.WORKLOADS["c3_1m_1080p_4views"]
    cap = 3 * wl.n
    ctx, cams, targets = setup("c3_1m_1080p_4views", capacity=cap)
    extent = runtime.scene_extent(wl.scene())  # from the initial positions (pipeline/mod.rs:214)
    try:
        for _ in range(3):
            ctx.train_step(list(range(len(cams))), cfg, use_graph=True)
        g0 = ctx.scene_download()
        st0 = ctx.adam_download()
        n = len(g0)
        # The reference's threshold (2e-4 on the pixel-space gradient norm of
        # a mean over 3 w h terms) is not reached at 1080p on this scene (the
        # 99th percentile of the statistics is ~2e-9): the rules are checked
        # with the threshold at the median of the non-zero statistics (half
        # of the seen Gaussians clone or split), and with threshold 0 and
        # max_gaussians 1.5M (the running cap binds part-way).
        avg = st0.grad_accum[:n] / np.maximum(st0.grad_denom[:n], 1)
        cfgs = {"median": dict(cfg, densify_grad_threshold=float(np.median(avg[avg > 0]))),
                "cap": dict(cfg, densify_grad_threshold=0.0, max_gaussians=1_500_000)}
        cfg = cfgs["median"]
        ctx.scene_densify(cfg, extent)
        g_dev = ctx.scene_download()
        st_dev = ctx.adam_download()
        reps = {}
        for case, div in (("median", 0.0), ("median", 1.6), ("cap", 0.0)):
            cfg = cfgs[case]
            buf = np.zeros((cap, 59), np.float32)
            buf[:n] = g0
            st = oracle.AdamState(n, cap=cap)
            st.m1[:n], st.m2[:n] = st0.m1[:n], st0.m2[:n]
            st.accum[:n], st.denom[:n] = st0.grad_accum[:n], st0.grad_denom[:n]
            nn = oracle.densify_and_prune(buf, n, st, cfg, extent, split_div=div)
            ref = buf[:nn]
            # the reference-API entry point on the same state
            api = np.zeros((cap, 59), np.float32)
            api[:n] = g0
            sta = runtime.AdamState(n, capacity=cap)
            sta.m1[:n], sta.m2[:n] = st0.m1[:n], st0.m2[:n]
            sta.grad_accum[:n], sta.grad_denom[:n] = st0.grad_accum[:n], st0.grad_denom[:n]
            ctx.set_split_scale_divisor(div)
            na = ctx.densify_and_prune(api, n, sta, cfg, extent)
            ctx.set_split_scale_divisor(0.0)
            assert na == nn
            logf_cols = [3, 8, 9, 10]
            exact_cols = [c for c in range(59) if c not in logf_cols]
            for got in ([api[:na]] + ([g_dev] if (case, div) == ("median", 0.0) else [])):
                assert len(got) == nn
                assert np.array_equal(got[:, exact_cols], ref[:, exact_cols])
                assert _ulps(got[:, logf_cols], ref[:, logf_cols]).max() <= 1
            assert sta.rows == st.rows and np.array_equal(sta.m1[:st.rows], st.m1[:st.rows])
            reps[f"{case}/{div}"] = {"before": n, "after": nn, "threshold": cfg["densify_grad_threshold"]}
        assert st_dev.rows == max(n, len(g_dev))
        assert np.all(st_dev.grad_accum[:st_dev.rows] == 0) and np.all(st_dev.grad_denom[:st_dev.rows] == 0)
        assert reps["median/0.0"]["after"] > n + n // 100  # the state densifies
        assert reps["cap/0.0"]["after"] < 1_500_000 + 2  # the running cap binds
        # training continues on the densified scene
        loss = ctx.train_step(list(range(len(cams))), cfgs["median"], use_graph=True)
        assert np.all(np.isfinite(loss))
    finally:
        ctx.close()
    report("c3_densify", reps)
