#!/usr/bin/env python3
"""Training-step benchmark on one MI355X.

A step is one iteration of the reference's training loop body
(pipeline/mod.rs:232-237) over a batch of views: forward and backward of every
view, gradients summed, one Adam step. The default workload is 1M Gaussians
(SH degree 3) with 4 views of 1920x1080 per step, run as the device-resident,
batched, hipGraph-replayed step. Inputs are synthetic and seeded (the scene
distributions of the workload table in rgsplat_amd/synthetic.py); targets are
renders of a second seeded scene.

Prints one JSON line. Besides the throughput it reports the roofline of the
dominant kernel (HIP-event timings of each stage on the context's stream,
algorithmic work counted on the device), the CPU oracle timed on the host on a
bounded sample of the same workload, the plain (one view per call, host
buffers) GPU path, and the gradient relative error against the CPU oracle at
the small parity size.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "train iters/sec (fwd+bwd+Adam) at 1M Gaussians 1080p; grad rel-err vs CPU"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3

# Algorithmic work per unit (per stage):
#   raster_fwd: 25 FP32 ops + 1 exp per evaluated (pixel, splat) pair
#   raster_bwd: 60 FP32 ops + 1 exp + 1 div per evaluated pair
#   loss_ssim:  1440 FLOP per pixel (30 blurs x 22 taps x 2 + elementwise)
#   project:    236 B read per Gaussian + 60 B written per Gaussian and view
#               (splat record 52 B, depth key and value 8 B)
#   geometry_adam: 1652 B per Gaussian (59 params x 28 B + 8 B statistics)
FWD_OPS_PER_PAIR = 26
BWD_OPS_PER_PAIR = 62
SSIM_FLOP_PER_PIXEL = 1440
# The VALU kernels use no matrix cores. 157.3 TFLOP/s is MI355X's published
# FP32 vector peak, which counts packed FP32 (two FMAs per lane per cycle);
# this build disables packed FP32, so its FP32 issue ceiling is half of it.
FP32_SCALAR_TFLOPS = FP32_PEAK_TFLOPS / 2


def measured_traffic(workload: str) -> dict:
    """HBM bytes per launch of each stage's main kernel, from the committed
    rocprofv3 PMC summary of this workload (FETCH_SIZE and WRITE_SIZE passes,
    each scaled by the counter's reading for a streaming kernel of known
    size). Empty when no summary exists for the workload."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.md")
    if not os.path.exists(path):
        return {}
    text = open(path).read()
    start = text.find("```json")
    if start < 0:
        return {}
    body = text[start + 7:text.index("```", start + 7)]
    table = json.loads(body)
    return {stage: rec["traffic_bytes"] for stage, rec in table.get("stages", {}).items()}


def sync():
    """Device-wide barrier. Every train_step already ends with a stream
    synchronisation; torch's is added when torch is loaded in the process."""
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_available():
        torch.cuda.synchronize()


def with_alpha(rgb):
    return np.concatenate([rgb, np.ones(rgb.shape[:2] + (1,), np.float32)], -1)


def cpu_info() -> dict:
    """Host CPU model and the threads this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count() or 1, "allowed": allowed}


def step_check_and_cpu_baseline(ctx, ids, cams, targets, cfg, threads, cpu=True):
    """One captured step of the timed configuration (graph replay, the
    adaptive depth cut) against the oracle, and the oracle's own step time.

    The step's summed gradients are compared with the oracle's
    render_and_backward of every view at the pre-step parameters (max over
    parameter groups of ||g - g_ref|| / ||g_ref||, and the per-component
    maximum), and with the oracle's backward driven by the loss gradient of
    the step's own image (the L1 term's signum flips where rounding moves a
    pixel across its target, which on a trained scene moves the gradient by
    ~1e-4 relative). The same oracle work times the CPU path: 1 warm-up view,
    then 3 full steps (every view's render_and_backward plus one adam_step),
    median of the 3."""
    import oracle
    from tests.step_parity import group_rel_err, component_rel_err
    oracle.set_threads(threads)
    g0 = ctx.scene_download()
    ctx.set_grad_capture(True)
    ctx.train_step(ids, cfg, use_graph=True)
    grads = ctx.step_grads()
    img = ctx.step_image(sum(c.width * c.height for c in cams))
    ctx.set_grad_capture(False)
    out = {}
    t0 = time.perf_counter()
    oracle.render_and_backward(g0, cams[0], targets[0])
    warm = time.perf_counter() - t0
    steps = []
    ref = ref_dev = None
    img_err = 0.0
    for rep in range(3 if cpu else 1):
        t0 = time.perf_counter()
        tot = np.zeros((len(g0), 60), np.float32 if rep else np.float64)
        extra = 0.0
        base = 0
        for c, t in zip(cams, targets):
            r = oracle.Render(g0, c, t)
            tot += r.grads()
            if rep == 0:
                e0 = time.perf_counter()
                npx = c.width * c.height
                rgb = r.image()[0]
                dimg = img[base:base + npx].reshape(c.height, c.width, 3)
                base += npx
                img_err = max(img_err, float(np.abs(dimg - rgb).max()))
                _, dl = oracle.loss_grad(dimg, t, c.width, c.height, 0.2)
                g_dev = r.backward_with_dl(dl)
                ref_dev = g_dev.astype(np.float64) if ref_dev is None else ref_dev + g_dev
                extra += time.perf_counter() - e0
            del r
        gg = np.ascontiguousarray(g0.copy())
        st = oracle.AdamState(len(gg))
        oracle.adam_step(gg, tot.astype(np.float32), st, cfg)
        steps.append(time.perf_counter() - t0 - extra)
        if ref is None:
            ref = tot
        print(f"[bench] oracle step {rep}: {steps[-1]:.2f} s", file=sys.stderr, flush=True)
    groups = group_rel_err(grads, ref)
    out["grad_rel_err"] = max(groups.values())
    out["grad_rel_err_groups"] = groups
    out["grad_component_rel_err"] = component_rel_err(grads, ref)
    same = group_rel_err(grads, ref_dev)
    out["grad_rel_err_same_dl"] = max(same.values())
    out["image_max_abs_err"] = img_err
    if cpu:
        med = float(np.median(steps))
        info = cpu_info()
        out["cpu_baseline"] = {
            "value": 1.0 / med, "unit": "iters/s", "cores": threads, "kind": "port",
            "cpu_model": info["model"], "host_nproc": info["nproc"], "host_allowed_cpus": info["allowed"],
            "sample": f"the oracle (OpenMP, {threads} threads) on this workload's full step: render_and_backward "
                      f"of all {len(cams)} views + 1 adam_step at the pre-step parameters of a timed-configuration "
                      f"step; 1 warm-up view ({warm:.2f} s), then 3 steps, median "
                      f"({', '.join(f'{x:.2f}' for x in steps)} s)"}
    return out


def plain_path_api(ctx, g, cams, targets, cfg):
    """The reference-API path on the GPU: one render_and_backward call per
    view (host buffers in and out) and one adam_step call: PCIe-bound."""
    from rgsplat_amd.runtime import AdamState
    gg = np.ascontiguousarray(g.copy())
    st = AdamState(len(gg))
    steps = 2
    for it in range(steps + 1):
        if it == 1:
            t0 = time.perf_counter()
        tot = np.zeros((len(gg), 60), np.float32)
        for c, t in zip(cams, targets):
            _, gr = ctx.render_and_backward(gg, c, t)
            tot += gr
        ctx.adam_step(gg, tot, st, cfg)
    t1 = time.perf_counter()
    return steps / (t1 - t0)


def plain_path_device(runtime, g, cams, targets, cfg, steps=3):
    """The plain device path: the scene resident in HBM, one view per launch
    sequence (train_step of a single view, eager, no graph), its own Adam
    step per view as the reference's loop does. Reported per iteration of the
    same views as the batched step (one eager single-view step per view)."""
    ctx = runtime.Context(0)
    ctx.scene_upload(g, capacity=len(g))
    ctx.set_views(cams, targets)
    for v in range(len(cams)):
        ctx.train_step([v], cfg, use_graph=False)
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        for v in range(len(cams)):
            ctx.train_step([v], cfg, use_graph=False)
    sync()
    t1 = time.perf_counter()
    ctx.close()
    return steps / (t1 - t0)


def model_files(runtime, wl, reps):
    """Wall time of the model-file entry points on the workload's scene (the
    device scene to a file and back; files under the temp directory)."""
    import tempfile
    g = wl.scene()
    ctx = runtime.Context(0)
    ctx.scene_upload(g)
    d = tempfile.mkdtemp()
    ply, spl = os.path.join(d, "scene.ply"), os.path.join(d, "scene.splat")
    res = {}
    for name, fn in (("scene_save_ply", lambda: ctx.scene_save_ply(ply)),
                     ("scene_load_ply", lambda: ctx.scene_load_ply(ply)),
                     ("scene_save_splat", lambda: ctx.scene_save_splat(spl))):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        res[name + "_ms"] = (time.perf_counter() - t0) * 1000.0 / reps
        print(f"[bench] {name} {res[name + '_ms']:.1f} ms", file=sys.stderr, flush=True)
    res["ply_bytes"] = os.path.getsize(ply)
    res["splat_bytes"] = os.path.getsize(spl)
    ok = np.array_equal(ctx.scene_download().view(np.uint32), g.view(np.uint32))
    ctx.close()
    os.remove(ply)
    os.remove(spl)
    os.rmdir(d)
    print(json.dumps({"metric": "model files (wall ms per call, file I/O included)", "gaussians": len(g),
                      "reps": reps, "roundtrip_bit_exact": bool(ok), **res}), flush=True)


def sfm_init(runtime, wl, reps, cpu=True):
    """Wall time of the SfM initialisation (gaussians_from_sfm) on a point
    cloud of the workload's size: the workload's scene positions as f64 SfM
    points with random colours, host arrays in, device scene out; the oracle
    on the same cloud on the host cores beside it."""
    g = wl.scene()
    rng = np.random.default_rng(11)
    xyz = g[:, 0:3].astype(np.float64)
    rgb = rng.integers(0, 256, (len(g), 3), dtype=np.uint8)
    del g
    ctx = runtime.Context(0)
    ctx.scene_from_sfm(xyz, rgb)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.scene_from_sfm(xyz, rgb)
    ms = (time.perf_counter() - t0) * 1000.0 / reps
    print(f"[bench] scene_from_sfm {ms:.1f} ms", file=sys.stderr, flush=True)
    res = {"scene_from_sfm_ms": ms}
    if cpu:
        import oracle
        t0 = time.perf_counter()
        want, _ = oracle.gaussians_from_sfm(xyz, rgb)
        res["cpu_oracle_ms"] = (time.perf_counter() - t0) * 1000.0
        res["cpu_threads"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        got = ctx.scene_download()
        a, b = got.view(np.uint32), want.view(np.uint32)
        other = np.ones(59, bool)
        other[8:11] = False
        res["bit_exact_except_log_scale"] = bool(np.array_equal(a[:, other], b[:, other]))
        res["log_scale_max_ulp"] = int(np.abs(a[:, 8:11].astype(np.int64) - b[:, 8:11].astype(np.int64)).max())
        print(f"[bench] oracle gaussians_from_sfm {res['cpu_oracle_ms']:.0f} ms", file=sys.stderr, flush=True)
    ctx.close()
    print(json.dumps({"metric": "SfM initialisation (wall ms per call, host arrays in, device scene out)",
                      "points": len(xyz), "reps": reps, **res}), flush=True)


def preview(runtime, wl, reps, cpu=True):
    """Wall time of the preview render (GpuRenderer::render) of the
    workload's scene, resident on the device, at each of its views; the
    oracle on the first view beside it."""
    g = wl.scene()
    cams = wl.cameras()
    ctx = runtime.Context(0)
    ctx.scene_upload(g)
    bg = (0.0, 0.0, 0.0)
    img = ctx.scene_preview_render(cams[0], bg)
    t0 = time.perf_counter()
    for _ in range(reps):
        for cam in cams:
            ctx.scene_preview_render(cam, bg)
    ms = (time.perf_counter() - t0) * 1000.0 / (reps * len(cams))
    print(f"[bench] scene_preview_render {ms:.2f} ms per view", file=sys.stderr, flush=True)
    res = {"scene_preview_render_ms_per_view": ms}
    if cpu:
        import oracle
        t0 = time.perf_counter()
        want = oracle.preview_render(g, cams[0], bg)
        res["cpu_oracle_ms_first_view"] = (time.perf_counter() - t0) * 1000.0
        res["cpu_threads"] = 1
        res["bit_exact_first_view"] = bool(np.array_equal(img.view(np.uint32), want.view(np.uint32)))
        print(f"[bench] oracle preview_render {res['cpu_oracle_ms_first_view']:.0f} ms", file=sys.stderr, flush=True)
    ctx.close()
    print(json.dumps({"metric": "preview render (wall ms per view, device scene in, host image out)",
                      "gaussians": len(g), "views": [[c.width, c.height] for c in cams], "reps": reps, **res}),
          flush=True)


def target_ingestion(runtime, wl, reps, cpu=True):
    """Wall time of registering the workload's views from decoded RGBA8
    images 1.5x the cameras' sizes: upload, resize (image 0.25's Triangle
    filter) and sRGB-to-linear conversion on the device; the oracle's resize
    and conversion of the first view beside it."""
    cams = wl.cameras()
    rng = np.random.default_rng(12)
    ims = [rng.integers(0, 256, (c.height * 3 // 2, c.width * 3 // 2, 4), dtype=np.uint8) for c in cams]
    ctx = runtime.Context(0)
    ctx.set_views_rgba8_sized(cams, ims)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.set_views_rgba8_sized(cams, ims)
    ms = (time.perf_counter() - t0) * 1000.0 / reps
    print(f"[bench] set_views_rgba8_sized {ms:.1f} ms for {len(cams)} views", file=sys.stderr, flush=True)
    res = {"set_views_rgba8_sized_ms": ms}
    if cpu:
        import oracle
        t0 = time.perf_counter()
        oracle.rgba8_to_linear(oracle.resize_rgba8(ims[0], cams[0].width, cams[0].height))
        res["cpu_oracle_ms_first_view"] = (time.perf_counter() - t0) * 1000.0
        res["cpu_threads"] = 1
    ctx.close()
    print(json.dumps({"metric": "target ingestion (wall ms per registration of all views, RGBA8 host images in)",
                      "views": [[c.width, c.height] for c in cams],
                      "source": [[im.shape[1], im.shape[0]] for im in ims], "reps": reps, **res}), flush=True)


def densifying_run(runtime, wl, iters, grad_threshold=None):
    """The workload's densifying schedule: `iters` iterations of one batched
    step over all its views, densify_and_prune every 100 iterations from 500
    (pipeline/mod.rs:240-242), the scene growing in place up to
    max_gaussians. Reports iterations/s including the densifies, their cost
    and N(t)."""
    from rgsplat_amd import pipeline
    g = wl.scene()
    cams = wl.cameras()
    ctx = runtime.Context(0)
    tscene = wl.target_scene()
    targets = [with_alpha(ctx.render_color(tscene, c)) for c in cams]
    del tscene
    cfg = workload_train_config(num_iterations=iters)
    if grad_threshold is not None:
        cfg["densify_grad_threshold"] = grad_threshold
    ctx.scene_upload(g, capacity=int(cfg["max_gaussians"]))
    ctx.set_views(cams, targets)
    tuned = ctx.autotune_raster(list(range(len(cams))))
    sync()
    t0 = time.perf_counter()
    _, log = pipeline.train(ctx, g, cams, None, cfg, all_views=True)
    sync()
    wall = time.perf_counter() - t0
    dens_s = sum(d[3] for d in log.densify)
    ctx.close()
    out = {"metric": "densifying training run (iterations/s including densify)", "value": iters / wall,
           "unit": "iters/s", "iterations": iters, "wall_s": wall, "densify_s": dens_s,
           "config": {"workload": wl.name, "views_per_step": len(cams), "densify_every": 100,
                      "densify_from_iter": cfg["densify_from_iter"], "max_gaussians": cfg["max_gaussians"],
                      "densify_grad_threshold": cfg["densify_grad_threshold"],
                      "raster_params": tuned},
           "n_t": [[d[0], d[1], d[2], round(d[3] * 1000.0, 2)] for d in log.densify],
           "n_t_fields": ["iteration", "gaussians_before", "gaussians_after", "densify_ms"],
           "final_gaussians": log.sizes[-1], "loss_first_last": [log.losses[0], log.losses[-1]]}
    print(f"[bench] densifying run {iters} iterations in {wall:.1f} s ({dens_s:.2f} s densify)", file=sys.stderr,
          flush=True)
    print(json.dumps(out), flush=True)


def workload_train_config(**over):
    from rgsplat_amd import synthetic as S
    cfg = S.benchmark_train_config()
    cfg.update(over)
    return cfg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3_1m_1080p_4views")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-path measurement")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-stage event timing")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default raster work decomposition")
    ap.add_argument("--io", action="store_true",
                    help="time the model files instead (PLY save/load and .splat save of the workload's scene)")
    ap.add_argument("--sfm", action="store_true",
                    help="time the SfM initialisation instead (gaussians_from_sfm on a cloud of the workload's size)")
    ap.add_argument("--preview", action="store_true",
                    help="time the preview renderer instead (GpuRenderer::render of the workload's scene)")
    ap.add_argument("--targets", action="store_true",
                    help="time the target ingestion instead (RGBA8 images resized and converted on the device)")
    ap.add_argument("--densify", type=int, default=0, metavar="ITERS",
                    help="run the workload's densifying schedule for ITERS iterations instead (C3: 2000)")
    ap.add_argument("--densify-grad-threshold", type=float, default=None,
                    help="densify_grad_threshold of the densifying run (default: the reference's 2e-4)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the step would shard by views "
              "(each GPU renders its share of the views, then a reduce-scatter of the gradient rows, Adam on "
              "each GPU's shard and an all-gather of the parameters), but no multi-GPU path is built",
              file=sys.stderr)
        sys.exit(2)

    from rgsplat_amd import runtime
    from rgsplat_amd import synthetic as S

    wl = S.WORKLOADS[args.workload]
    if args.io:
        model_files(runtime, wl, args.steps)
        return
    if args.sfm:
        sfm_init(runtime, wl, args.steps, cpu=not args.no_cpu)
        return
    if args.preview:
        preview(runtime, wl, args.steps, cpu=not args.no_cpu)
        return
    if args.targets:
        target_ingestion(runtime, wl, args.steps, cpu=not args.no_cpu)
        return
    if args.densify:
        densifying_run(runtime, wl, args.densify, args.densify_grad_threshold)
        return
    print(f"[bench] workload {wl.name}: {wl.n} Gaussians, views {wl.sizes}", file=sys.stderr, flush=True)
    g = wl.scene()
    cams = wl.cameras()
    ctx = runtime.Context(0)
    # the practical ceiling of the memory-bound stages: a plain streaming
    # device copy of 1 GiB (read + written bytes per second)
    copy_gbs = ctx.hbm_copy_probe(1 << 30, 10)
    print(f"[bench] HBM copy probe {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    tscene = wl.target_scene()
    targets = [with_alpha(ctx.render_color(tscene, c)) for c in cams]
    # The oracle check and CPU baseline run on the timed views, except for a
    # clustered scene, whose 1080p oracle backward would hold ~1.5e9 (tile,
    # splat) accumulators per view: there they run on a 128x128 window of
    # each view around its densest cluster (registered beside the timed views).
    check_cams, check_targets = cams, targets
    if wl.clusters:
        check_cams = S.cluster_crops(wl)
        check_targets = [with_alpha(ctx.render_color(tscene, c)) for c in check_cams]
    del tscene
    cfg = S.benchmark_train_config()
    ctx.scene_upload(g, capacity=len(g))
    ids = list(range(len(cams)))
    if check_cams is cams:
        ctx.set_views(cams, targets)
        check_ids = ids
    else:
        ctx.set_views(cams + check_cams, targets + check_targets)
        check_ids = list(range(len(cams), len(cams) + len(check_cams)))
    use_graph = not args.no_graph
    tuned = None
    if not args.no_autotune:
        tuned = ctx.autotune_raster(ids)
        print(f"[bench] raster autotune {tuned}", file=sys.stderr, flush=True)
    # The warm-up steps carry HIP events around every stage (in graph mode
    # they are event nodes of the captured step) to find the dominant stage;
    # the timed steps carry events around that stage only, so its duration
    # comes from the timed region itself at the cost of one event pair. All
    # stages are timed again afterwards, outside the timed region.
    profile = not args.no_profile
    if profile:
        ctx.set_profiling(True)
    print("[bench] warmup", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        ctx.train_step(ids, cfg, use_graph=use_graph)
    dom = None
    if profile:
        warm_stages, _, _ = ctx.stage_times()
        dom = max(warm_stages, key=lambda s: warm_stages[s])
        ctx.set_profiling_stages([dom])
        ctx.train_step(ids, cfg, use_graph=use_graph)  # capture the step with the dominant stage's events
        ctx.set_profiling_stages([dom])  # restart the counters for the timed steps
    sync()
    t0 = time.perf_counter()
    losses = None
    for _ in range(args.steps):
        losses = ctx.train_step(ids, cfg, use_graph=use_graph)
    sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1000.0 / args.steps
    print(f"[bench] {ms:.3f} ms/step, losses {losses}", file=sys.stderr, flush=True)

    out = {"metric": METRIC, "value": 1000.0 / ms, "unit": "iters/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene, targets rendered from a second seeded scene)",
           "config": {"workload": wl.name, "gaussians": wl.n, "sh_degree": 3, "views_per_step": len(cams),
                      "resolutions": sorted(set(f"{w}x{h}" for w, h in wl.sizes)),
                      "execution": "device-resident step, views batched per launch, "
                                   + ("hipGraph replay" if use_graph else "eager"),
                      "raster_params": dict(zip(("super_tile", "sparse_lanes"), ctx.raster_params()),
                                            batch=ctx.raster_batch(), queue=ctx.raster_queue()),
                      "raster_autotuned": tuned is not None}}

    if profile:
        timed, n_prof, n_graph = ctx.stage_times()
        # every stage, on as many further steps (outside the timed region)
        ctx.set_profiling(True)
        ctx.train_step(ids, cfg, use_graph=use_graph)
        ctx.set_profiling(True)
        for _ in range(args.steps):
            ctx.train_step(ids, cfg, use_graph=use_graph)
        stages, _, _ = ctx.stage_times()
        stages[dom] = timed[dom]
        ctx.set_profiling(False)
        # the same steps without the stage events, for the events' overhead
        ctx.train_step(ids, cfg, use_graph=use_graph)
        sync()
        u0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.train_step(ids, cfg, use_graph=use_graph)
        sync()
        out["uninstrumented_iters_per_s"] = args.steps / (time.perf_counter() - u0)
        K = ctx.step_pairs()
        cut = ctx.cut_stats()
        # the raster kernels' own work on the last step's lists (counting
        # instantiations of the same kernels)
        try:
            kernel_counts = ctx.raster_work_counts()
        except runtime.RgsError as e:
            kernel_counts = {"unavailable": str(e)}
        # Algorithmic work, independent of this build's binning: the (pixel,
        # splat) pairs the reference's forward and backward evaluate on its
        # own tile lists (each view rendered once more through the reference-
        # binning entry point on the trained scene).
        g_now = ctx.scene_download()
        pf = pb = 0
        for cam in cams:
            ctx.render_color(g_now, cam)
            _, f_, b_ = ctx.step_work()
            pf += f_
            pb += b_
        del g_now
        npix = sum(w * h for w, h in wl.sizes)
        ng = ctx.scene_size()
        # the tile sort's LSD passes: 8 bits each over the global 32x32-bin ids
        n_bins = sum(((w + 15) // 16 + 1) // 2 * (((h + 15) // 16 + 1) // 2) for w, h in wl.sizes)
        sort_passes = max(1, (int(n_bins - 1).bit_length() + 7) // 8)
        work = {
            "raster_fwd": ("valu", FWD_OPS_PER_PAIR * pf, "TFLOP/s"),
            "raster_bwd": ("valu", BWD_OPS_PER_PAIR * pb, "TFLOP/s"),
            "loss_ssim": ("valu", SSIM_FLOP_PER_PIXEL * npix, "TFLOP/s"),
            # the projection reads 59 parameters per Gaussian and writes 64 B
            # per (Gaussian, view): sort key and value, depth, screen mean and
            # conic, pair count, tile rect, conic a / opacity / colour
            "project": ("hbm", 236 * ng + 64 * ng * len(cams), "GB/s"),
            # the update reads and writes the 59 parameters and both Adam
            # moments (1416 B), reads and clears the nine raster accumulators
            # and reads the pair count of every view (76 B per view), and
            # updates the two densification statistics (16 B)
            "geometry_adam": ("hbm", (1432 + 76 * len(cams)) * ng, "GB/s"),
            # pair generation writes 8 B per (bin, splat) pair; each tile-sort
            # pass reads and writes key and value (16 B) and the histogram
            # pass reads the key again (4 B)
            "pair_gen": ("hbm", 8 * cut["front_pairs"], "GB/s"),
            "tile_sort": ("hbm", 20 * sort_passes * cut["front_pairs"], "GB/s"),
        }
        traffic = measured_traffic(wl.name)
        rl = {}
        for s, (bound, amount, unit) in work.items():
            sec = stages[s] / 1000.0
            if bound == "hbm":
                ach = amount / sec / 1e9 if sec > 0 else 0.0
                peak = HBM_PEAK_GBS
            else:
                ach = amount / sec / 1e12 if sec > 0 else 0.0
                peak = FP32_PEAK_TFLOPS
            rl[s] = {"bound": bound, "achieved": ach, "peak": peak, "unit": unit, "frac": ach / peak,
                     "traffic": traffic.get(s), "ms": stages[s]}
            if bound == "hbm":
                rl[s]["frac_of_copy"] = ach / copy_gbs
            if bound == "valu":
                rl[s]["frac_scalar_issue"] = ach / FP32_SCALAR_TFLOPS
                rl[s]["achieved_basis"] = ("reference-equivalent: the reference algorithm's evaluated (pixel, splat) "
                                           "pairs x ops per pair; the kernel culls most of them"
                                           if s != "loss_ssim" else "FLOP per pixel x pixels")
        if dom in rl:
            out["roofline"] = dict(rl[dom], kernel=dom)
            if traffic.get(dom) is not None:
                out["roofline"]["traffic_source"] = (f"profiles/hbm_traffic_{wl.name}.md (rocprofv3 FETCH_SIZE/"
                                                     "WRITE_SIZE: bytes between L2 and the memory fabric, which "
                                                     "the Infinity Cache may serve; an upper bound on HBM bytes)")
        else:
            out["roofline"] = {"kernel": dom, "bound": None, "achieved": None, "peak": None, "unit": None,
                               "frac": None, "traffic": traffic.get(dom), "ms": stages[dom]}
        out["hbm_copy_probe"] = {"gbs": copy_gbs, "bytes": 1 << 30, "frac_of_peak": copy_gbs / HBM_PEAK_GBS,
                                 "what": "plain streaming device copy (one 16-byte load and store per thread), read + written "
                                         "bytes per second: the practical ceiling of the memory-bound stages "
                                         "(frac_of_copy)"}
        out["stages_ms"] = stages
        out["stages_source"] = (f"{dom}: the timed steps (events around it only); the others: {args.steps} "
                                "further steps with events around every stage")
        out["stage_rooflines"] = rl
        out["work"] = {"sorted_pairs": K, "front_pairs": cut["front_pairs"], "back_pairs": cut["back_pairs"],
                       "flagged_bins": cut["flagged_tiles"], "fwd_pair_evals": pf, "bwd_pair_evals": pb,
                       "pair_evals_source": "reference tile lists (API render of the trained scene)",
                       "kernel_counts": kernel_counts,
                       "gaussians": ng, "timed_steps_with_events": n_prof, "of_which_graph_replays": n_graph}
        print(f"[bench] stages {json.dumps(stages)}", file=sys.stderr, flush=True)

    threads = min(16, cpu_info()["allowed"])
    try:
        chk = step_check_and_cpu_baseline(ctx, check_ids, check_cams, check_targets, cfg, threads,
                                          cpu=not args.no_cpu)
        out.update(chk)
        if check_cams is not cams:
            out["check_views"] = ("128x128 windows of the timed views around their densest clusters (the full "
                                  "3M-Gaussian scene, each tile listing tens of thousands of Gaussians)")
        out["grad_rel_err_source"] = ("one captured step of the timed configuration (graph replay, depth cut) at "
                                      "the full workload size after the timed steps, against the oracle's "
                                      "render_and_backward of every view, max over parameter groups; "
                                      "grad_rel_err_same_dl: against the oracle's backward of the loss gradient of "
                                      "the step's own image")
    except ImportError as e:  # the oracle may be unbuilt on a bare box
        out["grad_rel_err"] = None
        print(f"[bench] grad check skipped: {e}", file=sys.stderr)
    if not args.no_plain:
        print("[bench] plain paths", file=sys.stderr, flush=True)
        ctx.close()
        ctx = None
        plain = plain_path_device(runtime, g, cams, targets, cfg)
        out["plain_gpu_path_iters_per_s"] = plain
        out["plain_gpu_path"] = ("device-resident eager train_step of one view per launch sequence, one per view "
                                 "of the batch (each with its Adam step, as the reference's loop); per iteration "
                                 "over the batch's views")
        out["speedup_vs_plain_gpu_path"] = out["value"] / plain
        ctx = runtime.Context(0)
        out["api_gpu_path_iters_per_s"] = plain_path_api(ctx, g, cams, targets, cfg)
        out["api_gpu_path"] = ("render_and_backward per view with host buffers + adam_step (PCIe-bound: "
                               "parameters and gradients cross PCIe per view)")
    if "cpu_baseline" in out and check_cams is cams:
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    if ctx is not None:
        ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
