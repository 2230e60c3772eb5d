"""train_blocking (pipeline/mod.rs:203-265) over the device-resident step.

The reference's loop renders one camera per iteration on the CPU, takes an
Adam step, densifies and prunes every 100 iterations between
densify_from_iter and densify_until_iter, resets opacities every
opacity_reset_iter iterations and writes a PLY checkpoint every
save_interval iterations. Here the scene, the Adam moments and the targets
stay on the device: each iteration is one rgs_train_step for the camera the
reference's xorshift picks, densification is rgs_scene_densify with the
extent of the initial positions, and checkpoints are written from the device
scene.
"""
from __future__ import annotations

import os
import time
from dataclasses import dataclass, field

import numpy as np

from . import runtime, synthetic

XORSHIFT_SEED = 0x9E3779B97F4A7C15
_M64 = (1 << 64) - 1


def camera_sequence(n_cams: int, n: int, seed: int = XORSHIFT_SEED) -> list[int]:
    """The camera of each of the first n iterations (pipeline/mod.rs:224-230:
    xorshift64 with shifts 13, 7, 17, then `rng as usize % n_cams`)."""
    rng, out = seed, []
    for _ in range(n):
        rng ^= (rng << 13) & _M64
        rng ^= rng >> 7
        rng ^= (rng << 17) & _M64
        out.append(rng % n_cams)
    return out


@dataclass
class TrainLog:
    losses: list = field(default_factory=list)     # per iteration
    cameras: list = field(default_factory=list)    # per iteration
    sizes: list = field(default_factory=list)      # Gaussians after each iteration's densify
    checkpoints: list = field(default_factory=list)
    densify: list = field(default_factory=list)    # (iteration, Gaussians before, after, seconds)


def train(ctx: runtime.Context, gaussians, cameras, images, cfg: dict | None = None, bg=(0.0, 0.0, 0.0),
          save_interval: int = 0, out_dir: str | None = None, lambda_dssim: float = 0.2,
          use_graph: bool = True, all_views: bool = False) -> tuple[np.ndarray, TrainLog]:
    """Runs cfg['num_iterations'] iterations of train_blocking's loop on the
    device and returns the trained Gaussians (n, 59) and a log. `images` are
    the cameras' linear RGBA targets (h, w, 4), or None when the views are
    already registered with the context. The scene's capacity is
    cfg['max_gaussians'] (densify never grows past it, as in the reference).
    save_interval 0 writes no checkpoints. all_views: every iteration is one
    step over all the cameras (their gradients summed, one Adam step) instead
    of the reference's one camera per iteration; the logged loss is their
    mean."""
    cfg = dict(cfg or synthetic.train_config())
    g = np.ascontiguousarray(gaussians, np.float32).reshape(-1, 59)
    extent = runtime.scene_extent(g)  # estimate_scene_extent, once (pipeline/mod.rs:214)
    cap = max(int(cfg["max_gaussians"]), len(g), 1)
    ctx.scene_upload(g, capacity=cap)
    if images is not None:
        ctx.set_views(cameras, images)
    cams = camera_sequence(len(cameras), int(cfg["num_iterations"]))
    log = TrainLog()
    batch = list(range(len(cameras)))
    for it, cam in enumerate(cams):
        ids = batch if all_views else [cam]
        loss = ctx.train_step(ids, cfg, bg=bg, lambda_dssim=lambda_dssim, use_graph=use_graph)
        log.losses.append(float(np.mean(np.asarray(loss))))
        log.cameras.append(-1 if all_views else cam)
        if cfg["densify_from_iter"] <= it <= cfg["densify_until_iter"] and it % 100 == 0 and it > 0:
            n0 = ctx.scene_size()
            t0 = time.perf_counter()
            ctx.scene_densify(cfg, extent)
            log.densify.append((it, n0, ctx.scene_size(), time.perf_counter() - t0))
        if it % cfg["opacity_reset_iter"] == 0 and it > 0:
            ctx.scene_reset_opacities()
        if save_interval and it % save_interval == 0 and it > 0:
            path = os.path.join(out_dir or ".", f"ckpt_{it:06d}.ply")
            ctx.scene_save_ply(path)
            log.checkpoints.append(path)
        log.sizes.append(ctx.scene_size())
    return ctx.scene_download(), log


def train_from_sfm(ctx: runtime.Context, xyz, rgb, cameras, images_rgba8, cfg: dict | None = None,
                   bg=(0.0, 0.0, 0.0), save_interval: int = 1, out_dir: str = ".",
                   use_graph: bool = True) -> tuple[np.ndarray, TrainLog]:
    """Stages 3-5 of the reference's pipeline (pipeline/mod.rs:140-172,
    177-193, 270-296) from SfM points and decoded images: gaussians_from_sfm
    on the device, the targets resized and converted on the device
    (load_target after decoding), train_blocking's loop, then output.ply and
    output.splat in out_dir. As in the reference, save_interval is at least
    1. Parsing COLMAP's binary files and decoding the images stay with the
    caller."""
    g = ctx.gaussians_from_sfm(xyz, rgb)
    ctx.set_views_rgba8_sized(cameras, images_rgba8)
    out, log = train(ctx, g, cameras, None, cfg, bg=bg, save_interval=max(int(save_interval), 1),
                     out_dir=out_dir, use_graph=use_graph)
    ctx.scene_save_ply(os.path.join(out_dir, "output.ply"))
    ctx.scene_save_splat(os.path.join(out_dir, "output.splat"))
    return out, log
