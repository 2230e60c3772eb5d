"""The training driver (train_blocking, pipeline/mod.rs:203-265) on the
device: the loss falls, checkpoints hold the scene of their iteration and
the returned Gaussians are the device scene."""
import numpy as np
import pytest

import oracle
from rgsplat_amd import pipeline, synthetic as S

pytestmark = pytest.mark.gpu


def test_train_driver(tmp_path):
    from rgsplat_amd import runtime
    wl = S.scaled(S.WORKLOADS["c0_10k_256"], n=3000, sizes=[(96, 64), (80, 64), (64, 48)])
    g, cams = wl.scene(), wl.cameras()
    tscene = wl.target_scene()
    targets = [np.concatenate([oracle.render_color(tscene, c), np.ones((c.height, c.width, 1), np.float32)], -1)
               for c in cams]
    cfg = S.benchmark_train_config()
    cfg.update(num_iterations=160, densify_from_iter=0, densify_until_iter=1000, opacity_reset_iter=150,
               max_gaussians=8000)
    ctx = runtime.Context(0)
    out, log = pipeline.train(ctx, g, cams, targets, cfg, save_interval=50, out_dir=str(tmp_path))
    assert log.cameras == pipeline.camera_sequence(3, 160)
    assert len(log.losses) == 160 and np.all(np.isfinite(log.losses))
    # the loss falls until the opacity reset at 150 (which raises it again)
    assert np.mean(log.losses[140:150]) < np.mean(log.losses[:10])
    assert np.mean(log.losses[151:]) > np.mean(log.losses[140:150])
    assert [p.split("/")[-1] for p in log.checkpoints] == ["ckpt_000050.ply", "ckpt_000100.ply",
                                                           "ckpt_000150.ply"]
    for p, it in zip(log.checkpoints, (50, 100, 150)):
        assert len(oracle.load_ply(p)) == log.sizes[it]
    assert out.shape == (log.sizes[-1], 59)
    assert np.array_equal(out.view(np.uint32), ctx.scene_download().view(np.uint32))
    # the checkpoint at 150 follows that iteration's opacity reset (min(0.01, sigma))
    ck = oracle.load_ply(log.checkpoints[-1])
    assert np.isfinite(ck).all()
    assert np.all(1.0 / (1.0 + np.exp(-ck[:, 3].astype(np.float64))) <= 0.01 + 1e-6)
    # one densify, at iteration 100, logged with the sizes around it
    assert [d[0] for d in log.densify] == [100]
    assert log.densify[0][1] == log.sizes[99] and log.densify[0][2] == log.sizes[100]
    # all views per iteration (one batched step each)
    cfg.update(num_iterations=20)
    _, log2 = pipeline.train(ctx, g, cams, targets, cfg, all_views=True)
    assert log2.cameras == [-1] * 20 and np.all(np.isfinite(log2.losses))
    assert np.mean(log2.losses[-5:]) < np.mean(log2.losses[:5])
    ctx.close()


def test_train_from_sfm(tmp_path):
    """SfM points and RGBA8 images of another size to output.ply / .splat."""
    from rgsplat_amd import runtime
    from tests import sfm_points
    xyz, rgb = sfm_points.cloud(2000, 11)
    xyz *= 0.3
    cams = S.make_cameras(5, [(64, 48), (80, 64)], focal_scale=1.0)
    rng = np.random.default_rng(2)
    ims = [rng.integers(0, 256, (c.height * 2, c.width * 2, 4), dtype=np.uint8) for c in cams]
    cfg = S.benchmark_train_config()
    cfg.update(num_iterations=30, max_gaussians=4000)
    ctx = runtime.Context(0)
    out, log = pipeline.train_from_sfm(ctx, xyz, rgb, cams, ims, cfg, save_interval=10, out_dir=str(tmp_path))
    assert len(log.checkpoints) == 2 and np.all(np.isfinite(log.losses))
    assert np.array_equal(oracle.load_ply(tmp_path / "output.ply").view(np.uint32), out.view(np.uint32))
    oracle.save_splat(out, tmp_path / "cpu.splat")
    assert (tmp_path / "output.splat").read_bytes() == (tmp_path / "cpu.splat").read_bytes()
    ctx.close()
