"""The preview renderer (GpuRenderer::project / render, gpu/compute.rs and
gpu/splat.wgsl) through the C ABI: SplatInfo2D records and images
bit-identical to the oracle's."""
import numpy as np
import pytest

import oracle
from rgsplat_amd import synthetic as S
from tests import preview_cases as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from rgsplat_amd import runtime
    c = runtime.Context(0)
    yield c
    c.close()


def same(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def cases():
    out = []
    cam = S.look_at_origin(np.array([1.0, 0.5, 0.3]), 3.0, 80, 60, 90.0)
    g = C.scene(300, 4, opacity=(0.001, 0.99))
    g[::7, C.ROT:C.ROT + 4] *= 1.7
    g[0:10, C.OPL] = -7.0
    g[10:15, 0:3] = (np.asarray(cam.c2w, np.float64)[:3, 3] * 2.5).astype(np.float32)
    out.append(("mixed", g, cam))
    out.append(("dense", C.scene(3000, 5, opacity=(0.05, 0.99)),
                S.look_at_origin(np.array([0.2, 1.0, 0.4]), 3.5, 203, 117, 160.0)))
    out.append(("large_splats", C.scene(200, 6, opacity=(0.5, 0.99), spread=0.5),
                S.look_at_origin(np.array([0.0, 0.3, 1.0]), 2.0, 96, 96, 120.0)))
    g = C.scene(400, 7, opacity=(0.3, 0.9))
    g[200:400] = g[0:200]                         # equal depths: index order decides
    g[200:400, C.SH:C.SH + 48] = -g[0:200, C.SH:C.SH + 48]
    out.append(("ties", g, S.look_at_origin(np.array([0.5, 0.5, 1.0]), 3.0, 64, 64, 80.0)))
    out.append(("smoke", C.smoke_gaussian(), S.identity_camera(32, 32)))
    return out


@pytest.mark.parametrize("k", range(5))
def test_preview_project(ctx, k):
    name, g, cam = cases()[k]
    assert same(ctx.preview_project(g, cam), oracle.preview_project(g, cam)), name


@pytest.mark.parametrize("k", range(5))
def test_preview_render(ctx, k):
    name, g, cam = cases()[k]
    for bg in [(0.0, 0.0, 0.0), (0.25, 0.5, 1.0)]:
        assert same(ctx.preview_render(g, cam, bg), oracle.preview_render(g, cam, bg)), name


def test_gpu_smoke_port(ctx):
    """tests/gpu_smoke.rs:29-56 on the device."""
    cam = S.identity_camera(32, 32)
    sp = ctx.preview_project(C.smoke_gaussian(), cam)
    assert sp.shape == (1, 10) and sp[0, 8] > 0 and sp[0, 9] > 0 and np.isfinite(sp[0, :2]).all()
    assert ctx.preview_render(C.smoke_gaussian(), cam)[16, 16, 0] > 0


def test_scene_preview_and_edges(ctx):
    from rgsplat_amd.runtime import RgsError
    g = C.scene(20000, 8, opacity=(0.05, 0.99))
    cam = S.look_at_origin(np.array([0.3, 0.7, 1.0]), 3.0, 320, 200, 300.0)
    ctx.scene_upload(g, capacity=25000)
    img = ctx.scene_preview_render(cam, (0.1, 0.2, 0.3))
    assert same(img, oracle.preview_render(g, cam, (0.1, 0.2, 0.3)))
    assert same(img, ctx.preview_render(g, cam, (0.1, 0.2, 0.3)))
    empty = ctx.preview_render(np.zeros((0, 59), np.float32), cam, (0.1, 0.2, 0.3))
    assert same(empty, np.broadcast_to(np.float32([0.1, 0.2, 0.3]), (200, 320, 3)))
    assert ctx.preview_project(np.zeros((0, 59), np.float32), cam).shape == (0, 10)
    bad = S.identity_camera(32, 32)
    bad.c2w = np.zeros((4, 4), np.float32)
    with pytest.raises(RgsError, match="error 3"):
        ctx.preview_render(g[:10], bad)


def test_scene_preview_to_device(ctx):
    import torch
    g = C.scene(5000, 12, opacity=(0.05, 0.99))
    cam = S.look_at_origin(np.array([0.1, 0.9, 0.6]), 3.0, 200, 120, 180.0)
    ctx.scene_upload(g)
    out = torch.empty((120, 200, 3), dtype=torch.float32, device="cuda:0")
    ctx.scene_preview_render_device(cam, out, (0.3, 0.2, 0.1))
    assert same(out.cpu().numpy(), oracle.preview_render(g, cam, (0.3, 0.2, 0.1)))


def test_scene_preview_to_device_after_torch_work(ctx):
    """The render into a torch tensor is ordered after the work torch has
    queued on its own stream: a fill of `out` held back by a spin kernel
    must land before the image, not over it."""
    import torch
    g = C.scene(3000, 13, opacity=(0.05, 0.99))
    cam = S.look_at_origin(np.array([0.4, -0.7, 0.5]), 3.0, 160, 96, 150.0)
    ctx.scene_upload(g)
    want = oracle.preview_render(g, cam, (0.0, 0.0, 0.0))
    out = torch.empty((96, 160, 3), dtype=torch.float32, device="cuda:0")
    for _ in range(3):
        torch.cuda._sleep(20_000_000)  # ~10 ms of spinning on torch's stream
        out.fill_(7.0)
        ctx.scene_preview_render_device(cam, out)
        torch.cuda.synchronize()
        assert same(out.cpu().numpy(), want)
