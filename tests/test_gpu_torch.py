"""torch.autograd.Function over the HIP renderer against the CPU oracle."""
import numpy as np
import pytest

import oracle
from rgsplat_amd import synthetic as S

pytestmark = pytest.mark.gpu

# Gaussian3D offset -> GaussianGrad offset
G3D_FROM_FLAT = [0, 1, 2, 10, 3, 4, 5, 6, 7, 8, 9] + list(range(11, 59))


def test_autograd_render_matches_oracle():
    import torch
    from rgsplat_amd import torch_ops
    wl = S.scaled(S.WORKLOADS["c0_10k_256"], n=3000, sizes=[(128, 96)])
    g = wl.scene()
    cam = wl.cameras()[0]
    rgb_t = oracle.render_color(wl.target_scene(), cam)
    target = np.concatenate([rgb_t, np.ones(rgb_t.shape[:2] + (1,), np.float32)], -1)
    gt = torch.tensor(g, device="cuda", requires_grad=True)
    rgb = torch_ops.render(gt, cam)
    r = oracle.Render(g, cam, target)
    ref_rgb = r.image()[0]
    assert np.array_equal(rgb.detach().cpu().numpy(), ref_rgb)
    # feed the oracle's dL/dRGB as the upstream gradient
    rgb.backward(torch.tensor(r.dl(), device="cuda"))
    got = gt.grad.cpu().numpy()
    ref = r.grads()[:, G3D_FROM_FLAT]
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= 1e-3
    # interleaved forwards before backward (as gradcheck does) stay correct
    gt2 = torch.tensor(g, device="cuda", requires_grad=True)
    out = torch_ops.render(gt2, cam)
    torch_ops.render(gt2.detach() * 1.01, cam)
    out.backward(torch.tensor(r.dl(), device="cuda"))
    err2 = np.linalg.norm(gt2.grad.cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err2 <= 1e-3


def test_autograd_finite_differences():
    """backward.rs:907-982 through the torch binding: one Gaussian, 24x24,
    L1 against a black target (locally smooth), fp32 central differences
    with eps 1e-3; rel < 0.05 or abs < 3e-4 (the reference's bars)."""
    import math

    import torch
    from rgsplat_amd import torch_ops
    cam = S.identity_camera(24, 24)
    g = np.zeros((1, 59), np.float32)
    g[0, 0:3] = [0.1, -0.05, 2.0]
    g[0, 3] = math.log(0.6 / 0.4)
    q = np.array([0.9, 0.1, -0.2, 0.15])
    g[0, 4:8] = q / np.linalg.norm(q)
    g[0, 8:11] = [-2.2, -2.0, -2.1]
    for ch in range(3):
        g[0, 11 + ch * 16] = 0.4
        for k in range(1, 16):
            g[0, 11 + ch * 16 + k] = math.sin((ch * 16 + k) * 0.05) * 0.04

    def loss(x):
        return torch_ops.render(x, cam).abs().mean()
    gt = torch.tensor(g, device="cuda", requires_grad=True)
    loss(gt).backward()
    an = gt.grad.cpu().numpy()[0]
    eps = 1e-3

    def num(k):
        gp, gm = g.copy(), g.copy()
        gp[0, k] += eps
        gm[0, k] -= eps
        with torch.no_grad():
            return (loss(torch.tensor(gp, device="cuda")) - loss(torch.tensor(gm, device="cuda"))).item() / (2 * eps)
    for k in [0, 1, 2, 3, 8, 9, 10, 11, 12, 27, 44]:
        n = num(k)
        rel = abs(an[k] - n) / max(abs(an[k]), abs(n), 1e-4)
        assert rel < 0.05 or abs(an[k] - n) < 3e-4, (k, an[k], n)


def test_outputs_ordered_with_torch_streams():
    """The renderer runs on torch's current stream, including the default
    stream (handle 0): a torch reduction launched right after the render sees
    the finished image, on the default stream and on a side stream."""
    import math

    import torch
    from rgsplat_amd import torch_ops
    cam = S.identity_camera(64, 48)
    wl = S.scaled(S.WORKLOADS["c0_10k_256"], n=400, sizes=[(64, 48)])
    g = wl.scene()
    refs, xs = [], []
    for k in range(6):
        x = g.copy()
        x[:, 2] += 0.05 * k
        xs.append(x)
        refs.append(float(np.abs(oracle.render_color(x, cam)).astype(np.float64).mean()))
    side = torch.cuda.Stream()
    for stream in (torch.cuda.default_stream(), side):
        with torch.cuda.stream(stream):
            got = [torch_ops.render(torch.tensor(x, device="cuda"), cam).abs().mean() for x in xs]
            vals = [v.item() for v in got]
        for v, r in zip(vals, refs):
            assert math.isclose(v, r, rel_tol=1e-5), (v, r)
