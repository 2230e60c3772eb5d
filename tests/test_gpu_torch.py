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
    """fp32 central differences of a torch loss on top of the renderer."""
    import torch
    from rgsplat_amd import torch_ops
    cam = S.identity_camera(32, 24)
    g = np.zeros((2, 59), np.float32)
    g[:, 4] = 1.0
    g[0, 0:3] = [0.05, -0.02, 2.0]
    g[1, 0:3] = [-0.1, 0.05, 2.5]
    g[:, 3] = 0.5
    g[:, 8:11] = [[-2.2, -2.0, -2.1], [-2.0, -2.3, -2.2]]
    g[:, 11] = 0.4
    g[:, 27] = 0.3
    g[:, 43] = 0.2
    w = torch.linspace(0.1, 1.0, 24 * 32 * 3, device="cuda").reshape(24, 32, 3)

    def loss(x):
        return (torch_ops.render(x, cam) * w).sum()
    gt = torch.tensor(g, device="cuda", requires_grad=True)
    loss(gt).backward()
    an = gt.grad.cpu().numpy()
    eps = 1e-3
    for i in range(2):
        for k in [0, 1, 2, 3, 8, 9, 10, 11, 27]:
            gp, gm = g.copy(), g.copy()
            gp[i, k] += eps
            gm[i, k] -= eps
            with torch.no_grad():
                num = (loss(torch.tensor(gp, device="cuda")) - loss(torch.tensor(gm, device="cuda"))).item() / (2 * eps)
            assert abs(an[i, k] - num) <= 0.02 * max(abs(num), 1e-2), (i, k, an[i, k], num)
