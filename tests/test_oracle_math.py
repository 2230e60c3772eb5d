"""Pins for the oracle's math restatement (math/*.rs).

Each test names the reference unit test it ports (same inputs, same tolerance)
or the independent fact it checks (closed forms, finite differences, numpy
linear algebra) so that a dropped term, a sign or an index slip fails here.
"""
import math

import numpy as np
import pytest

from rgsplat_amd import synthetic as S


def f32(a):
    return np.ascontiguousarray(a, np.float32)


def test_sigmoid_logit_roundtrip(orc):
    # math/mod.rs:215-221
    for v in [-3.0, -1.0, 0.0, 1.0, 3.0]:
        p = orc.lib().orc_sigmoid(v)
        assert abs(orc.lib().orc_logit(p) - v) < 1e-5
    # closed form
    assert orc.lib().orc_sigmoid(0.0) == 0.5
    assert abs(orc.lib().orc_sigmoid(2.0) - 1 / (1 + math.exp(-2))) < 1e-7


def test_covariance_psd_trace(orc):
    # math/mod.rs:223-231: Gaussian3D::new -> identity rotation, log_scale -3
    g = np.zeros(59, np.float32)
    g[4] = 1.0
    g[8:11] = -3.0
    g[3] = -2.0
    cov = orc.call3("orc_covariance_3d", g, out_shape=(3, 3))
    assert np.trace((cov + cov.T) / 2) > 0
    # closed form: R = I -> diag(exp(-3)^2)
    np.testing.assert_allclose(cov, np.diag([math.exp(-6)] * 3), rtol=1e-6, atol=0)


def test_covariance_3d_matches_numpy(orc):
    r = S.Stream(7)
    for _ in range(20):
        g = np.zeros(59, np.float32)
        q = r.normal((4,)) * 2.0  # unnormalised on purpose
        g[4:8] = q
        g[8:11] = r.normal((3,), -1.0, 0.5)
        qn = q / np.linalg.norm(q)
        w, x, y, z = qn
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        M = R @ np.diag(np.exp(g[8:11].astype(np.float64)))
        cov = orc.call3("orc_covariance_3d", g, out_shape=(3, 3))
        np.testing.assert_allclose(cov, M @ M.T, rtol=2e-5, atol=1e-7)


def test_rotation_identity_and_orthonormal(orc):
    # quaternion.rs:78-88
    r = orc.call3("orc_to_rotation_matrix", [1, 0, 0, 0], out_shape=(3, 3))
    np.testing.assert_allclose(r, np.eye(3), atol=1e-6)
    r = orc.call3("orc_to_rotation_matrix", [0.5, 0.5, 0.5, 0.5], out_shape=(3, 3))
    np.testing.assert_allclose(r @ r.T, np.eye(3), atol=1e-5)
    # a quarter turn about z maps x to y (Hamilton convention)
    h = math.sqrt(0.5)
    r = orc.call3("orc_to_rotation_matrix", [h, 0, 0, h], out_shape=(3, 3))
    np.testing.assert_allclose(r @ [1, 0, 0], [0, 1, 0], atol=1e-6)
    # near-zero quaternion -> identity (quaternion.rs:12-14)
    r = orc.call3("orc_to_rotation_matrix", [0, 0, 0, 0], out_shape=(3, 3))
    np.testing.assert_allclose(r, np.eye(3))


def test_rotation_jacobian_fd(orc):
    # quaternion.rs:91-106 (eps 1e-4, tolerance 1e-2 on the Frobenius norm)
    q = orc.call3("orc_q_normalize", [0.3, 0.7, -0.2, 0.5], out_shape=(4,))
    jac = orc.call3("orc_rotation_jacobian", q, out_shape=(4, 3, 3))
    eps = 1e-4
    for k in range(4):
        qp, qm = q.copy(), q.copy()
        qp[k] += eps
        qm[k] -= eps
        num = (orc.call3("orc_mat_from_quat", qp, out_shape=(3, 3)) -
               orc.call3("orc_mat_from_quat", qm, out_shape=(3, 3))) / (2 * eps)
        assert np.linalg.norm(jac[k] - num) < 1e-2
    # exact in fp64: the formula is quadratic, so central FD is exact up to rounding
    qd = q.astype(np.float64)

    def mat(q):
        w, x, y, z = q
        return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                         [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                         [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    for k in range(4):
        e = np.zeros(4)
        e[k] = 1e-3
        num = (mat(qd + e) - mat(qd - e)) / 2e-3
        np.testing.assert_allclose(jac[k], num, atol=2e-6)


def test_cov3d_symmetric_psd(orc):
    # covariance.rs:92-96
    cov = orc.call3("orc_build_cov3d", [0.3, 0.2, 0.5], [0.5, 0.5, 0.5, 0.5], out_shape=(3, 3))
    np.testing.assert_allclose(cov, cov.T, atol=1e-6)
    assert np.linalg.det(cov.astype(np.float64)) > 0
    # det(Sigma) = prod(s)^2 for a rotation
    assert abs(np.linalg.det(cov.astype(np.float64)) - (0.3 * 0.2 * 0.5) ** 2) < 1e-7


def test_backprop_cov3d_fd(orc):
    # covariance.rs:99-121, plus the rotation part (not in the reference test)
    # against fp64 central differences of <W, Sigma(s, q)> in the tangent space.
    w = np.array([[0.7, 0.1, -0.2], [0.1, 0.4, 0.3], [-0.2, 0.3, 0.9]], np.float32)
    log_scale = np.array([-1.0, -0.5, 0.2])
    rot = orc.call3("orc_q_normalize", [0.4, 0.2, -0.7, 0.5], out_shape=(4,))
    scale = np.exp(log_scale).astype(np.float32)
    d_ls = np.zeros(3, np.float32)
    d_rot = np.zeros(4, np.float32)
    orc.lib().orc_backprop_cov3d(orc._p(f32(scale)), orc._p(rot), orc._p(w), orc._p(d_ls), orc._p(d_rot))

    def loss(ls, q):
        qn = q / np.linalg.norm(q)
        a, x, y, z = qn
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - a * z), 2 * (x * z + a * y)],
                      [2 * (x * y + a * z), 1 - 2 * (x * x + z * z), 2 * (y * z - a * x)],
                      [2 * (x * z - a * y), 2 * (y * z + a * x), 1 - 2 * (x * x + y * y)]])
        M = R @ np.diag(np.exp(ls))
        return float(np.sum(w.astype(np.float64) * (M @ M.T)))
    eps = 1e-3
    for k in range(3):
        e = np.zeros(3)
        e[k] = eps
        num = (loss(log_scale + e, rot.astype(np.float64)) - loss(log_scale - e, rot.astype(np.float64))) / (2 * eps)
        assert abs(d_ls[k] - num) < 1e-2  # reference tolerance
        assert abs(d_ls[k] - num) < 1e-5 * max(1.0, abs(num))
    qd = rot.astype(np.float64)
    num = np.zeros(4)
    for k in range(4):
        e = np.zeros(4)
        e[k] = 1e-5
        num[k] = (loss(log_scale, qd + e) - loss(log_scale, qd - e)) / 2e-5
    # the analytic gradient omits the radial term: compare tangent components
    tang = lambda v: v - np.dot(v, qd) * qd
    np.testing.assert_allclose(tang(d_rot.astype(np.float64)), tang(num), atol=2e-5)


def test_sh_dc_only(orc):
    # sh.rs:146-156
    c = np.zeros(48, np.float32)
    C0 = np.float32(0.28209479177387814)
    c[0] = np.float32(1.0) / C0
    out = orc.call3("orc_eval_sh_deg3", c, [0, 0, 1], out_shape=(3,))
    assert abs(out[0] - 1.0) < 1e-5 and abs(out[1] - 0.5) < 1e-5 and abs(out[2] - 0.5) < 1e-5


def _real_sh_closed_form(d):
    """Real SH to degree 3 from the normalisation sqrt((2l+1)/4pi (l-m)!/(l+m)!)
    and associated Legendre polynomials, in the sign/order convention of the
    3DGS basis (sh.rs:39-61)."""
    x, y, z = d
    pi = math.pi
    c0 = 0.5 * math.sqrt(1 / pi)
    c1 = math.sqrt(3 / (4 * pi))
    return np.array([
        c0,
        -c1 * y, c1 * z, -c1 * x,
        0.5 * math.sqrt(15 / pi) * x * y,
        -0.5 * math.sqrt(15 / pi) * y * z,
        0.25 * math.sqrt(5 / pi) * (2 * z * z - x * x - y * y),
        -0.5 * math.sqrt(15 / pi) * x * z,
        0.25 * math.sqrt(15 / pi) * (x * x - y * y),
        -0.25 * math.sqrt(35 / (2 * pi)) * y * (3 * x * x - y * y),
        0.5 * math.sqrt(105 / pi) * x * y * z,
        -0.25 * math.sqrt(21 / (2 * pi)) * y * (4 * z * z - x * x - y * y),
        0.25 * math.sqrt(7 / pi) * z * (2 * z * z - 3 * x * x - 3 * y * y),
        -0.25 * math.sqrt(21 / (2 * pi)) * x * (4 * z * z - x * x - y * y),
        0.25 * math.sqrt(105 / pi) * z * (x * x - y * y),
        -0.25 * math.sqrt(35 / (2 * pi)) * x * (x * x - 3 * y * y),
    ])


def test_sh_basis_closed_form_and_orthonormal(orc):
    r = S.Stream(11)
    dirs = r.normal((64, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    for d in dirs:
        b = orc.call3("orc_sh_basis", d, out_shape=(16,))
        np.testing.assert_allclose(b, _real_sh_closed_form(d), rtol=2e-6, atol=2e-7)
    # orthonormality on the sphere (Monte-Carlo-free: Lebedev-like Fibonacci grid)
    n = 20000
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = math.pi * (1 + 5 ** 0.5) * i
    pts = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    B = np.stack([_real_sh_closed_form(p) for p in pts])
    gram = B.T @ B * (4 * math.pi / n)
    np.testing.assert_allclose(gram, np.eye(16), atol=2e-3)


def test_sh_basis_grad_fd(orc):
    d = np.array([0.3, -0.5, 0.81])
    g = orc.call3("orc_sh_basis_grad", d, out_shape=(16, 3))
    for c in range(3):
        e = np.zeros(3)
        e[c] = 1e-6
        num = (_real_sh_closed_form(d + e) - _real_sh_closed_form(d - e)) / 2e-6
        np.testing.assert_allclose(g[:, c], num, atol=2e-6)


def test_aabb_and_project_covariance(orc):
    # projection.rs:138-165 closed form: isotropic 2x2 with variance v -> r = 3 sqrt(v)
    s2 = np.diag([4.0, 4.0, 0.0]).astype(np.float32)
    out = np.zeros(4, np.int32)
    ok = orc.lib().orc_aabb(50.0, 60.0, orc._p(s2), 3.0, 200, 100, orc._p(out))
    assert ok and list(out) == [44, 54, 57, 67]
    # clipped to the image and empty when fully outside
    ok = orc.lib().orc_aabb(-20.0, 60.0, orc._p(s2), 3.0, 200, 100, orc._p(out))
    assert not ok
    ok = orc.lib().orc_aabb(2.0, 99.5, orc._p(s2), 3.0, 200, 100, orc._p(out))
    assert ok and list(out) == [0, 93, 9, 100]
    # EWA: on-axis point at depth z, isotropic sigma s -> (f s / z)^2 + 0.3
    view = np.eye(4, dtype=np.float32).reshape(16)
    sig = np.eye(3, dtype=np.float32) * 0.04
    o = np.zeros(9, np.float32)
    orc.lib().orc_project_covariance(orc._p(sig), orc._p(f32([0, 0, 2.0])), orc._p(f32(view)), 100.0, 100.0,
                                     orc._p(o))
    np.testing.assert_allclose(o.reshape(3, 3)[:2, :2], np.diag([(100 * 0.2 / 2) ** 2 + 0.3] * 2), rtol=1e-6)
    # off-axis: J = [[f/z, 0, -f x/z^2], ...] -> Sigma2D = J Sigma J^T
    p = np.array([0.3, -0.2, 2.0])
    orc.lib().orc_project_covariance(orc._p(sig), orc._p(f32(p)), orc._p(f32(view)), 100.0, 120.0, orc._p(o))
    J = np.array([[100 / 2, 0, -100 * 0.3 / 4], [0, 120 / 2, 120 * 0.2 / 4]])
    ref = J @ (np.eye(3) * 0.04) @ J.T + 0.3 * np.eye(2)
    np.testing.assert_allclose(o.reshape(3, 3)[:2, :2], ref, rtol=1e-5)


def test_srgb_endpoints(orc):
    # rasterizer.rs:285-290 and images.rs:88-95
    L = orc.lib()
    assert abs(L.orc_linear_to_srgb(0.0)) < 1e-6
    assert abs(L.orc_linear_to_srgb(1.0) - 1.0) < 1e-6
    assert L.orc_linear_to_srgb(0.5) > 0.5
    assert abs(L.orc_srgb_to_linear(0.0)) < 1e-6
    assert abs(L.orc_srgb_to_linear(1.0) - 1.0) < 1e-5
    assert L.orc_srgb_to_linear(0.5) < 0.5
    for x in [0.01, 0.2, 0.7]:
        assert abs(L.orc_srgb_to_linear(L.orc_linear_to_srgb(x)) - x) < 1e-5


def test_view_matrix_is_inverse(orc):
    cams = S.make_cameras(3003, [(64, 48)] * 8, focal=50.0)
    for c in cams:
        v = orc.view_matrix(c)
        np.testing.assert_allclose(v.astype(np.float64) @ c.c2w.astype(np.float64), np.eye(4), atol=2e-6)
    bad = S.identity_camera(8, 8)
    bad.c2w = np.zeros((4, 4), np.float32)
    with pytest.raises(ValueError):
        orc.view_matrix(bad)


def test_powi_matches_repeated_squaring(orc):
    L = orc.lib()
    for b in [1, 2, 3, 7, 100, 1001]:
        got = L.orc_powi(0.999, b)
        assert abs(got - 0.999 ** b) < 1e-5 * (0.999 ** b) * b ** 0.5 + 1e-7
    assert L.orc_powi(0.9, 1) == np.float32(0.9)
