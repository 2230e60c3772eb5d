"""SfM point clouds for the initialisation tests (pipeline/sfm.rs:365-455),
with the values gaussians_from_sfm must give where they follow from the
mathematics alone."""
import numpy as np

SH_C0 = 0.28209479177387814


def line(n, h, seed=0):
    """n points on a line along x, spacing h, in shuffled input order."""
    rng = np.random.default_rng(seed)
    x = np.arange(n, dtype=np.float64) * h
    perm = rng.permutation(n)
    xyz = np.zeros((n, 3))
    xyz[:, 0] = x[perm]
    xyz[:, 1] = 2.0
    xyz[:, 2] = -1.0
    rgb = rng.integers(0, 256, (n, 3), dtype=np.uint8)
    return xyz, rgb, perm


def line_scales(n, h, perm):
    """Mean of the 3 nearest of the window: 4h/3 inside, 2h at the first
    point. The last point's window is cut at the end of the sorted list and
    holds only the min(n, 50) // 2 points before it: with k = min(3, that)
    neighbours at h, 2h, .., kh its scale is (k + 1) h / 2 (n >= 4)."""
    s = np.full(n, 4.0 * h / 3.0)
    s[perm == 0] = 2.0 * h
    k = min(3, min(n, 50) // 2)
    s[perm == n - 1] = (k + 1) * h / 2.0
    return s


def cloud(n, seed, dup_frac=0.0, int_grid=False):
    """A random cloud; dup_frac of the points repeat an earlier x (ties in the
    x order and in the binary search); int_grid puts x on a coarse grid."""
    rng = np.random.default_rng(seed)
    xyz = rng.normal(0, 3, (n, 3))
    if int_grid:
        xyz[:, 0] = np.round(xyz[:, 0] * 4) / 4
    if dup_frac > 0:
        k = int(n * dup_frac)
        src = rng.integers(0, n, k)
        dst = rng.integers(0, n, k)
        xyz[dst, 0] = xyz[src, 0]
    rgb = rng.integers(0, 256, (n, 3), dtype=np.uint8)
    return xyz, rgb


def window_scales(xyz):
    """estimate_initial_scales restated with numpy: a stable argsort for the
    stable sort_by, searchsorted(side='right') - 1 for the binary search (its
    own x is present, and the search ends on the last equal key), the
    50-point window from 25 before it, and the mean of the 3 smallest
    distances, summed ascending from -0.0."""
    n = len(xyz)
    order = np.argsort(xyz[:, 0], kind="stable")
    sx = xyz[order, 0]
    window = min(50, n)
    out = np.zeros(n, np.float32)
    for i in range(n):
        idx = int(np.searchsorted(sx, xyz[i, 0], side="right")) - 1
        start = max(idx - window // 2, 0)
        end = min(start + window, n)
        js = order[start:end]
        js = js[js != i]
        if len(js) == 0:
            out[i] = np.float32(0.01)
            continue
        d = np.sort(np.sqrt(((xyz[js] - xyz[i]) ** 2).sum(1)))
        k = min(3, len(d))
        acc = -0.0
        for v in d[:k]:
            acc += v
        out[i] = np.float32(acc / k)
    return out
