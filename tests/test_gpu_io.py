"""Model files (io/ply.rs) through the C ABI against the oracle and the
hand-derived records of tests/ply_files.py: files written on the device path
are byte-identical to the oracle's, files read give bit-identical records."""
import numpy as np
import pytest

import oracle
from rgsplat_amd import synthetic as S
from tests import ply_files as F

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


def scenes():
    return [F.special_scene(), F.random_scene(1000, 5), F.random_scene(65, 6), np.zeros((0, 59), np.float32)]


def test_save_ply_bytes(ctx, tmp_path):
    for k, g in enumerate(scenes()):
        ctx.save_ply(g, tmp_path / f"gpu{k}.ply")
        oracle.save_ply(g, tmp_path / f"cpu{k}.ply")
        assert (tmp_path / f"gpu{k}.ply").read_bytes() == (tmp_path / f"cpu{k}.ply").read_bytes(), k


def test_scene_save_ply_bytes(ctx, tmp_path):
    g = F.random_scene(3000, 7)
    ctx.scene_upload(g, capacity=5000)
    ctx.scene_save_ply(tmp_path / "scene.ply")
    oracle.save_ply(g, tmp_path / "cpu.ply")
    assert (tmp_path / "scene.ply").read_bytes() == (tmp_path / "cpu.ply").read_bytes()


def test_save_splat_bytes(ctx, tmp_path):
    for k, g in enumerate(scenes()):
        ctx.save_splat(g, tmp_path / f"gpu{k}.splat")
        oracle.save_splat(g, tmp_path / f"cpu{k}.splat")
        assert (tmp_path / f"gpu{k}.splat").read_bytes() == (tmp_path / f"cpu{k}.splat").read_bytes(), k
    g = F.random_scene(777, 8)
    ctx.scene_upload(g)
    ctx.scene_save_splat(tmp_path / "scene.splat")
    oracle.save_splat(g, tmp_path / "cpu.splat")
    assert (tmp_path / "scene.splat").read_bytes() == (tmp_path / "cpu.splat").read_bytes()


def test_load_ply_crafted(ctx, tmp_path):
    for path, expected in F.crafted_files(tmp_path):
        got = ctx.load_ply(path)
        assert same(got, expected), path.name
        assert same(got, oracle.load_ply(path)), path.name
        ctx.scene_load_ply(path, capacity=len(expected) + 3)
        assert same(ctx.scene_download(), expected), path.name


def test_load_ply_roundtrip(ctx, tmp_path):
    for k, g in enumerate(scenes()):
        oracle.save_ply(g, tmp_path / f"m{k}.ply")
        assert same(ctx.load_ply(tmp_path / f"m{k}.ply"), g), k


@pytest.mark.parametrize("kind", ["magic", "format", "type", "count", "eof", "short"])
def test_load_ply_rejects(ctx, tmp_path, kind):
    from rgsplat_amd.runtime import RgsError
    path = F.bad_files(tmp_path)[kind]
    with pytest.raises(RgsError, match="error 7"):
        ctx.load_ply(path)
    with pytest.raises(RgsError, match="error 6"):
        ctx.load_ply(tmp_path / "no_such_file.ply")


def test_checkpoint_after_training(ctx, tmp_path):
    """A trained device scene written as a PLY and loaded back resumes from
    the same parameters (the checkpoint the reference writes every N
    iterations, pipeline/mod.rs:248-256)."""
    from rgsplat_amd import runtime
    wl = S.scaled(S.WORKLOADS["c0_10k_256"], n=3000)
    g, cams = wl.scene(), wl.cameras()
    tg = [np.concatenate([oracle.render_color(wl.target_scene(), c), np.ones((c.height, c.width, 1), np.float32)], -1)
          for c in cams]
    cfg = S.benchmark_train_config()
    c = runtime.Context(0)
    c.scene_upload(g)
    c.set_views(cams, tg)
    for _ in range(3):
        c.train_step([0], cfg, use_graph=False)
    trained = c.scene_download()
    c.scene_save_ply(tmp_path / "ckpt.ply")
    assert same(oracle.load_ply(tmp_path / "ckpt.ply"), trained)
    c.scene_load_ply(tmp_path / "ckpt.ply")
    assert same(c.scene_download(), trained)
    c.close()


def test_load_ply_short_ascii_and_huge_counts(ctx, tmp_path):
    """An ASCII body with fewer lines than its vertex count: read_line at EOF
    gives empty lines, whose properties read as 0 (io/ply.rs:170-179). A
    header that claims 2^31 vertices costs no host memory: a binary body is
    checked against the file's length before anything is allocated
    (RGS_E_FORMAT), an ASCII one stores only its lines and reports the count
    to a caller without room for it (RGS_E_CAPACITY); no exception leaves the
    library."""
    import ctypes
    from rgsplat_amd import runtime
    props = ["x", "y", "z", "opacity", "f_dc_0", "rot_1"]
    head = ("ply\nformat ascii 1.0\nelement vertex {n}\n" + "".join(f"property float {p}\n" for p in props)
            + "end_header\n")
    short = tmp_path / "short.ply"
    short.write_text(head.format(n=5) + "1 2 3 0.5 0.1 0.2\n4 5 6\n")
    got = ctx.load_ply(short)
    assert len(got) == 5
    assert same(got, oracle.load_ply(short))
    ctx.scene_load_ply(short)
    assert same(ctx.scene_download(), got)
    L = runtime.lib()
    n = ctypes.c_size_t()
    big_ascii = tmp_path / "big_ascii.ply"
    big_ascii.write_text(head.format(n=2 ** 31) + "1 2 3 0.5 0.1 0.2\n")
    assert L.rgs_load_ply(ctx.h, bytes(big_ascii), None, 0, ctypes.byref(n)) == 5 and n.value == 2 ** 31
    small = np.zeros((4, 59), np.float32)
    assert L.rgs_load_ply(ctx.h, bytes(big_ascii), small.ctypes.data_as(ctypes.c_void_p), 4, ctypes.byref(n)) == 5
    big_bin = tmp_path / "big_bin.ply"
    big_bin.write_bytes(head.format(n=2 ** 31).replace("ascii", "binary_little_endian").encode()
                        + np.arange(12, dtype=np.float32).tobytes())
    assert L.rgs_load_ply(ctx.h, bytes(big_bin), None, 0, ctypes.byref(n)) == 7
    assert L.rgs_scene_load_ply(ctx.h, bytes(big_bin), 0) == 7
    # the context still works
    assert same(ctx.load_ply(short), got)
