"""The COLMAP readers and camera building (pipeline/sfm.rs:132-350,
pipeline/mod.rs:270-296) on files written here byte by byte, with the values
they must give derived by hand from the layouts, the rotation checked
against the oracle's quaternion::to_rotation_matrix and the poses against
their inverse."""
import ctypes
import struct

import numpy as np
import pytest

import oracle
from rgsplat_amd import colmap


def points_file():
    b = struct.pack("<Q", 2)
    b += struct.pack("<Q3d3BdQ", 7, 1.5, -2.25, 3.0, 10, 20, 30, 0.5, 2) + b"\x01" * 16
    b += struct.pack("<Q3d3BdQ", 9, -0.0, 1e300, 4.0, 255, 0, 128, 0.0, 0)
    return b


def cameras_file():
    b = struct.pack("<Q", 5)
    b += struct.pack("<IiQQ3d", 1, 0, 640, 480, 500.0, 320.0, 240.0)           # SIMPLE_PINHOLE
    b += struct.pack("<IiQQ4d", 2, 1, 800, 600, 700.0, 710.0, 400.5, 300.25)   # PINHOLE
    b += struct.pack("<IiQQ4d", 3, 2, 0, 0, 650.0, 330.0, 250.0, 0.1)          # SIMPLE_RADIAL, no size
    b += struct.pack("<IiQQ8d", 4, 4, 1024, 768, 900.0, 905.0, 512.0, 384.0, 0, 0, 0, 0)  # OPENCV
    b += struct.pack("<IiQQ4d", 5, 42, (1 << 32) + 16, 32, 1.0, 2.0, 3.0, 4.0)  # unknown model: 4 params
    return b


def images_file():
    b = struct.pack("<Q", 3)
    q = (np.cos(0.4), 0.0, np.sin(0.4), 0.0)
    b += struct.pack("<I4d3dI", 11, *q, 0.5, -1.0, 4.0, 2) + b"view_a.png\x00" + struct.pack("<Q", 2) + b"\x00" * 48
    b += struct.pack("<I4d3dI", 12, 0.5, 0.5, 0.5, 0.5, 0.0, 0.0, 2.0, 1) + b"b.jpg\x00" + struct.pack("<Q", 0)
    b += struct.pack("<I4d3dI", 13, 1.0, 0, 0, 0, 0, 0, 0, 99) + b"orphan.png\x00" + struct.pack("<Q", 0)
    return b


def test_points3d():
    p = colmap.parse_points3d_bin(points_file())
    assert p.ids.tolist() == [7, 9] and p.track_length.tolist() == [2, 0]
    assert p.xyz.tolist() == [[1.5, -2.25, 3.0], [-0.0, 1e300, 4.0]]
    assert p.rgb.tolist() == [[10, 20, 30], [255, 0, 128]] and p.error.tolist() == [0.5, 0.0]
    with pytest.raises(ValueError):
        colmap.parse_points3d_bin(points_file()[:-1])


def test_cameras():
    c = colmap.parse_cameras_bin(cameras_file(), 1920, 1080)
    assert (c[1].width, c[1].height, c[1].fx, c[1].fy, c[1].cx, c[1].cy) == (640, 480, 500, 500, 320, 240)
    assert (c[2].fx, c[2].fy, c[2].cx, c[2].cy) == (700, 710, 400.5, 300.25)
    assert (c[3].width, c[3].height, c[3].fx, c[3].fy) == (1920, 1080, 650, 650)
    assert (c[4].fx, c[4].fy, c[4].cx, c[4].cy) == (900, 905, 512, 384)
    assert (c[5].width, c[5].height, c[5].fx, c[5].fy, c[5].cx) == (16, 32, 1, 1, 2)  # `as u32`, f, cx, cy
    with pytest.raises(ValueError):
        colmap.parse_cameras_bin(cameras_file()[:-8], 1920, 1080)


def test_images_and_poses():
    ims = colmap.parse_images_bin(images_file())
    assert [i.name for i in ims] == ["view_a.png", "b.jpg", "orphan.png"]
    assert [i.camera_id for i in ims] == [2, 1, 99] and ims[0].tvec == (0.5, -1.0, 4.0)
    intr = colmap.parse_cameras_bin(cameras_file(), 1920, 1080)
    cams, paths = colmap.build_cameras(ims, intr, "/data/img")
    assert paths == ["/data/img/view_a.png", "/data/img/b.jpg"]  # the unknown camera is skipped
    lib = oracle.lib()
    for cam, img in zip(cams, ims):
        R = np.zeros(9, np.float32)
        q = np.array(img.qvec, np.float32)
        lib.orc_to_rotation_matrix(q.ctypes.data, R.ctypes.data)
        R = R.reshape(3, 3)
        assert np.array_equal(cam.c2w[:3, :3], R.T)
        t = np.array(img.tvec)
        np.testing.assert_allclose(cam.c2w[:3, 3], -(R.T.astype(np.float64) @ t), atol=1e-6)
        # world -> camera of the pose is [R | t]
        w2c = oracle.view_matrix(cam)
        np.testing.assert_allclose(w2c[:3, :3], R, atol=1e-6)
        np.testing.assert_allclose(w2c[:3, 3], t, atol=1e-5)
    assert (cams[0].fx, cams[1].fx) == (700, 500)


def test_load_sfm_output(tmp_path):
    model = tmp_path / "ws" / "sparse" / "0"
    model.mkdir(parents=True)
    (model / "points3D.bin").write_bytes(points_file())
    (model / "cameras.bin").write_bytes(cameras_file())
    (model / "images.bin").write_bytes(images_file())
    pts, cams, paths = colmap.load_sfm_output(str(tmp_path / "ws"), "imgs")
    assert len(pts.xyz) == 2 and len(cams) == 2 and paths[0].endswith("view_a.png")
    g, _ = oracle.gaussians_from_sfm(pts.xyz, pts.rgb)
    assert g.shape == (2, 59)
    with pytest.raises(FileNotFoundError):
        colmap.load_sfm_output(str(tmp_path / "none_here"), "imgs")
