"""The oracle's restatement of the model files (io/ply.rs), pinned by the
reference's own tests (io/ply.rs:305-345: PLY round trip, 32-byte .splat
records), by the header text and record layout save_ply writes, by
hand-derived load_ply results on crafted files (tests/ply_files.py) and by
closed-form .splat bytes."""
import struct

import numpy as np
import pytest

import oracle
from tests import ply_files as F


def same(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_ply_roundtrip_preserves_parameters(tmp_path):
    """io/ply.rs:309-333 (the reference's test, same inputs)."""
    g = F.default_record()
    g[F.POS:F.POS + 3] = [1.0, -2.0, 3.5]
    g[F.OPL] = 0.42
    g[F.LS:F.LS + 3] = [-1.0, -2.0, -0.5]
    q = np.array([0.3, 0.7, -0.2, 0.5], np.float32)
    qn = np.zeros(4, np.float32)
    oracle.lib().orc_q_normalize(oracle._p(q), oracle._p(qn))
    g[F.ROT:F.ROT + 4] = qn
    g[F.SH:] = np.arange(48, dtype=np.float32) * np.float32(0.01) - np.float32(0.2)
    original = np.stack([g, F.default_record()])
    oracle.save_ply(original, tmp_path / "model.ply")
    loaded = oracle.load_ply(tmp_path / "model.ply")
    assert loaded.shape == original.shape
    assert np.all(np.abs(loaded - original) < 1e-5)
    assert same(loaded, original)  # the binary format is lossless


def test_save_ply_layout(tmp_path):
    g = F.special_scene()
    oracle.save_ply(g, tmp_path / "m.ply")
    data = (tmp_path / "m.ply").read_bytes()
    head = F.reference_header(len(g))
    assert data[: len(head)] == head
    assert data[len(head):] == F.reference_body(g)


def test_splat_record_is_32_bytes(tmp_path):
    """io/ply.rs:336-344."""
    g = np.stack([F.default_record()] * 10)
    oracle.save_splat(g, tmp_path / "m.splat")
    assert (tmp_path / "m.splat").stat().st_size == 32 * 10


def test_splat_closed_forms(tmp_path):
    """Records whose bytes follow from the definitions: scale = exp(0) = 1,
    colour 0.5 + 0 * C0 -> round(127.5) = 128, sigmoid(0) = 0.5 -> 128, the
    identity quaternion -> (255, 128, 128, 128) (256 clamps to 255), a zero
    quaternion -> identity, NaN -> 0 through Rust's `as u8`."""
    g = np.stack([F.default_record() for _ in range(3)])
    for r in g:
        r[F.LS:F.LS + 3] = 0.0
        r[F.OPL] = 0.0
        r[F.POS:F.POS + 3] = [1.0, 2.0, 3.0]
    g[1, F.ROT:F.ROT + 4] = 0.0
    g[2, F.OPL] = np.nan
    g[2, F.SH] = 100.0      # clamps to 1 -> 255
    g[2, F.SH + 16] = -100.0  # clamps to 0 -> 0
    oracle.save_splat(g, tmp_path / "c.splat")
    data = (tmp_path / "c.splat").read_bytes()
    for i in range(3):
        rec = data[32 * i: 32 * (i + 1)]
        assert struct.unpack("<6f", rec[:24]) == (1.0, 2.0, 3.0, 1.0, 1.0, 1.0)
    assert data[24:32] == bytes([128, 128, 128, 128, 255, 128, 128, 128])
    assert data[32 + 24:32 + 32] == bytes([128, 128, 128, 128, 255, 128, 128, 128])
    assert data[64 + 24:64 + 32] == bytes([255, 0, 128, 0, 255, 128, 128, 128])


def test_load_ply_crafted(tmp_path):
    for path, expected in F.crafted_files(tmp_path):
        got = oracle.load_ply(path)
        assert same(got, expected), path.name


@pytest.mark.parametrize("kind", ["magic", "format", "type", "count", "eof", "short"])
def test_load_ply_rejects(tmp_path, kind):
    path = F.bad_files(tmp_path)[kind]
    with pytest.raises(ValueError):
        oracle.load_ply(path)
