"""The training driver's host logic (train_blocking, pipeline/mod.rs:203-265)
with a recording stand-in for the device context: the camera sequence, the
densify / opacity-reset / checkpoint cadence and the extent."""
import numpy as np

from rgsplat_amd import pipeline, runtime, synthetic


def xorshift_numpy(n_cams, n):
    """xorshift64 (13, 7, 17) with numpy's wrapping uint64 arithmetic."""
    r = np.array([pipeline.XORSHIFT_SEED], np.uint64)
    out = []
    for _ in range(n):
        r ^= r << np.uint64(13)
        r ^= r >> np.uint64(7)
        r ^= r << np.uint64(17)
        out.append(int(r[0] % np.uint64(n_cams)))
    return out


def test_camera_sequence():
    for n_cams in (1, 3, 7, 64, 1000):
        assert pipeline.camera_sequence(n_cams, 500) == xorshift_numpy(n_cams, 500)
    # the first state of the generator from the seed, by hand
    s = 0x9E3779B97F4A7C15
    s ^= (s << 13) % 2**64
    s ^= s >> 7
    s ^= (s << 17) % 2**64
    assert pipeline.camera_sequence(2**32, 1) == [s % 2**32]


class Recorder:
    def __init__(self):
        self.calls = []
        self.n = 0

    def scene_upload(self, g, capacity=None):
        self.calls.append(("upload", len(g), capacity))
        self.n = len(g)
        self.g = np.array(g)

    def set_views(self, cams, images):
        self.calls.append(("views", len(cams)))

    def train_step(self, ids, cfg, bg=None, lambda_dssim=None, use_graph=None):
        self.calls.append(("step", tuple(ids)))
        return np.array([0.5], np.float32)

    def scene_densify(self, cfg, extent):
        self.calls.append(("densify", extent))
        self.n += 1

    def scene_reset_opacities(self):
        self.calls.append(("reset",))

    def scene_save_ply(self, path):
        self.calls.append(("save", path))

    def scene_size(self):
        return self.n

    def scene_download(self):
        return self.g


def test_schedule():
    rec = Recorder()
    g = synthetic.make_scene(50, 3)
    cfg = synthetic.train_config(num_iterations=701, densify_from_iter=200, densify_until_iter=600,
                                 opacity_reset_iter=300, max_gaussians=1000)
    cams = [None] * 5
    out, log = pipeline.train(rec, g, cams, [None] * 5, cfg, save_interval=250, out_dir="/tmp/x")
    steps = [c for c in rec.calls if c[0] == "step"]
    assert [c[1][0] for c in steps] == pipeline.camera_sequence(5, 701)
    it_of = []  # iteration index of every non-step call
    it = -1
    for c in rec.calls:
        if c[0] == "step":
            it += 1
        elif c[0] in ("densify", "reset", "save"):
            it_of.append((c[0], it))
    assert [i for k, i in it_of if k == "densify"] == [200, 300, 400, 500, 600]
    assert [i for k, i in it_of if k == "reset"] == [300, 600]
    assert [i for k, i in it_of if k == "save"] == [250, 500]
    assert log.checkpoints == ["/tmp/x/ckpt_000250.ply", "/tmp/x/ckpt_000500.ply"]
    ext = [c[1] for c in rec.calls if c[0] == "densify"]
    assert all(e == runtime.scene_extent(g) for e in ext)
    assert rec.calls[0] == ("upload", 50, 1000)
    assert log.sizes[-1] == 55 and len(log.losses) == 701
    # densify within an iteration comes before the reset (pipeline/mod.rs:240-246)
    i300 = [k for k, i in it_of if i == 300]
    assert i300 == ["densify", "reset"]
