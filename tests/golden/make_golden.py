"""Writes tests/golden/c0_small.npz: inputs and the CPU oracle's outputs for a
small seeded workload (config-0 distributions, 300 Gaussians, 48x40 view).

Inputs: scene, camera, target. Outputs: rendered RGB, final T, last
contributor, tile lists (CSR of Gaussian indices), loss, GaussianGrad records,
and the parameters and Adam moments after one adam_step. The reference ships
no data files, so these vectors come from the oracle; the oracle itself is
pinned by the reference's own tests (tests/test_oracle_*.py).

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from rgsplat_amd import synthetic as S  # noqa: E402


def make():
    wl = S.scaled(S.WORKLOADS["c0_10k_256"], n=300, sizes=[(48, 40)])
    g = wl.scene()
    g[:, 2] *= 0.5  # pull the scene in so most of it lands in the small view
    cam = wl.cameras()[0]
    cam.fx = cam.fy = 60.0
    tg_scene = wl.target_scene()
    rgb_t = oracle.render_color(tg_scene, cam)
    target = np.concatenate([rgb_t, np.ones(rgb_t.shape[:2] + (1,), np.float32)], -1)
    r = oracle.Render(g, cam, target, lam=0.2)
    rgb, ft, last = r.image()
    off, gis = r.tiles()
    grads = r.grads()
    cfg = S.train_config()
    g1 = g.copy()
    st = oracle.AdamState(len(g1))
    oracle.adam_step(g1, grads, st, cfg)
    out = dict(scene=g, c2w=cam.c2w, intrinsics=np.array([cam.width, cam.height, cam.fx, cam.fy, cam.cx, cam.cy],
                                                          np.float64),
               target=target, rgb=rgb, final_t=ft, last_contrib=last, tile_offsets=off, tile_gis=gis,
               loss=np.float64(r.loss), grads=grads, adam_params=g1, adam_m1=st.m1, adam_m2=st.m2)
    path = os.path.join(ROOT, "tests", "golden", "c0_small.npz")
    np.savez_compressed(path, **out)
    return path


def load():
    d = dict(np.load(os.path.join(ROOT, "tests", "golden", "c0_small.npz")))
    w, h, fx, fy, cx, cy = d["intrinsics"]
    cam = S.Camera(int(w), int(h), float(fx), float(fy), float(cx), float(cy), d["c2w"])
    return d, cam


if __name__ == "__main__":
    print(make())
