"""Windows wider than 64 lanes (kernel_size 9..16, up to 256 neighbour
pixels) with exact lane-set culling in both sampler forms, and beyond 16
without lane sets, against the oracle bit for bit: dense and compact
layouts, both neighbour modes, per-point radii; the wave-wide form forced
for every multi-candidate ray (heavy = 1) and never used (heavy = 0)."""
import numpy as np
import pytest

import epcq.csrc as C
from epcq import synth
from test_gpu_sampling import View, check_dense, compact_vs_dense
from test_gpu_wave import heavy_threshold

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [9, 11, 15, 16, 17])
@pytest.mark.parametrize("mode", ["compat", "knn"])
@pytest.mark.parametrize("heavy", [0, 1])
def test_wide_window(dev, k, mode, heavy):
    pts = synth.box_scene(300_000, n_boxes=20, seed=8)
    cam = synth.look_at_camera([2.0, -2.6, 1.9], 120, 90, synth.nerf_focal(120) * 2.0)
    print(f"wide window k={k} {mode} heavy={heavy}", flush=True)
    v = View(pts, cam, dev, kernel_size=k, radius_threshold=0.012, num_point_cloud_per_sp=8)
    with heavy_threshold(heavy):
        live = check_dense(v, mode)
        compact_vs_dense(v, mode)
    assert live.sum() > 1000


@pytest.mark.parametrize("heavy", [0, 1])
def test_wide_window_per_point_radius(dev, heavy):
    pts = synth.box_scene(200_000, n_boxes=20, seed=9)
    radius = synth.log_uniform_radius(len(pts), seed=10, lo=2e-3, hi=2e-2)
    cam = synth.look_at_camera([-2.2, 2.4, 2.0], 100, 80, synth.nerf_focal(100) * 2.0)
    print(f"wide window per-point radius heavy={heavy}", flush=True)
    v = View(pts, cam, dev, radius=radius, kernel_size=13, radius_threshold=0.02, num_point_cloud_per_sp=16)
    with heavy_threshold(heavy):
        compact_vs_dense(v, "knn")
        rays = np.random.default_rng(4).integers(0, cam.width * cam.height, size=3000).astype(np.int32)
        compact_vs_dense(v, "knn", rays=rays)


class launch_shape:
    """Process-wide launch shape for the duration of a with-block."""

    def __init__(self, **kw):
        self.kw = kw

    def __enter__(self):
        self.old = C.get_tuning()
        C.set_tuning(**dict(self.old, **self.kw))

    def __exit__(self, *exc):
        C.set_tuning(**self.old)


@pytest.mark.parametrize("cap", [6000, 1024])
@pytest.mark.parametrize("mode", ["knn", "compat"])
def test_wide_window_staging_capacity(dev, cap, mode):
    """k = 15 with per-point radii through 16x16 tiles: a 6000-point request
    is clamped to what fits 64 KB of LDS with the tile's hash entries and
    cones (about 1800 points), a 1024-point one overflows more often and
    reads candidates from global memory; both against the oracle."""
    pts = synth.box_scene(250_000, n_boxes=20, seed=11)
    radius = synth.log_uniform_radius(len(pts), seed=12, lo=2e-3, hi=2e-2)
    cam = synth.look_at_camera([2.4, 2.0, 1.8], 96, 80, synth.nerf_focal(96) * 2.0)
    print(f"wide window k=15 staging cap={cap} {mode}", flush=True)
    v = View(pts, cam, dev, radius=radius, kernel_size=15, radius_threshold=0.02, num_point_cloud_per_sp=16)
    with launch_shape(tile_w=16, tile_h=16, tile_cap=cap, heavy=0):
        check_dense(v, mode)
        compact_vs_dense(v, mode)
