"""The wave-cooperative form of the sampler (a ray with many candidates is
sampled by a whole wave) against the oracle.  The `heavy` tuning field sets
the candidate count above which a ray takes that form; heavy = 1 sends every
ray with more than one candidate through it, heavy = 0 none, so each case
below checks the wave form itself, bit for bit, on the layouts, modes and
window sizes the one-lane form is tested on."""
import numpy as np
import pytest
import torch

import epcq.csrc as C
from epcq import synth
from epcq.batch import MultiViewSampler
from epcq.epcq import RenderOptions
from oracle import oracle
from scenes import to_dev
from test_gpu_batch import KEYS, compare, per_view
from test_gpu_sampling import View, check_dense, compact_vs_dense

pytestmark = pytest.mark.gpu


class heavy_threshold:
    """Process-wide `heavy` tuning for the duration of a with-block."""

    def __init__(self, heavy):
        self.heavy = heavy

    def __enter__(self):
        self.old = C.get_tuning()
        C.set_tuning(**dict(self.old, heavy=self.heavy))

    def __exit__(self, *exc):
        C.set_tuning(**self.old)


@pytest.mark.parametrize("kw", [
    dict(kernel_size=3, radius_threshold=0.02, num_point_cloud_per_sp=4),
    dict(kernel_size=4, radius_threshold=0.03, num_point_cloud_per_sp=8),
    dict(kernel_size=7, radius_threshold=0.05, num_point_cloud_per_sp=12, num_sample_points_per_ray=4),
    dict(kernel_size=5, radius_threshold=0.03, sdf2weight_gaussian_gamma=0.2, sdf2weight_gaussian_alpha=0.01,
         num_point_cloud_per_sp=8),
    dict(kernel_size=5, radius_threshold=0.04, t_intvl=0.001, num_point_cloud_per_sp=8),
    dict(kernel_size=1, radius_threshold=0.05, num_point_cloud_per_sp=3),
])
@pytest.mark.parametrize("mode", ["compat", "knn"])
def test_wave_form_options(dev, kw, mode):
    pts = synth.sphere_cloud(40_000, seed=5)
    cam = synth.look_at_camera([0.3, -3.5, 0.4], 90, 70, 110.0)  # the sphere touches the borders
    print(f"wave form {mode} {kw}", flush=True)
    v = View(pts, cam, dev, **kw)
    with heavy_threshold(1):
        check_dense(v, mode)
        check_dense(v, mode, cull=False)
        compact_vs_dense(v, mode)


@pytest.mark.parametrize("heavy", [1, 64])
@pytest.mark.parametrize("mode", ["compat", "knn"])
def test_wave_form_dense_view(dev, mode, heavy):
    """A 200x150 view of the 1M-point scene (~30 points per pixel, windows of
    up to thousands of candidates): the densest rays go to the wave form,
    the rest stay one per lane, within the same tiles."""
    pts = synth.box_scene(1_000_000, seed=0)
    cam = synth.look_at_camera([-2.1, 2.9, 1.7], 200, 150, synth.nerf_focal(200))
    print(f"wave form dense view {mode} heavy={heavy}", flush=True)
    v = View(pts, cam, dev, radius_threshold=0.005, num_point_cloud_per_sp=8)
    with heavy_threshold(heavy):
        check_dense(v, mode)
        compact_vs_dense(v, mode)


def test_wave_form_ray_list_and_features(dev):
    pts, cam = synth.box_scene(400_000, seed=7), synth.look_at_camera([-2.0, -2.5, 2.0], 400, 300, 555.5)
    v = View(pts, cam, dev, radius_threshold=0.008, num_point_cloud_per_sp=8)
    rays = np.random.default_rng(1).integers(-5, 400 * 300 + 5, size=20_000).astype(np.int32)
    feat = torch.randn(len(pts), 16, generator=torch.Generator().manual_seed(1)).to(dev)
    print("wave form: ray list + fp32 / bf16 features", flush=True)
    with heavy_threshold(1):
        compact_vs_dense(v, "knn", rays=rays, feat=feat)
        compact_vs_dense(v, "compat", rays=rays, feat=feat[:, :13].contiguous().to(torch.bfloat16))


def test_wave_form_per_point_radius_k16(dev):
    pts = synth.box_scene(300_000, n_boxes=40, seed=8)
    radius = synth.log_uniform_radius(len(pts), seed=9)
    cam = synth.look_at_camera([2.0, -2.0, 3.0], 320, 180, 444.0)
    v = View(pts, cam, dev, radius=radius, kernel_size=7, radius_threshold=0.01, num_point_cloud_per_sp=16)
    feat = torch.randn(len(pts), 32, generator=torch.Generator().manual_seed(2)).to(torch.bfloat16).to(dev)
    print("wave form: per-point radius K=16 + bf16 features", flush=True)
    with heavy_threshold(1):
        compact_vs_dense(v, "knn", feat=feat)
        compact_vs_dense(v, "knn", cull=False)


@pytest.mark.parametrize("mode", ["compat", "knn"])
def test_wave_form_batch(dev, mode):
    """The batched pass with heavy = 1 equals the per-view path with heavy =
    0 (one ray per lane), view by view."""
    box = to_dev(synth.box_scene(400_000, seed=1), dev)
    opt = RenderOptions(radius_threshold=0.005, num_point_cloud_per_sp=8)
    cams = synth.random_views(6, seed_eye=5, seed_res=6)
    pairs = [(box, c) for c in cams]
    print(f"wave form batch {mode}", flush=True)
    tuning = dict(C.get_tuning(), heavy=1)
    sampler = MultiViewSampler(opt, mode=mode, device=dev, tuning=tuning)
    sampler.set_views(pairs)
    sampler.run()
    torch.cuda.synchronize()
    with heavy_threshold(0):
        for i, (xyz, cam) in enumerate(pairs):
            compare(sampler.view_outputs(i), per_view(xyz, cam, opt, mode, dev), KEYS)
    # and one view straight against the oracle
    xyz, cam = pairs[2]
    pts = xyz.cpu().numpy()
    oi, od, oh = oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width)
    o = sampler.view_outputs(2)
    res = oracle.quick_sampling(pts, oracle.camera_from(cam), oracle.options_from(opt), oi, oh,
                                mode=oracle.MODE_KNN if mode == "knn" else oracle.MODE_COMPAT, with_nb=True)
    s = o["sample_t"].shape[1]
    assert np.array_equal(o["n_samples"].cpu().numpy(), res[3].sum(1))
    assert np.array_equal(o["nb_idx"].cpu().numpy(), res[1][:, :s])
    assert np.array_equal(o["sample_t"].cpu().numpy().view(np.int32), res[2][:, :s].view(np.int32))
