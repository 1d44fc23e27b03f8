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
from scenes import bits, to_dev
from test_gpu_batch import KEYS, compare, per_view
from test_gpu_sampling import View, check_dense, compact_vs_dense

pytestmark = pytest.mark.gpu


class tuning:
    """Process-wide tuning fields for the duration of a with-block."""

    def __init__(self, **kw):
        self.kw = kw

    def __enter__(self):
        self.old = C.get_tuning()
        C.set_tuning(**dict(self.old, **self.kw))

    def __exit__(self, *exc):
        C.set_tuning(**self.old)


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


def dense_rays(v, mode, uv, min_live=1000):
    """The dense drop-in layout for a training ray subset (pixel coordinates
    uv [R, 2]) against the oracle."""
    sp_xyz, pc_idx, sp_t, mask = v.npts.render_image_hashtable_topk(
        v.cs, v.opt, v.idx_d, None, v.hash_d, mode=mode, training_ray_index=to_dev(uv, v.dev))
    torch.cuda.synchronize()
    rays = (uv[:, 1] * v.cam.width + uv[:, 0]).astype(np.int32)
    o_xyz, o_idx, o_t, o_mask = oracle.quick_sampling(v.pts, v.ocam, v.oopt, v.oi, v.oh, rays=rays,
                                                      mode=oracle.MODE_KNN if mode == "knn" else oracle.MODE_COMPAT)
    np.testing.assert_array_equal(mask.cpu().numpy().reshape(o_mask.shape), o_mask)
    np.testing.assert_array_equal(pc_idx.cpu().numpy().reshape(o_idx.shape), o_idx)
    assert np.array_equal(bits(sp_t.cpu().numpy().reshape(o_t.shape)), bits(o_t))
    assert np.array_equal(bits(sp_xyz.cpu().numpy().reshape(o_xyz.shape)), bits(o_xyz))
    assert o_mask.sum() > min_live


@pytest.mark.parametrize("form", ["lane", "wave"])
def test_ray_list_forms(dev, form):
    """Ray lists sampled one lane per ray (list_wave = 0, no deferral) and one
    wave per ray (list_wave above the list's length), each bit for bit
    against the oracle: the dense drop-in layout of a training batch, the
    compact layout with fp32 and bf16 features and rays off the image, short
    lists that sample without bucket cones, and per-point radii with a
    121-lane window (4-word lane sets)."""
    lanes = form == "lane"
    knobs = dict(list_wave=0, heavy=0) if lanes else dict(list_wave=1 << 30)
    pts, cam = synth.box_scene(400_000, seed=7), synth.look_at_camera([-2.0, -2.5, 2.0], 400, 300, 555.5)
    v = View(pts, cam, dev, radius_threshold=0.008, num_point_cloud_per_sp=8)
    rng = np.random.default_rng(11)
    uv = np.stack([rng.integers(0, cam.width, 4096), rng.integers(0, cam.height, 4096)], axis=1).astype(np.int64)
    rays = rng.integers(-5, cam.width * cam.height + 5, size=5000).astype(np.int32)
    feat = torch.randn(len(pts), 16, generator=torch.Generator().manual_seed(12)).to(dev)
    print(f"ray list form={form}", flush=True)
    with tuning(**knobs):
        dense_rays(v, "compat", uv)
        dense_rays(v, "knn", uv)
        compact_vs_dense(v, "knn", rays=rays, feat=feat)
        compact_vs_dense(v, "compat", rays=rays, feat=feat[:, :13].contiguous().to(torch.bfloat16))
        # lists this short (64 rays x 25 window pixels, under a quarter of the
        # view's pixels) skip the bucket cones and the pixel-ordered copy
        dense_rays(v, "compat", uv[:64], min_live=5)
        compact_vs_dense(v, "knn", rays=rays[:64], feat=feat)
    pts = synth.box_scene(300_000, n_boxes=40, seed=8)
    radius = synth.log_uniform_radius(len(pts), seed=9)
    cam = synth.look_at_camera([2.0, -2.0, 3.0], 320, 180, 444.0)
    v = View(pts, cam, dev, radius=radius, kernel_size=11, radius_threshold=0.01, num_point_cloud_per_sp=16)
    rays = rng.integers(0, cam.width * cam.height, size=3000).astype(np.int32)
    print(f"ray list form={form}: per-point radius, k = 11, K = 16", flush=True)
    with tuning(**knobs):
        compact_vs_dense(v, "knn", rays=rays)
