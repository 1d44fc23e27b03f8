"""Per-point radii on the GPU against the oracle: radii that are zero,
negative, NaN, infinite, beyond the float range when squared or below it
(a point with such a radius is never or always within reach, exactly as
`dist < radius` in the oracle decides), in the one-lane and the
wave-cooperative forms, and per-point radii in the batched pass."""
import numpy as np
import pytest
import torch

from epcq import synth
from epcq.batch import MultiViewSampler
from epcq.epcq import NeuralPoints, RenderOptions
from oracle import oracle
from scenes import camera_spec, to_dev
from test_gpu_batch import KEYS, compare
from test_gpu_sampling import View, compact_vs_dense
from test_gpu_wave import heavy_threshold

pytestmark = pytest.mark.gpu

EDGE_RADII = np.array([0.0, -1.0, np.nan, np.inf, 1e30, 1e-40, 1e-23, 0.3, np.finfo(np.float32).max],
                      dtype=np.float32)


def edge_radius(n, seed):
    """Log-uniform radii in [0.01, 0.05] with every 37th point given one of
    EDGE_RADII."""
    r = synth.log_uniform_radius(n, seed=seed, lo=0.01, hi=0.05)
    at = np.arange(0, n, 37)
    r[at] = EDGE_RADII[np.arange(len(at)) % len(EDGE_RADII)]
    return r


@pytest.mark.parametrize("heavy", [0, 1])
@pytest.mark.parametrize("mode", ["compat", "knn"])
def test_edge_radii_vs_oracle(dev, mode, heavy):
    pts = synth.sphere_cloud(40_000, seed=5)
    radius = edge_radius(len(pts), seed=3)
    cam = synth.look_at_camera([0.3, -3.5, 0.4], 90, 70, 110.0)
    print(f"edge radii {mode} heavy={heavy}", flush=True)
    v = View(pts, cam, dev, radius=radius, kernel_size=5, radius_threshold=0.02, num_point_cloud_per_sp=8)
    feat = torch.randn(len(pts), 8, generator=torch.Generator().manual_seed(4)).to(dev)
    with heavy_threshold(heavy):
        out = compact_vs_dense(v, mode, feat=feat)
        compact_vs_dense(v, mode, cull=False)
    # the points with an infinite radius are reached by every ray whose
    # window holds their pixel, so the view has live samples
    assert out["n_samples"].sum().item() > 1000


def per_view_radius(xyz_d, radius_d, cam, opt, mode, dev):
    npts = NeuralPoints(pc_type="pc", points=xyz_d)
    cs = camera_spec(cam, dev)
    idx, _, table = npts.pts2plane(cs, opt)
    out = npts.knn_sampling(cs, opt, idx, table, radius=radius_d, mode=mode)
    out["sorted_idx"], out["hash"] = idx, table
    return out


@pytest.mark.parametrize("mode", ["compat", "knn"])
def test_batch_per_point_radius(dev, mode):
    """The batched pass with per-point radii equals the per-view path, view
    by view, and one view equals the oracle."""
    pts = synth.box_scene(300_000, n_boxes=40, seed=8)
    radius = synth.log_uniform_radius(len(pts), seed=9)
    box, rad_d = to_dev(pts, dev), to_dev(radius, dev)
    opt = RenderOptions(radius_threshold=0.01, kernel_size=7, num_point_cloud_per_sp=16)
    cams = synth.random_views(5, seed_eye=11, seed_res=12)
    pairs = [(box, c) for c in cams]
    print(f"batch with per-point radii, {mode}", flush=True)
    sampler = MultiViewSampler(opt, mode=mode, device=dev, radius=rad_d)
    sampler.set_views(pairs)
    sampler.run()
    torch.cuda.synchronize()
    for i, (xyz, cam) in enumerate(pairs):
        compare(sampler.view_outputs(i), per_view_radius(xyz, rad_d, cam, opt, mode, dev), KEYS)
    cam = cams[0]
    oi, od, oh = oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width)
    res = oracle.quick_sampling(pts, oracle.camera_from(cam), oracle.options_from(opt), oi, oh, radius=radius,
                                mode=oracle.MODE_KNN if mode == "knn" else oracle.MODE_COMPAT, with_nb=True)
    o = sampler.view_outputs(0)
    s = o["sample_t"].shape[1]
    n = o["n_samples"].cpu().numpy()
    assert np.array_equal(n, res[3].sum(1)) and n.sum() > 1000
    assert np.array_equal(o["nb_idx"].cpu().numpy(), res[1][:, :s])
    assert np.array_equal(o["sample_t"].cpu().numpy().view(np.int32), res[2][:, :s].view(np.int32))
