"""Every launch shape the autotuner can pick, against the oracle: tile
shapes, staging capacities (which decide which tiles overflow the LDS and
read global memory), the XCD tile order and the outlier threshold, on the
per-view sampler (whole image, bit for bit) and on the batched pass (every
view's hash table in full and random rays of every view)."""
import numpy as np
import pytest
import torch

import epcq.csrc as C
from epcq import autotune, synth
from epcq.batch import MultiViewSampler
from epcq.epcq import RenderOptions
from oracle import oracle
from scenes import to_dev
from test_gpu_configs import check_rows, oracle_rows
from test_gpu_sampling import View, check_dense

pytestmark = pytest.mark.gpu


def tunings():
    out = []
    for tw, th in autotune.SHAPES:
        out.append(dict(autotune.DEFAULT, tile_w=tw, tile_h=th))
    for cap in autotune.CAPS:
        for tw, th in ((32, 8), (16, 16)):
            out.append(dict(autotune.DEFAULT, tile_w=tw, tile_h=th, tile_cap=cap))
    for fl in autotune.FLAGS:
        for hv in autotune.HEAVY:
            out.append(dict(autotune.DEFAULT, tile_w=16, tile_h=16, flags=fl, heavy=hv))
    uniq = []
    for t in out:
        if t not in uniq:
            uniq.append(t)
    return uniq


def name(t):
    return f'{t["tile_w"]}x{t["tile_h"]}-cap{t["tile_cap"]}-f{t["flags"]}-h{t["heavy"]}'


@pytest.fixture(scope="module")
def dense_view(dev):
    # ~30 points per pixel: many tiles overflow small staging capacities
    pts = synth.box_scene(1_000_000, seed=0)
    cam = synth.look_at_camera([-2.1, 2.9, 1.7], 200, 150, synth.nerf_focal(200))
    return View(pts, cam, dev, radius_threshold=0.005, num_point_cloud_per_sp=8)


@pytest.mark.parametrize("tuning", tunings(), ids=name)
def test_per_view_tuning(dev, dense_view, tuning):
    old = C.get_tuning()
    print(f"per-view sampler, tuning {name(tuning)}", flush=True)
    try:
        autotune.apply(tuning)
        check_dense(dense_view, "compat")
        check_dense(dense_view, "knn")
    finally:
        C.set_tuning(**old)


@pytest.fixture(scope="module")
def batch_scene(dev):
    pts = synth.box_scene(400_000, seed=1)
    cams = synth.random_views(6, seed_eye=31, seed_res=32)
    cams.append(synth.look_at_camera([-2.1, 2.9, 1.7], 200, 150, synth.nerf_focal(200)))
    refs = []
    for c in cams:
        oi, _, oh = oracle.pts2plane(pts, c.w2c, c.K, c.height, c.width)
        refs.append((oi, oh))
    return pts, to_dev(pts, dev), cams, refs


@pytest.mark.parametrize("tuning", tunings(), ids=name)
def test_batch_tuning(dev, batch_scene, tuning):
    pts, xyz, cams, refs = batch_scene
    opt = RenderOptions(radius_threshold=0.005, num_point_cloud_per_sp=8)
    print(f"batched pass, tuning {name(tuning)}", flush=True)
    sampler = MultiViewSampler(opt, mode="knn", device=dev, max_views=8, tuning=tuning)
    sampler.set_views([(xyz, c) for c in cams])
    sampler.run()
    torch.cuda.synchronize()
    rng = np.random.default_rng(5)
    for i, c in enumerate(cams):
        o = sampler.view_outputs(i)
        oi, oh = refs[i]
        assert np.array_equal(o["sorted_idx"].cpu().numpy(), oi)
        assert np.array_equal(o["hash"].cpu().numpy(), oh)
        rays = rng.integers(0, c.width * c.height, size=400).astype(np.int32)
        check_rows(o, oracle_rows(pts, c, opt, "knn", rays, oi, oh), rays, "knn")
    sampler.close()
