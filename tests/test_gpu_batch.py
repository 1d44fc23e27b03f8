"""Batched multi-view sampling: every view's hash table, point lists and
compact samples equal the per-view path (itself checked against the oracle)
and the oracle directly; HIP-graph replays equal the eager pass, also after
the view set changes."""
import numpy as np
import pytest
import torch

import epcq.csrc as C
from epcq import synth
from epcq.batch import MultiViewSampler
from epcq.epcq import NeuralPoints, RenderOptions
from oracle import oracle
from scenes import camera_spec, to_dev

pytestmark = pytest.mark.gpu


def per_view(xyz_d, cam, opt, mode, dev):
    npts = NeuralPoints(pc_type="pc", points=xyz_d)
    cs = camera_spec(cam, dev)
    idx, _, table = npts.pts2plane(cs, opt)
    out = npts.knn_sampling(cs, opt, idx, table, mode=mode)
    out["sorted_idx"], out["hash"] = idx, table
    return out


def compare(a, b, keys):
    for k in keys:
        if k not in a:
            continue
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        assert x.shape == y.shape, k
        assert np.array_equal(x.view(np.int32) if x.dtype == np.float32 else x,
                              y.view(np.int32) if y.dtype == np.float32 else y), k


KEYS = ["sorted_idx", "hash", "n_samples", "sample_t", "nb_idx", "nb_t", "nb_dist"]


def mixed_views(seed, count):
    cams = synth.random_views(count, seed_eye=seed, seed_res=seed + 1)
    return cams


@pytest.mark.parametrize("mode", ["compat", "knn"])
def test_batch_equals_per_view(dev, mode):
    box = to_dev(synth.box_scene(400_000, seed=1), dev)
    sph = to_dev(synth.sphere_cloud(100_000, seed=2), dev)
    opt = RenderOptions(radius_threshold=0.005, num_point_cloud_per_sp=8)
    cams = mixed_views(5, 6)
    pairs = [(box, c) for c in cams] + [(sph, synth.look_at_camera([0.0, -4.0, 0.0], 320, 240, 444.0))]
    print(f"batch of {len(pairs)} views, {mode}", flush=True)
    sampler = MultiViewSampler(opt, mode=mode, device=dev)
    sampler.set_views(pairs)
    sampler.run()
    torch.cuda.synchronize()
    assert sampler.rays == sum(c.width * c.height for _, c in pairs)
    for i, (xyz, cam) in enumerate(pairs):
        ref = per_view(xyz, cam, opt, mode, dev)
        compare(sampler.view_outputs(i), ref, KEYS)
    # one view straight against the oracle
    xyz, cam = pairs[1]
    pts = xyz.cpu().numpy()
    oi, od, oh = oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width)
    o = sampler.view_outputs(1)
    assert np.array_equal(o["sorted_idx"].cpu().numpy(), oi)
    assert np.array_equal(o["hash"].cpu().numpy(), oh)
    res = oracle.quick_sampling(pts, oracle.camera_from(cam), oracle.options_from(opt), oi, oh,
                                mode=oracle.MODE_KNN if mode == "knn" else oracle.MODE_COMPAT, with_nb=True)
    s = o["sample_t"].shape[1]
    assert np.array_equal(o["n_samples"].cpu().numpy(), res[3].sum(1))
    assert np.array_equal(o["nb_idx"].cpu().numpy(), res[1][:, :s])
    assert np.array_equal(o["sample_t"].cpu().numpy().view(np.int32), res[2][:, :s].view(np.int32))


def test_batch_graph_replay_and_new_views(dev):
    box = to_dev(synth.box_scene(300_000, seed=3), dev)
    opt = RenderOptions(radius_threshold=0.005, num_point_cloud_per_sp=8)
    sampler = MultiViewSampler(opt, mode="knn", device=dev, max_views=64)
    first = [(box, c) for c in mixed_views(7, 5)]
    second = [(box, c) for c in mixed_views(9, 3)]
    sampler.reserve(max(sampler.required_bytes(first), sampler.required_bytes(second)))
    sampler.set_views(first)
    sampler.run()
    torch.cuda.synchronize()
    eager = [{k: v.clone() for k, v in sampler.view_outputs(i).items()} for i in range(len(first))]
    print("graph capture", flush=True)
    sampler.capture()
    for rep in range(3):
        for i in range(len(first)):  # clobber the outputs between replays
            for t in sampler.view_outputs(i).values():
                t.fill_(7)
        sampler.replay()
        torch.cuda.synchronize()
        x = torch.randn(1000, device=dev).sum().item()  # eager work between replays
        assert np.isfinite(x)
        for i in range(len(first)):
            compare(sampler.view_outputs(i), eager[i], KEYS)
    # a different, ragged view set through the same graph
    sampler.set_views(second)
    sampler.replay()
    torch.cuda.synchronize()
    for i, (xyz, cam) in enumerate(second):
        compare(sampler.view_outputs(i), per_view(xyz, cam, opt, "knn", dev), KEYS)
