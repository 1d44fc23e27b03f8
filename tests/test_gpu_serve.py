"""Serving loop: each request's result is the same alone and in a mixed,
continuously batched stream (and equals the per-view path); admission by
footprint; batches on several slots overlap."""
import numpy as np
import pytest
import torch

from epcq import synth
from epcq.epcq import NeuralPoints, RenderOptions
from epcq.serve import PointServer, poisson_trace, run_trace
from scenes import camera_spec, to_dev

pytestmark = pytest.mark.gpu

RES = [(200, 150), (400, 300), (320, 256), (160, 120)]


def clouds(dev):
    return {
        "sphere": to_dev(synth.sphere_cloud(60_000, seed=10), dev),
        "box20": to_dev(synth.box_scene(200_000, n_boxes=20, seed=11), dev),
        "box40": to_dev(synth.box_scene(300_000, n_boxes=40, seed=12), dev),
    }


def same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        assert x.shape == y.shape, k
        if x.dtype == np.float32:
            x, y = x.view(np.int32), y.view(np.int32)
        assert np.array_equal(x, y), k


def per_view(xyz, cam, opt, mode, dev):
    npts = NeuralPoints(pc_type="pc", points=xyz)
    cs = camera_spec(cam, dev)
    idx, _, table = npts.pts2plane(cs, opt)
    return npts.knn_sampling(cs, opt, idx, table, mode=mode)


@pytest.mark.parametrize("mode", ["compat", "knn"])
def test_alone_equals_in_mix(dev, mode):
    cl = clouds(dev)
    opt = RenderOptions(radius_threshold=0.005, num_point_cloud_per_sp=8)
    trace = poisson_trace(2000.0, 24, cl, seed=40, resolutions=RES)
    probe_t, probe_cloud, probe_cam = trace[9]
    print(f"serve {mode}: 24 requests, probe {probe_cloud} {probe_cam.width}x{probe_cam.height}", flush=True)
    # alone
    server = PointServer(cl, opt, mode=mode, slots=3, slot_bytes=96 << 20, max_views=8)
    r = server.submit(probe_cloud, probe_cam)
    server.drain()
    alone = {k: v.clone() for k, v in r.result.items()}
    server.close()
    # in the mix, arriving while other batches are on the GPU
    server = PointServer(cl, opt, mode=mode, slots=3, slot_bytes=96 << 20, max_views=8)
    reqs = run_trace(server, trace)
    assert len(reqs) == len(trace) and all(q.done for q in reqs)
    assert server.launches > 1
    torch.cuda.synchronize()
    same(reqs[9].result, alone)
    # every request equals the per-view path on its view
    for q in reqs[::5]:
        same(q.result, {k: v for k, v in per_view(cl[q.cloud], q.cam, opt, mode, dev).items()})
    server.close()


def test_admission_and_overlap(dev):
    cl = clouds(dev)
    opt = RenderOptions(radius_threshold=0.005, num_point_cloud_per_sp=8)
    big = synth.look_at_camera([0.0, -4.0, 0.0], 1600, 1200, synth.nerf_focal(1600))
    server = PointServer(cl, opt, mode="compat", slots=2, slot_bytes=64 << 20, max_views=4)
    fp = server.footprint("box40", big)
    assert fp > server.budget
    with pytest.raises(RuntimeError, match="exceeds a slot's arena"):
        server.submit("box40", big)
    with pytest.raises(RuntimeError, match="unknown cloud"):
        server.submit("nope", big)
    # a burst: batches are bounded by max_views and by the arena, both slots busy
    cams = synth.random_views(20, seed_eye=50, seed_res=51, resolutions=RES)
    reqs = [server.submit("box20" if i % 2 else "sphere", c) for i, c in enumerate(cams)]
    server.drain()
    assert all(q.done for q in reqs)
    assert server.launches >= 5
    assert server.max_inflight == 2
    per_batch = {}
    for q in reqs:
        per_batch.setdefault((q.slot, q.launched), []).append(q)
    for batch in per_batch.values():
        assert len(batch) <= 4
        assert sum(q.footprint for q in batch) <= server.budget
    same(reqs[3].result, per_view(cl["box20"], cams[3], opt, "compat", dev))
    server.close()
