"""The benchmark configurations at their full sizes against the oracle.

- The 1M-point box scene, one 800x800 view, r = 0.005, K = 8, with 16-dim
  fp32 point features and inverse-distance weights: the whole image through
  the tiled sampler, 4000 random rays of it against the oracle bit for bit.
- The mixed 64-view batch of the same cloud through one batched pass, with
  the bench's launch shape: every view's point lists and hash table against
  the oracle in full, and 500 random rays of every view.
- Serving: the four resident clouds (0.5M sphere, 1M / 20 boxes, 2M / 40
  boxes, 5M / 80 boxes), requests of every resolution of the batch's set
  arriving by a Poisson process into the continuous-batching server; every
  request's result on 300 random rays against the oracle, and a request
  served alone equal to the same request in the mix.
The oracle's cost bounds the ray subsets; every ray of an image is computed
independently, so a random subset checks the same code paths as the whole."""
import numpy as np
import pytest
import torch

from epcq import synth
from epcq.batch import MultiViewSampler
from epcq.epcq import NeuralPoints, RenderOptions
from epcq.serve import PointServer, poisson_trace, run_trace
from oracle import oracle
from scenes import bits, camera_spec, to_dev

pytestmark = pytest.mark.gpu

BENCH_TUNING = {"tile_w": 16, "tile_h": 16, "tile_cap": 1024, "flags": 1, "heavy": 256}


def oracle_rows(pts, cam, opt, mode, rays, oi=None, oh=None):
    if oi is None:
        oi, _, oh = oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width)
    return oracle.quick_sampling(pts, oracle.camera_from(cam), oracle.options_from(opt), oi, oh,
                                 mode=oracle.MODE_KNN if mode == "knn" else oracle.MODE_COMPAT, rays=rays,
                                 with_nb=True)


def check_rows(out, res, rays, mode):
    """Compact outputs `out` (whole image) at `rays` against the oracle's rows."""
    o_xyz, o_idx, o_t, o_mask, o_nbt, o_nbd = res
    s = out["sample_t"].shape[1]
    assert (o_mask[:, s:] == 0).all()
    r = torch.from_numpy(rays.astype(np.int64)).to(out["n_samples"].device)
    assert np.array_equal(out["n_samples"][r].cpu().numpy(), o_mask.sum(1))
    assert np.array_equal(out["nb_idx"][r].cpu().numpy(), o_idx[:, :s])
    assert np.array_equal(bits(out["sample_t"][r].cpu().numpy()), bits(o_t[:, :s]))
    if mode == "knn":
        assert np.array_equal(bits(out["nb_t"][r].cpu().numpy()), bits(o_nbt[:, :s]))
        assert np.array_equal(bits(out["nb_dist"][r].cpu().numpy()), bits(o_nbd[:, :s]))
    return o_mask


@pytest.mark.parametrize("mode", ["compat", "knn"])
def test_box_scene_800x800_features(dev, mode):
    pts, cams, kw = synth.workload_box_scene()
    cam = cams[0]
    opt = RenderOptions(**kw)
    npts = NeuralPoints(pc_type="pc", points=to_dev(pts, dev))
    cs = camera_spec(cam, dev)
    idx, _, table = npts.pts2plane(cs, opt)
    feat = torch.randn(len(pts), 16, generator=torch.Generator().manual_seed(1)).to(dev)
    print(f"box scene 800x800 {mode}, 16-dim f32 features, whole image", flush=True)
    out = npts.knn_sampling(cs, opt, idx, table, mode=mode, feat=feat)
    torch.cuda.synchronize()
    rays = np.random.default_rng(11).integers(0, cam.width * cam.height, size=4000).astype(np.int32)
    oi, oh = idx.cpu().numpy(), table.cpu().numpy()
    res = oracle_rows(pts, cam, opt, mode, rays, oi, oh)
    o_mask = check_rows(out, res, rays, mode)
    s = out["sample_t"].shape[1]
    w, agg = oracle.idw(pts, feat.cpu().numpy(), res[0][:, :s], res[1][:, :s])
    live = o_mask[:, :s].reshape(-1)
    r = torch.from_numpy(rays.astype(np.int64)).to(dev)
    gw = out["w"][r].cpu().numpy().reshape(-1, w.shape[1])
    ga = out["agg"][r].cpu().numpy().reshape(-1, agg.shape[1])
    assert np.array_equal(bits(gw[live]), bits(w[live]))
    assert np.array_equal(bits(ga[live]), bits(agg[live]))
    assert (gw[~live] == 0).all() and (ga[~live] == 0).all()
    assert live.sum() > 1000


@pytest.mark.parametrize("tuning", [BENCH_TUNING, None])
def test_batched_64_views_every_view(dev, tuning):
    pts, cams, kw = synth.workload_batched()
    opt = RenderOptions(**kw)
    xyz = to_dev(pts, dev)
    sampler = MultiViewSampler(opt, mode="compat", device=dev, max_views=64, tuning=tuning)
    sampler.set_views([(xyz, c) for c in cams])
    print(f"batched pass, 64 views, tuning {tuning}", flush=True)
    sampler.capture()
    sampler.replay()
    torch.cuda.synchronize()
    assert sampler.rays == sum(c.width * c.height for c in cams)
    rng = np.random.default_rng(21)
    for i, cam in enumerate(cams):
        o = sampler.view_outputs(i)
        oi, _, oh = oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width)
        assert np.array_equal(o["sorted_idx"].cpu().numpy(), oi), i
        assert np.array_equal(o["hash"].cpu().numpy(), oh), i
        rays = rng.integers(0, cam.width * cam.height, size=500).astype(np.int32)
        check_rows(o, oracle_rows(pts, cam, opt, "compat", rays, oi, oh), rays, "compat")
    sampler.close()


def serving_clouds(dev):
    return {
        "sphere-0.5M": to_dev(synth.sphere_cloud(500_000, seed=10), dev),
        "box20-1M": to_dev(synth.box_scene(1_000_000, n_boxes=20, seed=11), dev),
        "box40-2M": to_dev(synth.box_scene(2_000_000, n_boxes=40, seed=12), dev),
        "box80-5M": to_dev(synth.box_scene(5_000_000, n_boxes=80, seed=13), dev),
    }


def test_serving_four_clouds_against_oracle(dev):
    print("serving: generating the four clouds", flush=True)
    clouds = serving_clouds(dev)
    host = {k: v.cpu().numpy() for k, v in clouds.items()}
    opt = RenderOptions(radius_threshold=0.005, num_point_cloud_per_sp=8, topK=8)
    trace = poisson_trace(3000.0, 40, clouds, seed=20)
    res = {(c.width, c.height) for _, _, c in trace}
    assert len(res) == len(synth.RESOLUTIONS), res  # every resolution of the set appears
    server = PointServer(clouds, opt, mode="compat", slots=4, slot_bytes=1 << 30, max_views=8,
                         tuning=BENCH_TUNING)
    reqs = run_trace(server, trace)
    torch.cuda.synchronize()
    assert all(q.done for q in reqs) and server.launches > 1 and server.max_inflight > 1
    print(f"serving: {len(reqs)} requests in {server.launches} batches; checking each against the oracle",
          flush=True)
    rng = np.random.default_rng(23)
    for q in reqs:
        cam = q.cam
        rays = rng.integers(0, cam.width * cam.height, size=300).astype(np.int32)
        check_rows(q.result, oracle_rows(host[q.cloud], cam, opt, "compat", rays), rays, "compat")
    # the same request alone
    probe = reqs[len(reqs) // 2]
    mixed = {k: v.clone() for k, v in probe.result.items()}
    server.close()
    alone_server = PointServer(clouds, opt, mode="compat", slots=1, slot_bytes=1 << 30, max_views=8,
                               tuning=BENCH_TUNING)
    r = alone_server.submit(probe.cloud, probe.cam)
    alone_server.drain()
    torch.cuda.synchronize()
    for k in mixed:
        assert np.array_equal(bits(r.result[k].cpu().numpy()), bits(mixed[k].cpu().numpy())), k
    alone_server.close()
