"""The large configuration at full size: 20M points on 200 boxes and the
ground plane, per-point radius LogUniform[0.001, 0.01], one 1920x1080 view
of the 16, a window of k = 15 sized from the largest radius (225 neighbour
pixels, culled with 4-word lane sets), K = 16, 32-dim bf16 features.  The hash table is compared with the
oracle in full; the sampling on a random set of rays (the oracle's cost) and,
over the whole image, through properties that hold at any size and the
agreement of the tiled and ray-list samplers."""
import numpy as np
import pytest
import torch

import epcq.csrc as C
from epcq import synth
from epcq.epcq import NeuralPoints, RenderOptions
from oracle import oracle
from scenes import bits, camera_spec, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def large(dev):
    print("generating the 20M-point cloud", flush=True)
    pts, radius, cams, kw = synth.workload_large()
    cam = cams[0]
    opt = RenderOptions(**kw)
    xyz = to_dev(pts, dev)
    npts = NeuralPoints(pc_type="pc", points=xyz)
    cs = camera_spec(cam, dev)
    idx, depth, table = npts.pts2plane(cs, opt)
    torch.cuda.synchronize()
    return dict(pts=pts, radius=radius, cam=cam, opt=opt, npts=npts, cs=cs, idx=idx, table=table,
                radius_d=to_dev(radius, dev))


def test_large_hash_table_equals_oracle(large):
    c = large["cam"]
    print("oracle hash table, 20M points", flush=True)
    oi, od, oh = oracle.pts2plane(large["pts"], c.w2c, c.K, c.height, c.width)
    assert np.array_equal(large["idx"].cpu().numpy(), oi)
    assert np.array_equal(large["table"].cpu().numpy().reshape(-1, 2), oh)


def test_large_sampling_rays_equal_oracle(large, dev):
    c, opt = large["cam"], large["opt"]
    rng = np.random.default_rng(3)
    rays = rng.integers(0, c.width * c.height, size=4000).astype(np.int32)
    feat = torch.randn(len(large["pts"]), 32, generator=torch.Generator().manual_seed(4)).to(torch.bfloat16).to(dev)
    print("knn_sampling on 4000 rays, per-point radius, K=16, bf16 features", flush=True)
    out = large["npts"].knn_sampling(large["cs"], opt, large["idx"], large["table"], rays=to_dev(rays, dev),
                                     radius=large["radius_d"], mode="knn", feat=feat)
    torch.cuda.synchronize()
    oi = large["idx"].cpu().numpy()
    oh = large["table"].cpu().numpy().reshape(-1, 2)
    o_xyz, o_idx, o_t, o_mask, o_nbt, o_nbd = oracle.quick_sampling(
        large["pts"], oracle.camera_from(c), oracle.options_from(opt), oi, oh, radius=large["radius"],
        mode=oracle.MODE_KNN, rays=rays, with_nb=True)
    s = out["sample_t"].shape[1]
    assert (o_mask[:, s:] == 0).all()
    assert np.array_equal(out["n_samples"].cpu().numpy(), o_mask.sum(1))
    assert np.array_equal(out["nb_idx"].cpu().numpy(), o_idx[:, :s])
    for k, ref in (("sample_t", o_t[:, :s]), ("nb_t", o_nbt[:, :s]), ("nb_dist", o_nbd[:, :s])):
        assert np.array_equal(bits(out[k].cpu().numpy()), bits(ref)), k
    # features: bf16 inputs, fp32 accumulation, 1e-3 relative
    fh = feat.cpu().view(torch.int16).numpy().view(np.uint16)
    w, agg = oracle.idw(large["pts"], fh, o_xyz[:, :s], o_idx[:, :s], bf16=True)
    live = o_mask[:, :s].reshape(-1)
    gw = out["w"].cpu().numpy().reshape(-1, w.shape[1])[live]
    ga = out["agg"].cpu().numpy().reshape(-1, agg.shape[1])[live]
    np.testing.assert_allclose(gw, w[live], rtol=1e-3, atol=1e-6)
    np.testing.assert_allclose(ga, agg[live], rtol=1e-3, atol=1e-4)
    assert live.sum() > 100


def test_large_full_image_properties(large, dev):
    c, opt = large["cam"], large["opt"]
    print("knn_sampling over the full 1920x1080 image", flush=True)
    full = large["npts"].knn_sampling(large["cs"], opt, large["idx"], large["table"], radius=large["radius_d"],
                                      mode="knn")
    torch.cuda.synchronize()
    n = full["n_samples"].cpu().numpy()
    t = full["sample_t"].cpu().numpy()
    idx = full["nb_idx"].cpu().numpy()
    s = t.shape[1]
    assert n.min() >= 0 and n.max() <= s and n.sum() > 0
    cols = np.arange(s)[None, :]
    live = cols < n[:, None]
    assert (t[~live] == 0).all() and (idx[~live] == -1).all()
    # samples ascend along each ray
    assert (np.diff(np.where(live, t, np.inf), axis=1)[live[:, 1:]] > 0).all()
    # neighbours: valid point ids
    valid = idx >= 0
    assert (idx[valid] < len(large["pts"])).all()
    # the tiled sampler agrees with the ray-list sampler on a random subset
    rng = np.random.default_rng(5)
    rays = rng.integers(0, c.width * c.height, size=20000).astype(np.int32)
    part = large["npts"].knn_sampling(large["cs"], opt, large["idx"], large["table"], rays=to_dev(rays, dev),
                                      radius=large["radius_d"], mode="knn")
    torch.cuda.synchronize()
    for k in ("n_samples", "sample_t", "nb_idx", "nb_t", "nb_dist"):
        a = full[k].cpu().numpy()[rays]
        b = part[k].cpu().numpy()
        assert np.array_equal(bits(a), bits(b)), k
