"""The reference renderer's whole chain after the point cloud
(opt/render_image.py:58-116) on this package: pts2plane, sampling of a
training ray subset, FeatureQuery, Aggregator, clamp_t and render.  The
sampling is the oracle's bit for bit (checked here again); the image is
compared with the same chain run through the fp32 torch Aggregator and the
oracle's render, within the Aggregator's bf16 tolerance carried through the
compositing (5e-3 per sample colour and alpha: 2e-2 per pixel)."""
import numpy as np
import pytest
import torch

from aggregator_ref import aggregator_reference
from epcq import synth
from epcq.epcq import Aggregator, FeatureQuery, NeuralPoints, RenderOptions, clamp_t, render
from oracle import oracle
from scenes import bits, camera_spec, to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("subset", [True, False])
def test_render_chain(dev, subset):
    pts = synth.box_scene(150_000, n_boxes=12, seed=21)
    cam = synth.look_at_camera([2.2, -2.4, 2.0], 128, 96, synth.nerf_focal(128) * 1.6)
    N = len(pts)
    g = torch.Generator().manual_seed(22)
    npts = NeuralPoints(pc_type="pc", points=to_dev(pts, dev))
    npts.rgb_data = torch.nn.Parameter(torch.rand((N, 3), generator=g).to(dev), requires_grad=False)
    d = torch.randn((N, 3), generator=g)
    npts.dir_data = torch.nn.Parameter((d / d.norm(dim=-1, keepdim=True)).to(dev), requires_grad=False)
    npts.feat_data = torch.nn.Parameter(torch.randn((N, 32), generator=g).to(dev), requires_grad=False)
    opt = RenderOptions(radius_threshold=0.01, num_point_cloud_per_sp=8, num_sample_points_per_ray=8)
    cs = camera_spec(cam, dev)
    rng = np.random.default_rng(23)
    uv = None
    if subset:
        uv = np.stack([rng.integers(0, cam.width, 3000), rng.integers(0, cam.height, 3000)], axis=1)
        uv = torch.from_numpy(uv.astype(np.int64))
    torch.manual_seed(24)
    agg = Aggregator()
    with torch.no_grad():
        for m in agg.modules():
            if isinstance(m, torch.nn.Linear):
                m.bias.normal_(0.0, 0.1)
    agg = agg.to(dev)
    print(f"render chain subset={subset}", flush=True)
    with torch.no_grad():
        idx, depth, table = npts.pts2plane(cs, opt)
        sp_xyz, pc_idx, sp_t, mask = npts.render_image_hashtable_topk(
            cs, opt, idx, depth, table, training_ray_index=uv.to(dev) if subset else None)
        fq = FeatureQuery()(npts, sp_xyz, pc_idx, mask, cs, uv.to(dev) if subset else torch.zeros((0, 2)))
        alpha, rgb = agg(*fq[:6], mask, fq[6])
        rows = alpha.shape[0]
        t = sp_t.reshape(rows, -1)
        image = render(clamp_t(t), alpha, rgb, white_bg=True)
        torch.cuda.synchronize()
        # the sampling is the oracle's
        oi, od, oh = oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width)
        rays = None if not subset else (uv[:, 1] * cam.width + uv[:, 0]).numpy().astype(np.int32)
        o_xyz, o_idx, o_t, o_mask = oracle.quick_sampling(pts, oracle.camera_from(cam), oracle.options_from(opt),
                                                          oi, oh, mode=oracle.MODE_COMPAT, rays=rays)
        assert np.array_equal(mask.cpu().numpy().reshape(o_mask.shape), o_mask)
        assert np.array_equal(pc_idx.cpu().numpy().reshape(o_idx.shape), o_idx)
        assert np.array_equal(bits(sp_t.cpu().numpy().reshape(o_t.shape)), bits(o_t))
        assert int(mask.sum()) > 500
        # the same chain with the fp32 aggregator and the oracle's render
        ref_a, ref_c = aggregator_reference(agg.cpu(), *[x.cpu() for x in fq])
        agg.to(dev)
        live = mask.reshape(-1).cpu()
        fa = torch.zeros(live.numel()).masked_scatter_(live, ref_a).view(rows, -1)
        fc = torch.zeros((live.numel(), 3)).masked_scatter_(live[:, None].expand(-1, 3), ref_c).view(rows, -1, 3)
        want = oracle.render(fa.numpy(), fc.numpy(), True, t=t.cpu().numpy())
    got = image.cpu().numpy()
    assert got.shape == (rows, 3) and np.isfinite(got).all()
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-2)
    # rays without samples keep the white background
    empty = ~mask.reshape(rows, -1).any(-1).cpu().numpy()
    assert (got[empty] == 1.0).all()
