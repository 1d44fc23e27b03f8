"""HIP quick_sampling / knn_sampling / quick_query / idw against the oracle:
indices, counts and ordering bit-exact; fp32 depths, points and distances
bit-exact (the tolerance the path allows is 1e-6 relative; the shared
operation order makes them identical)."""
import numpy as np
import pytest
import torch

import epcq.csrc as C
from epcq import synth
from epcq.epcq import NeuralPoints, RenderOptions
from oracle import oracle
from scenes import bits, camera_spec, to_dev

pytestmark = pytest.mark.gpu

FP_RTOL = 1e-6  # fp32 depths / distances tolerance of the path


class View:
    """A point cloud + camera on the device and in the oracle's form, with
    the oracle's hash table."""

    def __init__(self, pts, cam, dev, radius=None, **opt_kw):
        self.pts, self.cam, self.dev = pts, cam, dev
        self.opt = RenderOptions(**opt_kw)
        self.ocam = oracle.camera_from(cam)
        self.oopt = oracle.options_from(self.opt)
        self.oi, self.od, self.oh = oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width)
        self.npts = NeuralPoints(pc_type="pc", points=to_dev(pts, dev))
        self.cs = camera_spec(cam, dev)
        self.idx_d = to_dev(self.oi, dev)
        self.hash_d = to_dev(self.oh, dev)
        self.radius = radius
        self.radius_d = to_dev(radius, dev) if radius is not None else None


def assert_fp_equal(g, o):
    g, o = np.asarray(g), np.asarray(o)
    if not np.array_equal(bits(g), bits(o)):
        np.testing.assert_allclose(g, o, rtol=FP_RTOL, atol=0)
        raise AssertionError("fp32 outputs within tolerance but not bit-identical")


def check_dense(v, mode="compat", cull=True):
    H, W = v.cam.height, v.cam.width
    sp_xyz, pc_idx, sp_t, mask = v.npts.render_image_hashtable_topk(v.cs, v.opt, v.idx_d, None, v.hash_d, mode=mode,
                                                                    cull=cull)
    torch.cuda.synchronize()
    o_xyz, o_idx, o_t, o_mask = oracle.quick_sampling(v.pts, v.ocam, v.oopt, v.oi, v.oh,
                                                      mode=oracle.MODE_KNN if mode == "knn" else oracle.MODE_COMPAT)
    S, Kc = v.opt.num_sample_points_per_ray, v.opt.num_point_cloud_per_sp
    assert sp_xyz.shape == (H, W, S, 3) and pc_idx.shape == (H, W, S, Kc)
    assert sp_t.shape == (H, W, S) and mask.shape == (H, W, S) and mask.dtype == torch.bool
    np.testing.assert_array_equal(mask.cpu().numpy().reshape(o_mask.shape), o_mask)
    np.testing.assert_array_equal(pc_idx.cpu().numpy().reshape(o_idx.shape), o_idx)
    assert_fp_equal(sp_t.cpu().numpy().reshape(o_t.shape), o_t)
    assert_fp_equal(sp_xyz.cpu().numpy().reshape(o_xyz.shape), o_xyz)
    return o_mask


def test_quick_sampling_sphere_config(dev):
    """100k-point sphere, 400x400, r = 0.01, K = 8 (the CPU-runnable workload)."""
    pts, cams, kw = synth.workload_sphere_small()
    print("quick_sampling sphere 400x400", flush=True)
    live = check_dense(View(pts, cams[0], dev, **kw))
    assert live.sum() > 10_000


def test_quick_sampling_box_config(dev):
    """1M-point box scene, 800x800, r = 0.005, K = 8."""
    pts, cams, kw = synth.workload_box_scene()
    print("quick_sampling box 800x800", flush=True)
    live = check_dense(View(pts, cams[0], dev, **kw))
    assert live.sum() > 50_000


@pytest.mark.parametrize("kw", [
    dict(kernel_size=3, radius_threshold=0.02, num_point_cloud_per_sp=4),
    dict(kernel_size=4, radius_threshold=0.03, num_point_cloud_per_sp=8),
    dict(kernel_size=7, radius_threshold=0.05, num_point_cloud_per_sp=12, num_sample_points_per_ray=4),
    dict(kernel_size=5, radius_threshold=0.03, num_sample_points_per_ray=1, num_point_cloud_per_sp=8),
    dict(kernel_size=5, radius_threshold=0.03, sdf2weight_gaussian_gamma=0.2, sdf2weight_gaussian_alpha=0.01,
         num_point_cloud_per_sp=8),
    dict(kernel_size=5, radius_threshold=0.04, t_intvl=0.001, num_point_cloud_per_sp=8),
    dict(kernel_size=1, radius_threshold=0.05, num_point_cloud_per_sp=3),
])
@pytest.mark.parametrize("mode", ["compat", "knn"])
def test_quick_sampling_options(dev, kw, mode):
    pts = synth.sphere_cloud(40_000, seed=5)
    cam = synth.look_at_camera([0.3, -3.5, 0.4], 90, 70, 110.0)  # the sphere touches the borders
    print(f"quick_sampling {mode} {kw}", flush=True)
    v = View(pts, cam, dev, **kw)
    check_dense(v, mode)
    check_dense(v, mode, cull=False)


@pytest.mark.parametrize("mode", ["compat", "knn"])
def test_dense_small_view_culling(dev, mode):
    """A 200x150 view of the 1M-point scene (~30 points per pixel): the
    bucket culling skips most neighbour pixels; results stay exact."""
    pts = synth.box_scene(1_000_000, seed=0)
    cam = synth.look_at_camera([-2.1, 2.9, 1.7], 200, 150, synth.nerf_focal(200))
    print(f"dense small view {mode}", flush=True)
    v = View(pts, cam, dev, radius_threshold=0.005, num_point_cloud_per_sp=8)
    live = check_dense(v, mode)
    assert live.sum() > 5_000
    compact_vs_dense(v, mode)


def test_quick_sampling_box_knn(dev):
    pts, cam = synth.box_scene(300_000, seed=6), synth.look_at_camera([2.5, -2.0, 2.2], 320, 240, 444.4)
    print("quick_sampling knn box 320x240", flush=True)
    check_dense(View(pts, cam, dev, radius_threshold=0.01, num_point_cloud_per_sp=8), "knn")


def compact_vs_dense(v, mode, rays=None, feat=None, radius=None, cull=True):
    H, W = v.cam.height, v.cam.width
    out = v.npts.knn_sampling(v.cs, v.opt, v.idx_d, v.hash_d, rays=to_dev(rays, v.dev) if rays is not None else None,
                              radius=v.radius_d, mode=mode, feat=feat, cull=cull)
    torch.cuda.synchronize()
    res = oracle.quick_sampling(v.pts, v.ocam, v.oopt, v.oi, v.oh, radius=v.radius,
                                mode=oracle.MODE_KNN if mode == "knn" else oracle.MODE_COMPAT, rays=rays,
                                with_nb=True)
    o_xyz, o_idx, o_t, o_mask, o_nbt, o_nbd = res
    s_max = out["sample_t"].shape[1]
    n = out["n_samples"].cpu().numpy()
    np.testing.assert_array_equal(n, o_mask.sum(1))
    assert n.max() <= s_max
    assert (o_mask[:, s_max:] == 0).all()
    assert_fp_equal(out["sample_t"].cpu().numpy(), o_t[:, :s_max])
    np.testing.assert_array_equal(out["nb_idx"].cpu().numpy(), o_idx[:, :s_max])
    if mode == "knn":
        assert_fp_equal(out["nb_t"].cpu().numpy(), o_nbt[:, :s_max])
        assert_fp_equal(out["nb_dist"].cpu().numpy(), o_nbd[:, :s_max])
    if feat is not None:
        fh = feat.cpu()
        bf16 = fh.dtype == torch.bfloat16
        fnp = fh.view(torch.int16).numpy().view(np.uint16) if bf16 else fh.numpy()
        w, agg = oracle.idw(v.pts, fnp, o_xyz[:, :s_max], o_idx[:, :s_max], bf16=bf16)
        live = o_mask[:, :s_max].reshape(-1)
        gw = out["w"].cpu().numpy().reshape(-1, w.shape[1])
        ga = out["agg"].cpu().numpy().reshape(-1, agg.shape[1])
        assert_fp_equal(gw[live], w[live])
        assert_fp_equal(ga[live], agg[live])
        assert (gw[~live] == 0).all() and (ga[~live] == 0).all()
    return out


@pytest.mark.parametrize("mode", ["compat", "knn"])
def test_knn_sampling_compact_equals_dense(dev, mode):
    pts, cams, kw = synth.workload_sphere_small()
    v = View(pts, cams[0], dev, **kw)
    print(f"knn_sampling compact {mode}", flush=True)
    out = compact_vs_dense(v, mode)
    assert out["sample_t"].shape[1] == 2  # the emission bound for these options


def test_knn_sampling_ray_subset_and_features(dev):
    pts, cam = synth.box_scene(400_000, seed=7), synth.look_at_camera([-2.0, -2.5, 2.0], 400, 300, 555.5)
    v = View(pts, cam, dev, radius_threshold=0.008, num_point_cloud_per_sp=8)
    rng = np.random.default_rng(1)
    rays = rng.integers(-5, 400 * 300 + 5, size=20_000).astype(np.int32)  # includes out-of-range ids
    feat = torch.randn(len(pts), 16, generator=torch.Generator().manual_seed(1)).to(dev)
    print("knn_sampling subset + fp32 features", flush=True)
    compact_vs_dense(v, "knn", rays=rays, feat=feat)
    print("knn_sampling subset + bf16 features (compat)", flush=True)
    compact_vs_dense(v, "compat", rays=rays, feat=feat[:, :13].contiguous().to(torch.bfloat16))


def test_knn_sampling_per_point_radius_k16(dev):
    pts = synth.box_scene(300_000, n_boxes=40, seed=8)
    radius = synth.log_uniform_radius(len(pts), seed=9)
    cam = synth.look_at_camera([2.0, -2.0, 3.0], 320, 180, 444.0)
    v = View(pts, cam, dev, radius=radius, kernel_size=9, radius_threshold=0.01, num_point_cloud_per_sp=16)
    feat = torch.randn(len(pts), 32, generator=torch.Generator().manual_seed(2)).to(torch.bfloat16).to(dev)
    print("knn_sampling per-point radius K=16", flush=True)
    compact_vs_dense(v, "knn", feat=feat)
    compact_vs_dense(v, "knn", cull=False)


def test_quick_query_vs_oracle(dev):
    pts, cam = synth.sphere_cloud(100_000, seed=11), synth.look_at_camera([0.0, -4.0, 0.0], 400, 400, 555.5)
    for kw in (dict(), dict(kernel_size=7, num_point_per_ray=64), dict(kernel_size=3, radius_threshold=0.05)):
        v = View(pts, cam, dev, **kw)
        rays = np.random.default_rng(3).integers(-10, 400 * 400 + 10, size=9000).astype(np.int32)
        print(f"quick_query {kw}", flush=True)
        g = C.quick_query_for_nearby_point(v.npts._to_cpp(), v.cs._to_cpp(), v.opt._to_cpp(), v.idx_d,
                                           v.hash_d, to_dev(rays, dev)).cpu().numpy()
        o = oracle.quick_query(pts, v.ocam, v.oopt, v.oi, v.oh, rays)
        np.testing.assert_array_equal(g, o)
        assert (g >= 0).sum() > 0


def test_idw_features_dense(dev):
    pts, cams, kw = synth.workload_sphere_small()
    v = View(pts, cams[0], dev, **kw)
    sp_xyz, pc_idx, _, mask = v.npts.render_image_hashtable_topk(v.cs, v.opt, v.idx_d, None, v.hash_d)
    m = mask.reshape(-1)
    sp = sp_xyz.reshape(-1, 3)[m].contiguous()
    idx = pc_idx.reshape(-1, pc_idx.shape[-1])[m].contiguous()
    feat = torch.randn(len(pts), 16, generator=torch.Generator().manual_seed(4)).to(dev)
    w, agg = C.idw_features(v.npts.xyz_data, feat, sp, idx)
    ow, oagg = oracle.idw(pts, feat.cpu().numpy(), sp.cpu().numpy(), idx.cpu().numpy())
    assert_fp_equal(w.cpu().numpy(), ow)
    assert_fp_equal(agg.cpu().numpy(), oagg)


def test_device_expf_matches_oracle(dev):
    x = np.concatenate([np.linspace(-104, 89, 200_001, dtype=np.float32),
                        np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-30, -1e-30], np.float32)])
    g = C.debug_expf(to_dev(x, dev)).cpu().numpy()
    o = np.array([oracle.expf(v) for v in x], np.float32)
    np.testing.assert_array_equal(bits(g), bits(o))


def test_sampling_deterministic(dev):
    pts, cams, kw = synth.workload_box_scene(300_000)
    v = View(pts, cams[0], dev, **kw)
    a = v.npts.render_image_hashtable_topk(v.cs, v.opt, v.idx_d, None, v.hash_d)
    b = v.npts.render_image_hashtable_topk(v.cs, v.opt, v.idx_d, None, v.hash_d)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
