"""HIP hash-table build, scatter_hashtable and scatter against the oracle:
indices, counts and order bit-exact, depths bit-exact."""
import numpy as np
import pytest
import torch

from epcq import synth
from oracle import oracle
from scenes import bits, to_dev

pytestmark = pytest.mark.gpu


def gpu_build(pts, cam, dev, sort=False):
    import epcq.csrc as C
    out = C.build_hashtable(to_dev(pts, dev), to_dev(cam.w2c, dev), to_dev(cam.K, dev),
                            [cam.height, cam.width], sort)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in out]


def check_build(pts, cam, dev, sort=False):
    gi, gd, gh = gpu_build(pts, cam, dev, sort)
    oi, od, oh = oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width, sort=sort)
    np.testing.assert_array_equal(gh, oh)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(bits(gd), bits(od))
    return gi, gd, gh


@pytest.mark.parametrize("sort", [False, True])
def test_build_sphere_400(dev, sort):
    pts, cams, _ = synth.workload_sphere_small()
    print(f"build sphere 100k 400x400 sort={sort}", flush=True)
    check_build(pts, cams[0], dev, sort)


def test_build_box_scene_800(dev):
    pts, cams, _ = synth.workload_box_scene()
    print("build box 1M 800x800", flush=True)
    check_build(pts, cams[0], dev)


@pytest.mark.parametrize("W,H", [(1920, 1080), (1600, 1200), (200, 150), (1, 1), (3, 1000), (1000, 3)])
def test_build_resolutions(dev, W, H):
    pts = synth.box_scene(300_000, n_boxes=20, seed=4)
    cam = synth.look_at_camera([1.0, -3.0, 2.5], W, H, synth.nerf_focal(max(W, 8)))
    print(f"build box 300k {W}x{H}", flush=True)
    check_build(pts, cam, dev, sort=(W == 200))


def test_build_edge_cases(dev):
    cam = synth.look_at_camera([0.0, -4.0, 0.0], 64, 48, 80.0)
    print("build edge: empty", flush=True)
    gi, gd, gh = gpu_build(np.zeros((0, 3), np.float32), cam, dev)
    assert gi.size == 0 and (gh == -1).all()
    print("build edge: one point", flush=True)
    check_build(np.array([[0.0, 0.0, 0.0]], np.float32), cam, dev)
    print("build edge: all off-image", flush=True)
    check_build(np.array([[50.0, 0.0, 0.0], [0.0, 0.0, 50.0]] * 100, np.float32), cam, dev)
    print("build edge: hot pixel 60k points", flush=True)
    hot = np.zeros((60_000, 3), np.float32)
    hot[:, 1] = np.linspace(-0.5, 0.5, 60_000, dtype=np.float32)  # all on the optical axis
    gi, gd, gh = check_build(hot, cam, dev, sort=True)
    busy = gh[:, 1] > 0
    assert gh[busy, 1].sum() == 60_000 and busy.sum() <= 4 and gh[:, 1].max() > 20_000
    print("build edge: hot pixel + background", flush=True)
    mix = np.concatenate([hot[:20_000], synth.sphere_cloud(50_000, seed=3)])
    check_build(mix, cam, dev, sort=False)


def long_runs(n, levels, offsets=(0.0,)):
    """n points on the optical axis of a camera at (0, -4, 0), depths from 7
    down to -3 in point order (the insertion sort's worst case; negative depths
    are points behind the camera, which project onto the centre pixel too),
    rounded to 1/levels so that many depths repeat; point i is shifted
    sideways by offsets[i % len(offsets)] to spread the runs over pixels."""
    y = np.round(np.linspace(3.0, -7.0, n) * levels) / levels
    pts = np.zeros((n, 3), np.float32)
    pts[:, 1] = y
    pts[:, 0] = np.asarray(offsets, np.float32)[np.arange(n) % len(offsets)]
    return pts


def test_depth_sort_long_runs_vs_oracle(dev):
    """Runs far longer than one thread sorts (32 points) go through the block-
    wide radix sort: the fused build and scatter_hashtable against the oracle."""
    import epcq.csrc as C
    cam = synth.look_at_camera([0.0, -4.0, 0.0], 64, 48, 80.0)
    pts = long_runs(24_000, 100, offsets=(0.0, 0.0, 0.03, -0.05))
    print("depth sort: fused build, runs of ~12k", flush=True)
    gi, gd, gh = check_build(pts, cam, dev, sort=True)
    assert gh[:, 1].max() > 5000 and (gd[gd != -1] < 0).any()
    print("depth sort: scatter_hashtable", flush=True)
    ix, ks, ds = presorted_inputs(pts, cam)
    gi, gd, gh = [t.cpu().numpy() for t in C.scatter_hashtable(to_dev(ix, dev), to_dev(ks, dev), to_dev(ds, dev),
                                                             (cam.height, cam.width), True)]
    oi, od, oh = oracle.scatter_hashtable(ix, ks, ds, cam.height, cam.width, True)
    np.testing.assert_array_equal(gh, oh)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(bits(gd), bits(od))


def test_depth_sort_hot_pixel_properties(dev):
    """300k points in (almost) one pixel, depths descending in point order
    with ties: each run equals the stable argsort of the oracle's projected
    depths (the oracle's bubble sort would take minutes at this size)."""
    import epcq.csrc as C
    cam = synth.look_at_camera([0.0, -4.0, 0.0], 64, 48, 80.0)
    pts = long_runs(300_000, 1000)
    key, depth = oracle.project(pts, cam.w2c, cam.K, cam.height, cam.width)
    P = cam.height * cam.width
    inside = np.nonzero(key < P)[0]
    print("depth sort: 300k points on the optical axis", flush=True)
    gi, gd, gh = gpu_build(pts, cam, dev, sort=True)
    assert gh[:, 1].max() > 250_000
    for p in np.unique(key[inside]):
        sel = inside[key[inside] == p]
        want = sel[np.argsort(depth[sel], kind="stable")]
        s, c = gh[p]
        assert c == sel.size
        np.testing.assert_array_equal(gi[s:s + c], want)
        np.testing.assert_array_equal(bits(gd[s:s + c]), bits(depth[want]))
    assert (gi[inside.size:] == -1).all()
    print("depth sort: scatter_hashtable, the same runs", flush=True)
    ix, ks, ds = presorted_inputs(pts, cam)
    si, sd, sh = [t.cpu().numpy() for t in C.scatter_hashtable(to_dev(ix, dev), to_dev(ks, dev), to_dev(ds, dev),
                                                             (cam.height, cam.width), True)]
    np.testing.assert_array_equal(si, gi)
    np.testing.assert_array_equal(bits(sd), bits(gd))
    np.testing.assert_array_equal(sh, gh)


def test_build_deterministic(dev):
    pts, cam = synth.box_scene(500_000, seed=9), synth.look_at_camera([2.0, 2.0, 2.0], 640, 512, 888.8)
    a = gpu_build(pts, cam, dev)
    b = gpu_build(pts, cam, dev)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(bits(x), bits(y))


def test_build_via_neural_points_api(dev):
    from scenes import camera_spec
    from epcq.epcq import NeuralPoints, RenderOptions
    pts, cams, _ = synth.workload_sphere_small()
    npts = NeuralPoints(pc_type="pc", points=to_dev(pts, dev))
    idx, depth, table = npts.pts2plane(camera_spec(cams[0], dev), RenderOptions(), sort=False)
    oi, od, oh = oracle.pts2plane(pts, cams[0].w2c, cams[0].K, 400, 400)
    np.testing.assert_array_equal(idx.cpu().numpy(), oi)
    np.testing.assert_array_equal(table.cpu().numpy(), oh)
    assert table.shape == (400 * 400, 2) and table.dtype == torch.int32


def presorted_inputs(pts, cam):
    key, depth = oracle.project(pts, cam.w2c, cam.K, cam.height, cam.width)
    ks, ix, ds = oracle.sort_by_key(key, depth)
    return ix, ks, ds


@pytest.mark.parametrize("sort", [False, True])
def test_scatter_hashtable_vs_oracle(dev, sort):
    import epcq.csrc as C
    pts, cam = synth.sphere_cloud(100_000, seed=1), synth.look_at_camera([0.0, -4.0, 0.0], 300, 200, 400.0)
    ix, ks, ds = presorted_inputs(pts, cam)
    # a shuffled tie order inside pixels (as torch.sort may leave it)
    rng = np.random.default_rng(0)
    order = np.lexsort((rng.permutation(ks.size), ks))
    ix, ks, ds = ix[order], ks[order], ds[order]
    print(f"scatter_hashtable sort={sort}", flush=True)
    gi, gd, gh = [t.cpu().numpy() for t in C.scatter_hashtable(to_dev(ix, dev), to_dev(ks, dev), to_dev(ds, dev),
                                                             (cam.height, cam.width), sort)]
    oi, od, oh = oracle.scatter_hashtable(ix, ks, ds, cam.height, cam.width, sort)
    np.testing.assert_array_equal(gh, oh)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(bits(gd), bits(od))


@pytest.mark.parametrize("bins", [1, 4, 8])
def test_scatter_vs_oracle(dev, bins):
    import epcq.csrc as C
    pts, cam = synth.sphere_cloud(60_000, seed=2), synth.look_at_camera([0.0, -4.0, 0.0], 120, 100, 160.0)
    ix, ks, ds = presorted_inputs(pts, cam)
    ds = ds.copy()
    ds[::97] = -1.0  # skipped entries
    print(f"scatter bins={bins}", flush=True)
    gd, gi = [t.cpu().numpy() for t in C.scatter(to_dev(ix, dev), to_dev(ks, dev), to_dev(ds, dev),
                                               (cam.height, cam.width), bins)]
    od, oi = oracle.scatter(ix, ks, ds, cam.height, cam.width, bins)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(bits(gd), bits(od))
