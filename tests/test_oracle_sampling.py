"""Pins the oracle's sampling functions (quick_sampling in compat and knn
mode, quick_query, idw) with hand-built scenes whose answers follow from the
reference's rules by hand (epcq/csrc/epcq_render_topk.cu:10-339,
epcq/epcq.py:434-452), and with set-based brute force on small random
scenes.

The hand-built scenes use a camera at the origin looking down +z with
identity rotation, focal f and principal point (W/2, H/2): the ray of the
centre pixel is o = 0, d = (0, 0, 1) exactly, so a point (a, b, z) has ray
depth t = z and point-to-ray distance sqrt(a^2 + b^2) with no rounding in
play.
"""
import numpy as np
import pytest

from oracle import oracle

W = H = 9
CENTRE = (H // 2) * W + W // 2


def axis_camera(f=2.0, width=W, height=H):
    c2w = np.eye(4, dtype=np.float32)[:3]
    K = np.array([[f, 0, width / 2], [0, f, height / 2], [0, 0, 1]], np.float32)
    cam = oracle.camera(c2w, c2w, K, f, f, width / 2, height / 2, width, height)
    return cam, c2w, K


def sample(pts, mode=oracle.MODE_COMPAT, rays=(CENTRE,), f=2.0, **opt):
    pts = np.asarray(pts, np.float32).reshape(-1, 3)
    cam, w2c, K = axis_camera(f)
    idx, _, table = oracle.pts2plane(pts, w2c, K, H, W)
    kw = dict(radius_threshold=0.01, num_point_cloud_per_sp=8, kernel_size=5)
    kw.update(opt)
    o = oracle.options(**kw)
    sp_xyz, pc_idx, sp_t, mask, nb_t, nb_d = oracle.quick_sampling(
        pts, cam, o, idx, table, mode=mode, rays=np.asarray(rays, np.int32), with_nb=True)
    return dict(sp_xyz=sp_xyz[0], pc_idx=pc_idx[0], sp_t=sp_t[0], mask=mask[0], nb_t=nb_t[0], nb_dist=nb_d[0])


def test_axis_ray_is_exact():
    cam, _, _ = axis_camera()
    o, d = oracle.ray(cam, W // 2, H // 2)
    assert o.tolist() == [0.0, 0.0, 0.0] and d.tolist() == [0.0, 0.0, 1.0]


def test_two_hits_in_one_slot_give_one_sample():
    r = sample([(0, 0, 2.0), (0.001, 0, 2.004)])
    assert r["mask"].tolist() == [True] + [False] * 15
    assert r["sp_t"][0] == np.float32(2.0) and np.all(r["sp_t"][1:] == 0)
    assert r["sp_xyz"][0].tolist() == [0.0, 0.0, 2.0]
    assert r["pc_idx"][0].tolist() == [0, 1, -1, -1, -1, -1, -1, -1]
    assert np.all(r["pc_idx"][1:] == -1)


def test_single_hits_per_slot_give_no_sample():
    # slots 0 and 1 each hold one hit: a slot needs at least two
    r = sample([(0, 0, 2.0), (0, 0, 2.02)])
    assert not r["mask"].any() and np.all(r["sp_t"] == 0) and np.all(r["pc_idx"] == -1)


def test_two_slots_two_samples():
    # hits at t = 2 (x2) and t = 2.0625 (x2): slots 0 and int(0.0625 / 0.01) = 6
    r = sample([(0, 0, 2.0), (0.002, 0, 2.0), (0.001, 0, 2.0625), (0, 0.003, 2.0625)])
    assert r["mask"].tolist()[:3] == [True, True, False]
    assert r["sp_t"][0] == np.float32(2.0)
    assert 2.0599 < r["sp_t"][1] < 2.0601  # tmin + 6 * t_intvl
    assert r["pc_idx"][0][:3].tolist() == [0, 1, -1]
    assert r["pc_idx"][1][:3].tolist() == [2, 3, -1]


def test_transmittance_stops_the_walk():
    # sdf = 0 so alpha = gamma exactly; gamma 0.3: weights 0.3, 0.21, 0.12, 0.03
    # then transmittance -0.2 < 1e-3 ends the walk: four of five slots emit
    zs = [2.0, 2.125, 2.25, 2.375, 2.5]
    pts = [(0, 0, z) for z in zs for _ in range(2)]
    r = sample(pts, sdf2weight_gaussian_gamma=0.3)
    assert r["mask"].sum() == 4
    # default gamma 0.9: 0.9, then 0.09; transmittance -0.8 ends it
    assert sample(pts)["mask"].sum() == 2
    # the output capacity S also ends it
    assert sample(pts, sdf2weight_gaussian_gamma=0.3, num_sample_points_per_ray=3)["mask"].sum() == 3


def test_transmittance_break_precedes_the_weight_test():
    # gamma 3, alpha_p 0.01: slot 0 (sdf 0.010991) has alpha = 3 exp(-1.0991)
    # ~ 0.9995, leaving transmittance ~ 5e-4 < 1e-3, which ends the walk; the
    # next slot (alpha 3, weight ~ 1.5e-3 > 1e-3) would otherwise emit
    a = 0.010991
    pts = [(a, 0, 2.0), (-a, 0, 2.0), (0, 0, 2.0625), (0, 0, 2.0625)]
    r = sample(pts, radius_threshold=0.05, sdf2weight_gaussian_gamma=3.0, sdf2weight_gaussian_alpha=0.01)
    assert r["mask"].sum() == 1 and r["sp_t"][0] == np.float32(2.0)


def test_slot_cap_of_128():
    # t - tmin = 1.5 gives slot 150 (dropped); 1.25 gives slot 125 (kept)
    far = sample([(0, 0, 2.0), (0, 0, 2.0), (0, 0, 3.5), (0, 0, 3.5)])
    assert far["mask"].sum() == 1
    near = sample([(0, 0, 2.0), (0, 0, 2.0), (0, 0, 3.25), (0, 0, 3.25)])
    assert near["mask"].sum() == 2 and near["sp_t"][1] == np.float32(3.25)


def test_weight_threshold():
    # alpha_p = 0.001: sdf 0.009 gives alpha = 0.9 exp(-9) ~ 1.1e-4 < 1e-3 (not
    # emitted); sdf 0.001 gives 0.9 exp(-1) ~ 0.33 (emitted)
    weak = sample([(0.009, 0, 2.0), (0, 0.009, 2.0)], sdf2weight_gaussian_alpha=0.001)
    assert not weak["mask"].any()
    strong = sample([(0.001, 0, 2.0), (0, 0.001, 2.0)], sdf2weight_gaussian_alpha=0.001)
    assert strong["mask"].sum() == 1


def test_radius_is_strict():
    # distance exactly 0.5 = r is not a hit
    pts = [(0.5, 0, 2.0), (0, 0.5, 2.0)]
    assert not sample(pts, radius_threshold=0.5)["mask"].any()
    assert sample(pts, radius_threshold=0.50001)["mask"].sum() == 1


def test_48_hits_per_bucket():
    # one pixel holds 48 hits at t = 2 followed (in point order) by two at t = 1:
    # the bucket stops after 48 hits, so tmin stays 2
    late = [(0, 0, 2.0)] * 48 + [(0, 0, 1.0)] * 2
    r = sample(late)
    assert r["mask"].sum() == 1 and r["sp_t"][0] == np.float32(2.0)
    # the same points with the t = 1 pair first: tmin = 1, slots 0 and 100
    early = [(0, 0, 1.0)] * 2 + [(0, 0, 2.0)] * 48
    r = sample(early)
    assert r["mask"].sum() == 2
    assert r["sp_t"][0] == np.float32(1.0) and r["sp_t"][1] == np.float32(2.0)


# lateral offsets of ten points at t = 2, in point (= visit) order
QUIRK = [5e-3, 4e-3, 3e-3, 2e-3, 1e-3, 6e-3, 7e-3, 8e-3, 0.5e-3, 0.9e-3]


def test_compat_buffer_fill_phase_quirk():
    # quick_search_nearby_pc (:44-72) tracks the minimum while filling, so the
    # 9th point (0.5e-3) evicts the closest one (point 4, 1e-3), and the 10th
    # (0.9e-3) then evicts the true maximum (point 7, 8e-3)
    r = sample([(d, 0, 2.0) for d in QUIRK])
    assert r["mask"].sum() == 1 and r["sp_t"][0] == np.float32(2.0)
    assert r["pc_idx"][0].tolist() == [0, 1, 2, 3, 8, 5, 6, 9]
    # fewer outputs than the 8-entry buffer: the buffer's first entries
    r4 = sample([(d, 0, 2.0) for d in QUIRK], num_point_cloud_per_sp=4)
    assert r4["pc_idx"][0].tolist() == [0, 1, 2, 3]


def test_knn_mode_true_nearest_with_depth_and_distance():
    r = sample([(d, 0, 2.0) for d in QUIRK], mode=oracle.MODE_KNN)
    assert r["pc_idx"][0].tolist() == [8, 9, 4, 3, 2, 1, 0, 5]
    want = np.array(sorted(QUIRK)[:8], np.float32)
    assert np.array_equal(r["nb_dist"][0], want)
    assert np.all(r["nb_t"][0] == np.float32(2.0))
    # ties by point index
    r = sample([(1e-3, 0, 2.0), (0, 1e-3, 2.0), (-1e-3, 0, 2.0)], mode=oracle.MODE_KNN)
    assert r["pc_idx"][0][:4].tolist() == [0, 1, 2, -1]


def test_knn_mode_counts_border_pixels_once():
    # at pixel (0, 0) with k = 3 the clamped window visits pixel (0, 0) four
    # times; compat sees each of its points four times, knn once
    cam, w2c, K = axis_camera()
    d = np.array([(0.5 - W / 2) / 2.0, (0.5 - H / 2) / 2.0, 1.0])
    d /= np.linalg.norm(d)
    pts = np.array([3.0 * d, 3.0 * d + [0.001, 0, 0]], np.float32)
    compat = sample(pts, rays=(0,), kernel_size=3)
    knn = sample(pts, rays=(0,), kernel_size=3, mode=oracle.MODE_KNN)
    assert compat["mask"].sum() == 1 and knn["mask"].sum() == 1
    assert sorted(compat["pc_idx"][0].tolist()) == [0, 0, 0, 0, 1, 1, 1, 1]
    assert sorted(knn["pc_idx"][0][:2].tolist()) == [0, 1] and np.all(knn["pc_idx"][0][2:] == -1)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_knn_mode_equals_brute_force_set(seed):
    """Every live sample's knn neighbours are the K nearest (distance, index)
    among all points that project into the sample's clamped pixel window,
    found by scanning all points (no hash table)."""
    rng = np.random.default_rng(seed)
    Wd, Hd, f = 24, 18, 30.0
    n = 3000
    pts = np.empty((n, 3), np.float32)
    pts[:, :2] = rng.uniform(-0.3, 0.3, (n, 2))
    pts[:, 2] = rng.choice([2.0, 2.03], n) + rng.normal(0, 0.002, n)
    c2w = np.eye(4, dtype=np.float32)[:3]
    K = np.array([[f, 0, Wd / 2], [0, f, Hd / 2], [0, 0, 1]], np.float32)
    cam = oracle.camera(c2w, c2w, K, f, f, Wd / 2, Hd / 2, Wd, Hd)
    idx, _, table = oracle.pts2plane(pts, c2w, K, Hd, Wd)
    key, _ = oracle.project(pts, c2w, K, Hd, Wd)
    kc, r0, k = 8, 0.02, 5
    o = oracle.options(radius_threshold=r0, num_point_cloud_per_sp=kc, kernel_size=k)
    sp_xyz, pc_idx, _, mask, _, nb_d = oracle.quick_sampling(pts, cam, o, idx, table, mode=oracle.MODE_KNN,
                                                             with_nb=True)
    pu, pv = key % Wd, key // Wd
    inside = key < Wd * Hd
    live = checked = 0
    for q in range(Wd * Hd):
        u, v = q % Wd, q // Wd
        win = inside & (np.abs(pu - u) <= 2) & (np.abs(pv - v) <= 2)
        for s in np.nonzero(mask[q])[0]:
            live += 1
            d = np.linalg.norm(pts[win].astype(np.float64) - sp_xyz[q, s].astype(np.float64), axis=1)
            ids = np.nonzero(win)[0]
            ok = d < r0
            order = np.lexsort((ids[ok], d[ok]))
            want = ids[ok][order][:kc]
            got = pc_idx[q, s][pc_idx[q, s] >= 0]
            dd = d[ok][order][:kc + 1]
            if np.any(np.diff(dd) < 1e-6):
                continue  # a float-level tie, which float64 ranking cannot decide
            assert got.tolist() == want.tolist(), (q, s)
            checked += 1
    assert live > 50 and checked > 0.8 * live


def test_quick_query_lane_order_caps_and_border_duplicates():
    cam, w2c, K = axis_camera()
    o_pts = [(0, 0, 2.0)] * 40  # 40 hits in the centre pixel
    pts = np.array(o_pts + [(0.0005, 0, 2.0)], np.float32)
    idx, _, table = oracle.pts2plane(pts, w2c, K, H, W)
    o = oracle.options(radius_threshold=0.01, kernel_size=5, num_point_per_ray=256)
    out = oracle.quick_query(pts, cam, o, idx, table, np.array([CENTRE, -1, W * H], np.int32))
    # 32 hits per bucket, in bucket order; invalid ray ids give empty rows
    assert out[0][:32].tolist() == list(range(32)) and np.all(out[0][32:] == -1)
    assert np.all(out[1] == -1) and np.all(out[2] == -1)
    o = oracle.options(radius_threshold=0.01, kernel_size=5, num_point_per_ray=10)
    out = oracle.quick_query(pts, cam, o, idx, table, np.array([CENTRE], np.int32))
    assert out[0].tolist() == list(range(10))
    # pixel (0, 0), k = 3: four clamped lanes visit it, each appends its hits
    d = np.array([(0.5 - W / 2) / 2.0, (0.5 - H / 2) / 2.0, 1.0])
    d /= np.linalg.norm(d)
    pts = np.array([3.0 * d, 3.2 * d], np.float32)
    idx, _, table = oracle.pts2plane(pts, w2c, K, H, W)
    o = oracle.options(radius_threshold=0.01, kernel_size=3, num_point_per_ray=16)
    out = oracle.quick_query(pts, cam, o, idx, table, np.array([0], np.int32))
    assert out[0][:8].tolist() == [0, 1] * 4 and np.all(out[0][8:] == -1)


def test_quick_query_equals_brute_force_hits():
    rng = np.random.default_rng(5)
    Wd, Hd, f = 20, 16, 25.0
    n = 2000
    pts = np.empty((n, 3), np.float32)
    pts[:, :2] = rng.uniform(-0.3, 0.3, (n, 2))
    pts[:, 2] = rng.uniform(1.8, 2.2, n)
    c2w = np.eye(4, dtype=np.float32)[:3]
    K = np.array([[f, 0, Wd / 2], [0, f, Hd / 2], [0, 0, 1]], np.float32)
    cam = oracle.camera(c2w, c2w, K, f, f, Wd / 2, Hd / 2, Wd, Hd)
    idx, _, table = oracle.pts2plane(pts, c2w, K, Hd, Wd)
    key, _ = oracle.project(pts, c2w, K, Hd, Wd)
    r0 = 0.01
    o = oracle.options(radius_threshold=r0, kernel_size=3, num_point_per_ray=256)
    rays = np.arange(2 * Wd + 2, 2 * Wd + 12, dtype=np.int32)  # interior pixels: no clamping
    out = oracle.quick_query(pts, cam, o, idx, table, rays)
    for row, q in zip(out, rays):
        u, v = q % Wd, q // Wd
        x = ((u + 0.5) - Wd / 2) / f
        y = ((v + 0.5) - Hd / 2) / f
        dvec = np.array([x, y, 1.0]) / np.linalg.norm([x, y, 1.0])
        p = pts.astype(np.float64)
        dist = np.linalg.norm(p - (p @ dvec)[:, None] * dvec, axis=1)
        win = (key < Wd * Hd) & (np.abs(key % Wd - u) <= 1) & (np.abs(key // Wd - v) <= 1)
        near = np.abs(dist - r0) < 1e-6
        want = set(np.nonzero(win & (dist < r0) & ~near)[0].tolist())
        got = set(row[row >= 0].tolist())
        maybe = set(np.nonzero(win & near)[0].tolist())
        assert want <= got <= want | maybe, q


def test_idw_closed_form():
    # sample at the origin, neighbours at distances 1, 2, 4 and one empty entry
    xyz = np.array([(1, 0, 0), (0, 2, 0), (0, 0, 4)], np.float32)
    feat = np.array([(1.0, 0.0), (0.0, 2.0), (-0.5, 1.0)], np.float32)
    sp = np.zeros((1, 3), np.float32)
    nb = np.array([[0, 1, 2, -1]], np.int32)
    w, agg = oracle.idw(xyz, feat, sp, nb)
    inv = np.array([1 / (1 + 1e-6), 1 / (2 + 1e-6), 1 / (4 + 1e-6), 0.0])
    want_w = inv / (inv.sum() + 1e-6)
    assert np.allclose(w[0], want_w, rtol=1e-6, atol=0)
    assert abs(want_w[0] - 4 / 7) < 1e-5
    want_agg = want_w[:3] @ feat.astype(np.float64)
    assert np.allclose(agg[0], want_agg, rtol=1e-6, atol=1e-7)
    # the same through bf16 features (values exact in bf16)
    bits = (feat.view(np.uint32) >> 16).astype(np.uint16)
    w2, agg2 = oracle.idw(xyz, bits, sp, nb, bf16=True)
    assert np.array_equal(w2, w) and np.array_equal(agg2, agg)
    # no neighbour at all: zero weights and features
    w3, agg3 = oracle.idw(xyz, feat, sp, np.full((1, 4), -1, np.int32))
    assert np.all(w3 == 0) and np.all(agg3 == 0)
