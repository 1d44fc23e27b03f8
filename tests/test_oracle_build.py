"""Pins the oracle's hash-table build (A4-A8) and per-pixel top-B (A7) with
closed forms, invariants and brute force."""
import numpy as np
import pytest

from oracle import oracle


def cam_axes(W, H, f):
    """Identity-rotation camera at the origin looking down +z."""
    w2c = np.zeros((3, 4), np.float32)
    w2c[:, :3] = np.eye(3)
    K = np.array([[f, 0, W / 2], [0, f, H / 2], [0, 0, 1]], np.float32)
    return w2c, K


def test_projection_closed_form():
    W, H, f = 10, 8, 5.0
    w2c, K = cam_axes(W, H, f)
    # screen x = f*X/Z + W/2 ; place points at chosen screen coordinates, Z = 2
    Z = 2.0

    def at(sx, sy, z=Z):
        return [(sx - W / 2) * z / f, (sy - H / 2) * z / f, z]

    pts = np.array([
        at(0.5, 0.5),     # pixel (0, 0)
        at(9.5, 7.5),     # pixel (9, 7) -> key 79
        at(3.0, 4.0),     # exactly integer screen coords -> (3, 4) -> 43
        at(-0.5, 2.5),    # sx in (-1, 0): truncation toward zero keeps u = 0 -> key 20
        at(-1.0, 2.5),    # sx = -1 -> u = -1, off-image
        at(10.0, 1.5),    # sx = W -> off-image
        at(5.5, 8.0),     # sy = H -> off-image
        [0.1, 0.1, 0.0],  # on the camera plane: division by zero -> off-image
        at(4.5, 3.5, -2.0),  # behind the camera: projected through the centre, no culling
    ], np.float32)
    key, depth = oracle.project(pts, w2c, K, H, W)
    P = H * W
    assert key.tolist()[:8] == [0, 79, 43, 20, P, P, P, P]
    # behind the camera: X = (4.5-5)*-2/5 = 0.2, Y = (3.5-4)*-2/5 = 0.2, Z = -2
    # sx = (5*0.2 + 5*-2) / -2 = 4.5 -> u = 4 ; sy = (5*0.2 + 4*-2)/-2 = 3.5 -> v = 3
    assert key[8] == 3 * W + 4
    np.testing.assert_array_equal(depth, pts[:, 2])


def brute_force_table(key, depth, P, sort):
    """Counting sort written from the definition: pixel lists in point-index
    order (or (depth, index) order), row-major starts, (-1,-1) for empty
    pixels, -1 tail."""
    n = key.size
    lists = [[] for _ in range(P)]
    for i in range(n):
        if key[i] < P:
            lists[key[i]].append(i)
    idx = np.full(n, -1, np.int32)
    dep = np.full(n, -1.0, np.float32)
    table = np.full((P, 2), -1, np.int32)
    pos = 0
    for p in range(P):
        run = lists[p]
        if sort:
            run = sorted(run, key=lambda i: (depth[i], i))
        if run:
            table[p] = (pos, len(run))
        for i in run:
            idx[pos] = i
            dep[pos] = depth[i]
            pos += 1
    return idx, dep, table


def random_scene(n, seed, W, H, f, spread=1.5):
    rng = np.random.default_rng(seed)
    pts = rng.normal(size=(n, 3)).astype(np.float32) * spread
    pts[:, 2] = np.abs(pts[:, 2]) + 1.0
    w2c, K = cam_axes(W, H, f)
    return pts, w2c, K


@pytest.mark.parametrize("sort", [False, True])
@pytest.mark.parametrize("n,W,H,f,seed", [(500, 7, 5, 3.0, 0), (3000, 16, 9, 6.0, 1), (1, 4, 4, 2.0, 2),
                                          (2000, 1, 1, 1.0, 3), (800, 33, 2, 4.0, 4)])
def test_pts2plane_equals_brute_force(n, W, H, f, seed, sort):
    pts, w2c, K = random_scene(n, seed, W, H, f)
    if seed == 0:  # duplicate depths inside pixels exercise the (depth, index) tie order
        pts[1::3] = pts[0:-1:3][: pts[1::3].shape[0]]
    key, depth = oracle.project(pts, w2c, K, H, W)
    idx, dep, table = oracle.pts2plane(pts, w2c, K, H, W, sort=sort)
    bi, bd, bt = brute_force_table(key, depth, H * W, sort)
    np.testing.assert_array_equal(idx, bi)
    np.testing.assert_array_equal(dep.view(np.int32), bd.view(np.int32))
    np.testing.assert_array_equal(table, bt)


def test_pts2plane_invariants():
    W, H = 40, 30
    pts, w2c, K = random_scene(20000, 5, W, H, 12.0)
    key, depth = oracle.project(pts, w2c, K, H, W)
    idx, dep, table = oracle.pts2plane(pts, w2c, K, H, W)
    P = W * H
    n_in = int((key < P).sum())
    cnt = table[:, 1]
    assert cnt[cnt > 0].sum() == n_in
    assert ((table[:, 0] == -1) == (table[:, 1] == -1)).all()
    assert not (cnt == 0).any()
    starts = np.concatenate([[0], np.cumsum(np.maximum(cnt, 0))[:-1]])
    np.testing.assert_array_equal(table[cnt > 0, 0], starts[cnt > 0])
    assert (idx[n_in:] == -1).all() and (dep[n_in:] == -1).all()
    # every in-image point appears exactly once, in its own pixel's run
    assert sorted(idx[:n_in].tolist()) == sorted(np.nonzero(key < P)[0].tolist())
    for p in np.nonzero(cnt > 0)[0][:200]:
        s, c = table[p]
        run = idx[s:s + c]
        assert (key[run] == p).all() and (np.diff(run) > 0).all()


def test_pts2plane_special_cases():
    W, H = 6, 4
    w2c, K = cam_axes(W, H, 3.0)
    # empty input
    idx, dep, table = oracle.pts2plane(np.zeros((0, 3), np.float32), w2c, K, H, W)
    assert idx.size == 0 and (table == -1).all()
    # everything off-image
    pts = np.array([[100.0, 0, 1], [0, -100.0, 1], [0, 0, -0.0]], np.float32)
    idx, dep, table = oracle.pts2plane(pts, w2c, K, H, W)
    assert (idx == -1).all() and (table == -1).all()
    # every point in one pixel: one run of n, in index order
    pts = np.tile(np.array([[0.01, 0.01, 1.0]], np.float32), (50, 1))
    pts[:, 2] = np.linspace(2.0, 1.0, 50, dtype=np.float32)
    idx, dep, table = oracle.pts2plane(pts, w2c, K, H, W)
    p = np.nonzero(table[:, 1] > 0)[0]
    assert p.size == 1 and table[p[0], 0] == 0 and table[p[0], 1] == 50
    np.testing.assert_array_equal(idx, np.arange(50))
    idx, dep, table = oracle.pts2plane(pts, w2c, K, H, W, sort=True)
    np.testing.assert_array_equal(idx, np.arange(50)[::-1])  # depth ascending


def test_scatter_hashtable_reference_semantics_on_presorted_input():
    """scatter_hashtable keeps the given within-pixel order (a stable bubble
    sort by depth when sort is set) and leaves off-image positions at -1."""
    keys = np.array([0, 0, 2, 2, 2, 5, 6, 6], np.int32)  # P = 6 -> key 6 is the sentinel
    pidx = np.array([7, 3, 1, 9, 4, 0, 2, 5], np.int32)
    depth = np.array([0.5, 0.4, 1.0, 0.2, 1.0, 0.3, 9.0, 9.0], np.float32)
    oi, od, oh = oracle.scatter_hashtable(pidx, keys, depth, 2, 3, sort=False)
    np.testing.assert_array_equal(oi, [7, 3, 1, 9, 4, 0, -1, -1])
    np.testing.assert_array_equal(oh, [[0, 2], [-1, -1], [2, 3], [-1, -1], [-1, -1], [5, 1]])
    oi, od, oh = oracle.scatter_hashtable(pidx, keys, depth, 2, 3, sort=True)
    np.testing.assert_array_equal(oi, [3, 7, 9, 1, 4, 0, -1, -1])  # ties (1.0, 1.0) keep order 1, 4
    np.testing.assert_array_equal(od, np.array([0.4, 0.5, 0.2, 1.0, 1.0, 0.3, -1, -1], np.float32))
    # the last pixel's run may reach the end of the array
    oi, od, oh = oracle.scatter_hashtable(pidx[:6], keys[:6], depth[:6], 2, 3, sort=False)
    np.testing.assert_array_equal(oh[5], [5, 1])


def test_scatter_hashtable_sort_long_runs_negative_depths():
    """The depth sort of a long run equals a stable argsort by depth: depths
    descending in input order (the bubble sort's worst case), many ties, and
    negative depths (points behind the camera project into the image too)."""
    rng = np.random.default_rng(11)
    n = 3000
    keys = np.sort(rng.integers(0, 3, size=n)).astype(np.int32)
    depth = np.round(np.linspace(6.0, -3.0, n, dtype=np.float32) * 20) / 20
    depth = depth.astype(np.float32)
    pidx = rng.permutation(n).astype(np.int32)
    oi, od, oh = oracle.scatter_hashtable(pidx, keys, depth, 1, 3, sort=True)
    for p in range(3):
        s, c = oh[p]
        sel = np.nonzero(keys == p)[0]
        order = sel[np.argsort(depth[sel], kind="stable")]
        np.testing.assert_array_equal(oi[s:s + c], pidx[order])
        np.testing.assert_array_equal(od[s:s + c], depth[order])


def test_scatter_top_b_brute_force():
    rng = np.random.default_rng(3)
    H, W, B = 3, 4, 3
    n = 60
    keys = np.sort(rng.integers(0, H * W + 1, size=n)).astype(np.int32)
    depth = rng.uniform(0.1, 5.0, size=n).astype(np.float32)
    depth[::7] = -1.0  # negative depths are skipped
    pidx = rng.permutation(n).astype(np.int32)
    od, oi = oracle.scatter(pidx, keys, depth, H, W, B)
    for p in range(H * W):
        sel = [(depth[i], pidx[i]) for i in range(n) if keys[i] == p and depth[i] >= 0]
        sel = sorted(sel)[:B]  # distinct random depths: no ties
        np.testing.assert_array_equal(od[p, :len(sel)], np.array([d for d, _ in sel], np.float32))
        np.testing.assert_array_equal(oi[p, :len(sel)], [i for _, i in sel])
        assert (oi[p, len(sel):] == -1).all() and (od[p, len(sel):] == -1).all()


def test_scatter_tie_rule():
    """Equal depths: earlier entries win (strict < replacement), and the
    first of equal maxima is evicted."""
    keys = np.zeros(5, np.int32)
    depth = np.array([2.0, 1.0, 2.0, 1.0, 0.5], np.float32)
    pidx = np.array([10, 11, 12, 13, 14], np.int32)
    od, oi = oracle.scatter(pidx, keys, depth, 1, 1, 3)
    # fill: (2,10) (1,11) (2,12) max=first 2.0 at slot 0; (1,13) replaces slot 0;
    # new max = slot 2 (2,12); (0.5,14) replaces it -> {13, 11, 14} sorted by depth (stable)
    np.testing.assert_array_equal(oi[0], [14, 13, 11])
    np.testing.assert_array_equal(od[0], np.array([0.5, 1.0, 1.0], np.float32))
