"""Pins the oracle's scalar pieces: exp, neighbour offsets, rays, projection
onto a ray.  Each check compares with mathematics computed independently
(float64 / closed forms / the reference's float formula)."""
import math

import numpy as np
import pytest

from oracle import oracle


def ulp_distance(a, b):
    ia = np.array([a], np.float32).view(np.int32)[0].astype(np.int64)
    ib = np.array([b], np.float32).view(np.int32)[0].astype(np.int64)
    return abs(ia - ib)


@pytest.mark.parametrize("lo,hi,count", [(-20.0, 20.0, 4001), (-0.1, 0.1, 2001), (-87.0, 88.7, 3001),
                                         (-103.9, -87.4, 501)])
def test_expf_error_bound(lo, hi, count):
    worst = 0
    for x in np.linspace(lo, hi, count, dtype=np.float32):
        got = oracle.expf(x)
        exact = np.float32(math.exp(float(x)))  # correctly rounded reference value
        if exact == 0.0 or not np.isfinite(exact):
            continue
        if abs(float(exact)) < 1.2e-38:  # subnormal: absolute error within half a subnormal step twice
            assert abs(float(got) - math.exp(float(x))) <= 2 * 1.4e-45
            continue
        worst = max(worst, ulp_distance(got, exact))
    assert worst <= 2, f"exp error {worst} ulp"


def test_expf_special_values():
    assert oracle.expf(0.0) == 1.0
    assert oracle.expf(1.0) == np.float32(math.e)
    assert oracle.expf(89.0) == np.inf
    assert oracle.expf(-104.0) == 0.0
    assert np.isnan(oracle.expf(float("nan")))
    # exp(-x) * exp(x) ~ 1
    for x in (0.3, 1.7, 11.0):
        assert abs(float(oracle.expf(x)) * float(oracle.expf(-x)) - 1.0) < 4e-7


@pytest.mark.parametrize("k", list(range(1, 16)))
def test_neighbor_offsets_match_reference_float_formula(k):
    """calculate_neighbor_uv (cuda_util.cuh:382-388) evaluated literally in
    float32: kernel_offset = ceil(k / 2.0f), dv = ceil((float)(l+1)/k) - ko,
    du = l % k - (ko - 1)."""
    ko = int(np.ceil(np.float32(k) / np.float32(2.0)))
    seen = set()
    for lane in range(k * k):
        dv = int(np.ceil(np.float32(lane + 1) / np.float32(k))) - ko
        du = lane % k - (ko - 1)
        assert oracle.neighbor_offset(k, lane) == (du, dv)
        seen.add((du, dv))
    assert len(seen) == k * k  # a full k x k window, row-major
    if k == 5:
        assert oracle.neighbor_offset(5, 0) == (-2, -2)
        assert oracle.neighbor_offset(5, 12) == (0, 0)
        assert oracle.neighbor_offset(5, 24) == (2, 2)
    if k == 4:  # even windows are asymmetric
        assert oracle.neighbor_offset(4, 0) == (-1, -1)
        assert oracle.neighbor_offset(4, 15) == (2, 2)


def identity_camera(W=8, H=6, f=4.0, eye=(0.0, 0.0, 0.0)):
    c2w = np.zeros((3, 4), np.float32)
    c2w[:, :3] = np.eye(3)
    c2w[:, 3] = eye
    w2c = np.zeros((3, 4), np.float32)
    w2c[:, :3] = np.eye(3)
    w2c[:, 3] = -np.asarray(eye, np.float32)
    K = np.array([[f, 0, W / 2], [0, f, H / 2], [0, 0, 1]], np.float32)
    return oracle.camera(c2w, w2c, K, f, f, W / 2, H / 2, W, H), c2w, w2c, K


def test_ray_closed_form():
    cam, *_ = identity_camera(W=8, H=6, f=4.0, eye=(1.0, 2.0, 3.0))
    for u, v in [(0, 0), (3, 2), (7, 5), (4, 3)]:
        o, d = oracle.ray(cam, u, v)
        x = (u + 0.5 - 4.0) / 4.0
        y = (v + 0.5 - 3.0) / 4.0
        n = math.sqrt(x * x + y * y + 1.0)
        np.testing.assert_array_equal(o, np.array([1, 2, 3], np.float32))
        np.testing.assert_allclose(d, [x / n, y / n, 1.0 / n], rtol=0, atol=2e-7)
        assert abs(float(np.dot(d.astype(np.float64), d)) - 1.0) < 1e-6


def test_ray_pixel_centre_uses_double_precision():
    # (u + 0.5 - cx) / fx evaluated in double and rounded once to float
    cam, *_ = identity_camera(W=1000, H=2, f=3.0)
    cam.cx = np.float32(1.0 / 3.0)
    o, d = oracle.ray(cam, 777, 0)
    x = np.float32((777 + 0.5 - float(np.float32(1.0 / 3.0))) / 3.0)
    y = np.float32((0 + 0.5 - 1.0) / 3.0)
    s = np.float32(np.float32(x * x) + np.float32(y * y)) + np.float32(1.0)
    rn = np.float32(1.0) / np.float32(np.sqrt(s))
    np.testing.assert_array_equal(d, np.array([x * rn, y * rn, rn], np.float32))


def test_point_ray_projection_closed_form():
    o = np.array([0.5, -1.0, 2.0], np.float32)
    d = np.array([0.0, 0.6, 0.8], np.float32)
    # on the ray at t = 2.5
    p = o + np.float32(2.5) * d
    t, dist = oracle.point_ray(o, d, p)
    assert abs(t - 2.5) < 1e-6 and dist < 1e-6
    # 0.1 off the ray along x (perpendicular to d) at t = 4
    p = o + np.float32(4.0) * d + np.array([0.1, 0.0, 0.0], np.float32)
    t, dist = oracle.point_ray(o, d, p)
    assert abs(t - 4.0) < 2e-6 and abs(dist - 0.1) < 1e-6
    # behind the origin: negative t
    p = o - np.float32(1.5) * d
    t, dist = oracle.point_ray(o, d, p)
    assert abs(t + 1.5) < 1e-6 and dist < 1e-6


def test_point_ray_random_vs_float64():
    rng = np.random.default_rng(7)
    for _ in range(300):
        o = rng.normal(size=3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        p = (o + rng.uniform(1, 5) * d + rng.normal(scale=0.01, size=3)).astype(np.float32)
        t, dist = oracle.point_ray(o, d, p)
        o64, d64, p64 = o.astype(np.float64), d.astype(np.float64), p.astype(np.float64)
        t64 = np.dot(p64 - o64, d64) / np.dot(d64, d64)
        dist64 = np.linalg.norm(p64 - (o64 + t64 * d64))
        assert abs(t - t64) <= 1e-6 * max(1.0, abs(t64))
        assert abs(dist - dist64) <= 2e-6
