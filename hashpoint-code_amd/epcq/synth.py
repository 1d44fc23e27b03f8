"""Deterministic synthetic scenes and cameras for the epcq workloads.

Every coordinate is produced with IEEE float64 arithmetic only (+, -, *, /,
sqrt and frexp, which are correctly rounded on every host), from a
counter-based SplitMix64 stream, so the same seed gives bit-identical point
clouds on any machine.  Gaussian variates come from Marsaglia's polar method
with a logarithm evaluated from its atanh series; directions uniform on the
sphere come from rejection sampling in the cube.

Scenes (sizes and seeds of the benchmark workloads):
  sphere_cloud   points on the unit sphere + N(0, 0.005^2) jitter
  box_scene      points on the faces of random axis-aligned boxes (sides
                 U[0.1, 0.5], centres U[-1, 1]^3) and on the plane z = -1,
                 x, y in [-2, 2], allocated by area, + N(0, sigma^2) jitter
Cameras look at the origin in the OpenCV convention (x right, y down, z
forward) with K = [[f, 0, W/2], [0, f, H/2], [0, 0, 1]].
"""
import hashlib
import math
from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_LN2 = 0.6931471805599453


def _mix64(z):
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based SplitMix64: draw k = mix64(key + (counter + 1) * golden)."""

    def __init__(self, seed, stream=0):
        with np.errstate(over="ignore"):
            base = np.array([np.uint64(seed) * _GOLDEN + np.uint64(stream) * _M1], dtype=np.uint64)
        self.key = _mix64(base)[0]
        self.counter = 0

    def bits(self, n):
        idx = np.arange(self.counter + 1, self.counter + 1 + n, dtype=np.uint64)
        self.counter += n
        with np.errstate(over="ignore"):
            return _mix64(self.key + idx * _GOLDEN)

    def uniform(self, n, lo=0.0, hi=1.0):
        """n float64 uniforms in [lo, hi) from the top 53 bits."""
        u = (self.bits(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n):
        """n standard normals (polar method)."""
        out = np.empty(n, dtype=np.float64)
        got = 0
        while got < n:
            need = n - got
            m = int(need * 0.65) + 64  # pairs; acceptance is pi/4
            u = self.uniform(2 * m, -1.0, 1.0).reshape(m, 2)
            s = u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1]
            ok = (s > 0.0) & (s < 1.0)
            u, s = u[ok], s[ok]
            f = np.sqrt(-2.0 * log_series(s) / s)
            z = (u * f[:, None]).reshape(-1)
            take = min(need, z.size)
            out[got:got + take] = z[:take]
            got += take
        return out

    def unit_vectors(self, n):
        """n directions uniform on the unit sphere (cube rejection)."""
        out = np.empty((n, 3), dtype=np.float64)
        got = 0
        while got < n:
            need = n - got
            m = int(need * 2.0) + 64  # acceptance is pi/6
            c = self.uniform(3 * m, -1.0, 1.0).reshape(m, 3)
            r2 = c[:, 0] * c[:, 0] + c[:, 1] * c[:, 1] + c[:, 2] * c[:, 2]
            ok = (r2 > 1e-12) & (r2 <= 1.0)
            c = c[ok] / np.sqrt(r2[ok])[:, None]
            take = min(need, c.shape[0])
            out[got:got + take] = c[:take]
            got += take
        return out


def log_series(x):
    """Natural log of positive finite float64 values: x = m 2^e with m in
    [0.5, 1); log m = 2 atanh((m - 1)/(m + 1)) summed to 1e-19."""
    m, e = np.frexp(x)
    y = (m - 1.0) / (m + 1.0)
    y2 = y * y
    term = y
    acc = y
    for k in range(1, 24):
        term = term * y2
        acc = acc + term / (2 * k + 1)
    return 2.0 * acc + e.astype(np.float64) * _LN2


def sphere_cloud(n, seed=0, jitter=0.005):
    """n points on the unit sphere with per-coordinate N(0, jitter^2) noise."""
    st = Stream(seed, 1)
    p = st.unit_vectors(n)
    p = p + jitter * st.normal(3 * n).reshape(n, 3)
    return p.astype(np.float32)


def box_scene(n, n_boxes=20, seed=0, jitter=0.002, plane_z=-1.0, plane_half=2.0):
    """n points on n_boxes random boxes and the ground plane, by area."""
    st = Stream(seed, 2)
    sides = st.uniform(3 * n_boxes, 0.1, 0.5).reshape(n_boxes, 3)
    centres = st.uniform(3 * n_boxes, -1.0, 1.0).reshape(n_boxes, 3)
    # faces: (centre, axis of the normal, sign, half extents of the two free axes)
    faces = []
    for b in range(n_boxes):
        c, s = centres[b], sides[b]
        for ax in range(3):
            a1, a2 = [a for a in range(3) if a != ax]
            for sign in (-1.0, 1.0):
                faces.append((c, ax, sign, s[ax] / 2.0, a1, a2, s[a1], s[a2]))
    areas = np.array([f[6] * f[7] for f in faces] + [(2 * plane_half) ** 2], dtype=np.float64)
    cdf = np.cumsum(areas / areas.sum())
    cdf[-1] = 1.0
    which = np.searchsorted(cdf, st.uniform(n), side="right")
    which = np.minimum(which, len(areas) - 1)
    a = st.uniform(n)
    b = st.uniform(n)
    pts = np.empty((n, 3), dtype=np.float64)
    plane = which == len(faces)
    pts[plane, 0] = -plane_half + 2 * plane_half * a[plane]
    pts[plane, 1] = -plane_half + 2 * plane_half * b[plane]
    pts[plane, 2] = plane_z
    if faces:
        fc = np.array([f[0] for f in faces])
        fax = np.array([f[1] for f in faces])
        fsign = np.array([f[2] for f in faces])
        fhalf = np.array([f[3] for f in faces])
        fa1 = np.array([f[4] for f in faces])
        fa2 = np.array([f[5] for f in faces])
        fs1 = np.array([f[6] for f in faces])
        fs2 = np.array([f[7] for f in faces])
        on = ~plane
        w = which[on]
        rows = np.nonzero(on)[0]
        cen = fc[w]
        q = cen.copy()
        q[np.arange(w.size), fax[w]] = cen[np.arange(w.size), fax[w]] + fsign[w] * fhalf[w]
        q[np.arange(w.size), fa1[w]] = cen[np.arange(w.size), fa1[w]] + (a[on] - 0.5) * fs1[w]
        q[np.arange(w.size), fa2[w]] = cen[np.arange(w.size), fa2[w]] + (b[on] - 0.5) * fs2[w]
        pts[rows] = q
    pts = pts + jitter * st.normal(3 * n).reshape(n, 3)
    return pts.astype(np.float32)


def exp_series(x):
    """exp(x) for float64 arrays, |x| < 700: x = k ln2 + r with |r| <= ln2 / 2,
    exp(r) from its Taylor series, times 2^k (exact)."""
    x = np.asarray(x, dtype=np.float64)
    k = np.floor(x / _LN2 + 0.5)
    r = x - k * _LN2
    term = np.ones_like(r)
    acc = np.ones_like(r)
    for j in range(1, 22):
        term = term * r / j
        acc = acc + term
    return np.ldexp(acc, k.astype(np.int64))


def _dot3(a, b):
    return float(a[0]) * float(b[0]) + float(a[1]) * float(b[1]) + float(a[2]) * float(b[2])


def log_uniform_radius(n, seed, lo=1e-3, hi=1e-2):
    """Per-point radius exp(U[ln lo, ln hi]) = lo * exp(u ln(hi / lo))."""
    u = Stream(seed, 3).uniform(n)
    span = float(log_series(np.array([hi / lo]))[0])
    return (lo * exp_series(u * span)).astype(np.float32)


@dataclass
class Camera:
    c2w: np.ndarray  # [3, 4] float32
    w2c: np.ndarray  # [3, 4] float32
    K: np.ndarray    # [3, 3] float32
    fx: float
    fy: float
    cx: float
    cy: float
    width: int
    height: int


def look_at_camera(eye, width, height, focal):
    """Camera at `eye` looking at the origin (OpenCV axes)."""
    eye = np.asarray(eye, dtype=np.float64)
    fwd = -eye / math.sqrt(_dot3(eye, eye))
    up = np.array([0.0, 0.0, 1.0])
    if abs(_dot3(fwd, up)) > 0.999:
        up = np.array([0.0, 1.0, 0.0])
    right = np.cross(fwd, up)
    right = right / math.sqrt(_dot3(right, right))
    down = np.cross(fwd, right)
    R = np.stack([right, down, fwd], axis=1)  # columns: camera axes in world
    c2w = np.concatenate([R, eye[:, None]], axis=1)
    t = np.array([-_dot3(R[:, i], eye) for i in range(3)])
    w2c = np.concatenate([R.T, t[:, None]], axis=1)
    K = np.array([[focal, 0.0, width / 2.0], [0.0, focal, height / 2.0], [0.0, 0.0, 1.0]])
    return Camera(c2w.astype(np.float32), w2c.astype(np.float32), K.astype(np.float32), float(np.float32(focal)),
                  float(np.float32(focal)), float(np.float32(width / 2.0)), float(np.float32(height / 2.0)),
                  int(width), int(height))


RESOLUTIONS = [(200, 150), (400, 300), (640, 512), (800, 800), (1600, 1200)]


def nerf_focal(width):
    """NeRF-synthetic field of view: f = 1111 * W / 800."""
    return 1111.0 * width / 800.0


def random_views(n_views, seed_eye, seed_res, radius=4.0, resolutions=RESOLUTIONS):
    """Cameras at radius*normalize(N(0, I3)) with resolutions drawn uniformly."""
    dirs = Stream(seed_eye, 4).normal(3 * n_views).reshape(n_views, 3)
    dirs = dirs / np.sqrt((dirs * dirs).sum(axis=1))[:, None]
    picks = (Stream(seed_res, 5).uniform(n_views) * len(resolutions)).astype(np.int64)
    picks = np.minimum(picks, len(resolutions) - 1)
    cams = []
    for i in range(n_views):
        W, H = resolutions[picks[i]]
        cams.append(look_at_camera(radius * dirs[i], W, H, nerf_focal(W)))
    return cams


def sha256(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


# ---- the benchmark workloads -------------------------------------------------
def workload_sphere_small():
    """100k-point sphere, one 400x400 view at distance 4, f = 555.5, r = 0.01."""
    pts = sphere_cloud(100_000, seed=0, jitter=0.005)
    cam = look_at_camera([0.0, -4.0, 0.0], 400, 400, 555.5)
    return pts, [cam], dict(radius_threshold=0.01, num_point_cloud_per_sp=8, topK=8)


def workload_box_scene(n=1_000_000):
    """1M-point 20-box scene, one 800x800 view, f = 1111, r = 0.005."""
    pts = box_scene(n, n_boxes=20, seed=0, jitter=0.002)
    eye = 4.0 * np.array([1.0, -1.0, 1.0]) / math.sqrt(3.0)
    cam = look_at_camera(eye, 800, 800, 1111.0)
    return pts, [cam], dict(radius_threshold=0.005, num_point_cloud_per_sp=8, topK=8)


def workload_batched(n=1_000_000, n_views=64):
    """The 1M-point box scene seen by 64 random views of mixed resolution."""
    pts = box_scene(n, n_boxes=20, seed=0, jitter=0.002)
    cams = random_views(n_views, seed_eye=2, seed_res=3)
    return pts, cams, dict(radius_threshold=0.005, num_point_cloud_per_sp=8, topK=8)


def window_for_radius(r_max, focal, distance):
    """Window size k (odd) whose half-width covers the largest radius r_max
    projected at the viewing distance: k = 2 ceil(r_max f / distance) + 1."""
    return 2 * int(math.ceil(r_max * focal / distance)) + 1


def workload_large(n=20_000_000, n_views=16):
    """20M points on 200 boxes + plane, per-point radius, 16 views 1920x1080;
    the window is sized from the largest radius at the eye distance 4
    (k = 15 for r_max = 0.01, f = 2666)."""
    pts = box_scene(n, n_boxes=200, seed=30, jitter=0.002)
    radius = log_uniform_radius(n, seed=31)
    dirs = Stream(32, 4).normal(3 * n_views).reshape(n_views, 3)
    dirs = dirs / np.sqrt((dirs * dirs).sum(axis=1))[:, None]
    f = nerf_focal(1920)
    cams = [look_at_camera(4.0 * d, 1920, 1080, f) for d in dirs]
    k = window_for_radius(0.01, f, 4.0)
    return pts, radius, cams, dict(radius_threshold=0.01, num_point_cloud_per_sp=16, topK=16, kernel_size=k)
