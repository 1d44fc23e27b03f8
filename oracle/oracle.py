"""ctypes binding of the CPU oracle (liboracle_epcq.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU-baseline
leg of bench.py as the checker.  The product never imports this module.
Arrays are numpy, C-contiguous, host memory.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_epcq.so")

MODE_COMPAT = 0
MODE_KNN = 1


def build():
    """Compiles the oracle with its Makefile (g++, OpenMP, no contraction)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle_epcq.so"], check=True)
    return LIB_PATH


class OrcCamera(ctypes.Structure):
    _fields_ = [("c2w", ctypes.c_float * 12), ("w2c", ctypes.c_float * 12), ("K", ctypes.c_float * 9),
                ("fx", ctypes.c_float), ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32)]


class OrcOptions(ctypes.Structure):
    _fields_ = [("kernel_size", ctypes.c_int32), ("t_intvl", ctypes.c_float),
                ("radius_threshold", ctypes.c_float), ("sigma_threshold", ctypes.c_float),
                ("background_brightness", ctypes.c_float), ("stop_threshold", ctypes.c_float),
                ("num_sample_points_per_ray", ctypes.c_int32), ("num_point_cloud_per_sp", ctypes.c_int32),
                ("num_point_per_ray", ctypes.c_int32), ("sdf2weight_gaussian_alpha", ctypes.c_float),
                ("sdf2weight_gaussian_gamma", ctypes.c_float), ("topK", ctypes.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_expf.restype = ctypes.c_float
        _lib.orc_expf.argtypes = [ctypes.c_float]
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else ctypes.c_void_p(0)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def set_threads(n):
    lib().orc_set_threads(int(n))


def camera(c2w, w2c, K, fx, fy, cx, cy, width, height):
    cam = OrcCamera()
    cam.c2w[:] = [float(v) for v in np.asarray(c2w, dtype=np.float32)[:3, :4].reshape(-1)]
    cam.w2c[:] = [float(v) for v in np.asarray(w2c, dtype=np.float32)[:3, :4].reshape(-1)]
    cam.K[:] = [float(v) for v in np.asarray(K, dtype=np.float32).reshape(-1)]
    cam.fx, cam.fy, cam.cx, cam.cy = float(fx), float(fy), float(cx), float(cy)
    cam.width, cam.height = int(width), int(height)
    return cam


def camera_from(c):
    """OrcCamera from any object with c2w, w2c, K, fx, fy, cx, cy, width, height."""
    def host(x):
        return x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)
    return camera(host(c.c2w), host(c.w2c), host(c.K), float(c.fx), float(c.fy), float(c.cx), float(c.cy),
                  int(c.width), int(c.height))


def options(**kw):
    defaults = dict(kernel_size=5, t_intvl=0.01, radius_threshold=0.016, sigma_threshold=1e-4,
                    background_brightness=1.0, stop_threshold=0.001, num_sample_points_per_ray=16,
                    num_point_cloud_per_sp=4, num_point_per_ray=256, sdf2weight_gaussian_alpha=0.2,
                    sdf2weight_gaussian_gamma=0.9, topK=8)
    defaults.update({k: v for k, v in kw.items() if k in defaults})
    o = OrcOptions()
    for name, ctype in OrcOptions._fields_:
        setattr(o, name, int(defaults[name]) if ctype is ctypes.c_int32 else float(defaults[name]))
    return o


def options_from(opt):
    return options(**{name: getattr(opt, name) for name, _ in OrcOptions._fields_})


def expf(x):
    return np.float32(lib().orc_expf(float(x)))


def neighbor_offset(k, lane):
    du, dv = ctypes.c_int32(), ctypes.c_int32()
    lib().orc_neighbor_offset(int(k), int(lane), ctypes.byref(du), ctypes.byref(dv))
    return du.value, dv.value


def ray(cam, u, v):
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    lib().orc_ray(ctypes.byref(cam), int(u), int(v), _p(o), _p(d))
    return o, d


def point_ray(o, d, p):
    t, dist = ctypes.c_float(), ctypes.c_float()
    o, d, p = _c(o, np.float32), _c(d, np.float32), _c(p, np.float32)
    lib().orc_point_ray(_p(o), _p(d), _p(p), ctypes.byref(t), ctypes.byref(dist))
    return np.float32(t.value), np.float32(dist.value)


def project(xyz, w2c, K, H, W):
    xyz = _c(xyz, np.float32)
    n = xyz.shape[0]
    key = np.empty(n, np.int32)
    depth = np.empty(n, np.float32)
    lib().orc_project(_p(xyz), ctypes.c_int64(n), _p(_c(np.asarray(w2c)[:3, :4], np.float32)),
                      _p(_c(K, np.float32)), int(H), int(W), _p(key), _p(depth))
    return key, depth


def sort_by_key(key, depth):
    key, depth = _c(key, np.int32), _c(depth, np.float32)
    n = key.size
    ks = np.empty(n, np.int32)
    ix = np.empty(n, np.int32)
    ds = np.empty(n, np.float32)
    lib().orc_sort_by_key(_p(key), _p(depth), ctypes.c_int64(n), _p(ks), _p(ix), _p(ds))
    return ks, ix, ds


def scatter_hashtable(point_idx, keys, depth, H, W, sort):
    point_idx, keys, depth = _c(point_idx, np.int32), _c(keys, np.int32), _c(depth, np.float32)
    n = keys.size
    oi = np.empty(n, np.int32)
    od = np.empty(n, np.float32)
    oh = np.empty((H * W, 2), np.int32)
    lib().orc_scatter_hashtable(_p(point_idx), _p(keys), _p(depth), ctypes.c_int64(n), int(H), int(W),
                                int(bool(sort)), _p(oi), _p(od), _p(oh))
    return oi, od, oh


def pts2plane(xyz, w2c, K, H, W, sort=False):
    xyz = _c(xyz, np.float32)
    n = xyz.shape[0]
    oi = np.empty(n, np.int32)
    od = np.empty(n, np.float32)
    oh = np.empty((H * W, 2), np.int32)
    lib().orc_pts2plane(_p(xyz), ctypes.c_int64(n), _p(_c(np.asarray(w2c)[:3, :4], np.float32)),
                        _p(_c(K, np.float32)), int(H), int(W), int(bool(sort)), _p(oi), _p(od), _p(oh))
    return oi, od, oh


def scatter(point_idx, keys, depth, H, W, bins):
    point_idx, keys, depth = _c(point_idx, np.int32), _c(keys, np.int32), _c(depth, np.float32)
    n = keys.size
    od = np.empty((H * W, bins), np.float32)
    oi = np.empty((H * W, bins), np.int32)
    lib().orc_scatter(_p(point_idx), _p(keys), _p(depth), ctypes.c_int64(n), int(H), int(W), int(bins), _p(od),
                      _p(oi))
    return od, oi


def quick_sampling(xyz, cam, opt, sorted_idx, hash_table, radius=None, mode=MODE_COMPAT, rays=None,
                   with_nb=False):
    """Dense outputs (sp_xyz [R,S,3], pc_idx [R,S,Kc], sp_t [R,S], mask [R,S]
    bool[, nb_t, nb_dist [R,S,Kc]]); R = H*W or len(rays)."""
    xyz = _c(xyz, np.float32)
    sorted_idx = _c(sorted_idx, np.int32)
    hash_table = _c(hash_table, np.int32)
    radius = _c(radius, np.float32) if radius is not None else None
    rays = _c(rays, np.int32) if rays is not None else None
    R = rays.size if rays is not None else cam.width * cam.height
    S, Kc = opt.num_sample_points_per_ray, opt.num_point_cloud_per_sp
    sp_xyz = np.empty((R, S, 3), np.float32)
    pc_idx = np.empty((R, S, Kc), np.int32)
    sp_t = np.empty((R, S), np.float32)
    mask = np.empty((R, S), np.uint8)
    nb_t = np.empty((R, S, Kc), np.float32) if with_nb else None
    nb_d = np.empty((R, S, Kc), np.float32) if with_nb else None
    lib().orc_quick_sampling(_p(xyz), ctypes.c_int64(xyz.shape[0]), ctypes.byref(cam), ctypes.byref(opt),
                             _p(sorted_idx), _p(hash_table), _p(radius), int(mode), _p(rays), ctypes.c_int64(R),
                             _p(sp_xyz), _p(pc_idx), _p(sp_t), _p(mask), _p(nb_t), _p(nb_d))
    out = (sp_xyz, pc_idx, sp_t, mask.astype(bool))
    return out + (nb_t, nb_d) if with_nb else out


def quick_query(xyz, cam, opt, sorted_idx, hash_table, rays):
    xyz = _c(xyz, np.float32)
    rays = _c(rays, np.int32)
    out = np.empty((rays.size, opt.num_point_per_ray), np.int32)
    lib().orc_quick_query(_p(xyz), ctypes.c_int64(xyz.shape[0]), ctypes.byref(cam), ctypes.byref(opt),
                          _p(_c(sorted_idx, np.int32)), _p(_c(hash_table, np.int32)), _p(rays),
                          ctypes.c_int64(rays.size), _p(out))
    return out


def idw(xyz, feat, sp_xyz, nb_idx, bf16=False):
    """feat: float32 [n,F], or uint16 bf16 bit patterns [n,F] with bf16=True."""
    xyz = _c(xyz, np.float32)
    feat = _c(feat, np.uint16 if bf16 else np.float32)
    sp_xyz = _c(sp_xyz, np.float32).reshape(-1, 3)
    nb_idx = _c(nb_idx, np.int32)
    K = nb_idx.shape[-1]
    nb_idx = nb_idx.reshape(-1, K)
    M = sp_xyz.shape[0]
    F = feat.shape[1]
    w = np.empty((M, K), np.float32)
    agg = np.empty((M, F), np.float32)
    lib().orc_idw(_p(xyz), _p(feat), int(bf16), int(F), _p(sp_xyz), _p(nb_idx), ctypes.c_int64(M), int(K), _p(w),
                  _p(agg))
    return w, agg


def clamp_t(t, vsize=0.004):
    t = _c(t, np.float32)
    R, S = t.shape
    out = np.empty((R, 0 if S == 1 else S), np.float32)
    lib().orc_clamp_t(_p(t), ctypes.c_int64(R), int(S), ctypes.c_float(vsize), _p(out))
    return out


def render(sigma, rgb, white_bg=False, t=None, delta=None, vsize=0.004):
    sigma, rgb = _c(sigma, np.float32), _c(rgb, np.float32)
    R, S = sigma.shape
    t = _c(t, np.float32) if t is not None else None
    delta = _c(delta, np.float32) if delta is not None else None
    out = np.empty((R, 3), np.float32)
    lib().orc_render(_p(t), _p(delta), ctypes.c_float(vsize), _p(sigma), _p(rgb), ctypes.c_int64(R), int(S),
                     int(bool(white_bg)), _p(out))
    return out


def render_backward(sigma, rgb, grad_out, white_bg=False, t=None, delta=None, vsize=0.004):
    sigma, rgb, grad_out = _c(sigma, np.float32), _c(rgb, np.float32), _c(grad_out, np.float32)
    R, S = sigma.shape
    t = _c(t, np.float32) if t is not None else None
    delta = _c(delta, np.float32) if delta is not None else None
    gs = np.empty((R, S), np.float32)
    gc = np.empty((R, S, 3), np.float32)
    lib().orc_render_backward(_p(t), _p(delta), ctypes.c_float(vsize), _p(sigma), _p(rgb), _p(grad_out),
                              ctypes.c_int64(R), int(S), int(bool(white_bg)), _p(gs), _p(gc))
    return gs, gc
