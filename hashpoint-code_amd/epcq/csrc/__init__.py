"""`epcq.csrc`: the op module of the epcq path, backed by libepcq_hip.so.

Drop-in replacement of the reference's pybind11 extension `epcq.csrc`
(epcq/csrc/epcq.cpp:5-55): the same four ops (`scatter`, `scatter_hashtable`,
`quick_sampling`, `quick_query_for_nearby_point`) with the same tensor
signatures, return tuples and error texts, and the same four spec classes with
read/write attributes.  It adds `build_hashtable` (fused pts2plane),
`knn_sampling` (compact K-nearest sampling with ray depths, distances and IDW
features), `idw_features`, and `max_live_samples`.

Each op checks its tensors on the host, allocates its outputs with torch on
the input's device, and calls the C ABI (include/epcq_abi.h) through ctypes
on the current HIP stream of that device.  The ABI never allocates device
memory and never synchronises.  There is no CPU fallback: if the HIP library
is missing this module fails to import.
"""
import ctypes
import os

import torch

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libepcq_hip.so")

MODE_COMPAT = 0
MODE_KNN = 1
_MODES = {"compat": MODE_COMPAT, "knn": MODE_KNN, MODE_COMPAT: MODE_COMPAT, MODE_KNN: MODE_KNN}


class _Camera(ctypes.Structure):
    _fields_ = [
        ("c2w", ctypes.c_void_p),
        ("w2c", ctypes.c_void_p),
        ("fx", ctypes.c_float),
        ("fy", ctypes.c_float),
        ("cx", ctypes.c_float),
        ("cy", ctypes.c_float),
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
    ]


class _Options(ctypes.Structure):
    _fields_ = [
        ("kernel_size", ctypes.c_int32),
        ("t_intvl", ctypes.c_float),
        ("radius_threshold", ctypes.c_float),
        ("sigma_threshold", ctypes.c_float),
        ("background_brightness", ctypes.c_float),
        ("stop_threshold", ctypes.c_float),
        ("num_sample_points_per_ray", ctypes.c_int32),
        ("num_point_cloud_per_sp", ctypes.c_int32),
        ("num_point_per_ray", ctypes.c_int32),
        ("sdf2weight_gaussian_alpha", ctypes.c_float),
        ("sdf2weight_gaussian_gamma", ctypes.c_float),
        ("topK", ctypes.c_int32),
    ]


class _Tuning(ctypes.Structure):
    _fields_ = [("tile_w", ctypes.c_int32), ("tile_h", ctypes.c_int32), ("tile_cap", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("heavy", ctypes.c_int32), ("list_wave", ctypes.c_int32)]


TUNE_XCD_ORDER = 1


class _ViewDesc(ctypes.Structure):
    _fields_ = [("xyz", ctypes.c_void_p), ("n", ctypes.c_int64), ("c2w", ctypes.c_float * 12),
                ("w2c", ctypes.c_float * 12), ("K", ctypes.c_float * 9), ("fx", ctypes.c_float),
                ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32)]


class _BatchViewOut(ctypes.Structure):
    _fields_ = [("sorted_idx", ctypes.c_void_p), ("hash", ctypes.c_void_p), ("rays", ctypes.c_int64),
                ("ray_offset", ctypes.c_int64), ("n_samples", ctypes.c_void_p), ("sample_t", ctypes.c_void_p),
                ("nb_idx", ctypes.c_void_p), ("nb_t", ctypes.c_void_p), ("nb_dist", ctypes.c_void_p)]


_AGG_FIELDS = ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4", "wa", "ba", "c1", "cb1", "c2", "cb2", "c3", "cb3",
               "c4", "cb4")


class _AggParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in _AGG_FIELDS]


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_F32 = ctypes.c_float
_SZ = ctypes.c_size_t

# name -> (restype, argtypes); the declarations of include/epcq_abi.h
ABI_SYMBOLS = {
    "epcq_abi_version": (_I32, []),
    "epcq_last_error": (ctypes.c_char_p, []),
    "epcq_set_tuning": (_I32, [ctypes.POINTER(_Tuning)]),
    "epcq_get_tuning": (_I32, [ctypes.POINTER(_Tuning)]),
    "epcq_squared_threshold": (_F32, [_F32]),
    "epcq_scatter_hashtable_workspace_size": (_SZ, [_I64, _I32]),
    "epcq_scatter_hashtable": (_I32, [_P, _P, _P, _I64, _I32, _I32, _I32, _P, _P, _P, _P, _SZ, _P]),
    "epcq_scatter": (_I32, [_P, _P, _P, _I64, _I32, _I32, _I32, _P, _P, _P]),
    "epcq_build_hashtable_workspace_size": (_SZ, [_I64, _I32, _I32]),
    "epcq_build_hashtable": (_I32, [_P, _I64, _P, _P, _I32, _I32, _I32, _P, _P, _P, _P, _SZ, _P]),
    "epcq_build_hashtable_noalloc": (_I32, [_P, _I64, _P, _P, _I32, _I32, _I32, _P, _P, _P, _P]),
    "epcq_quick_sampling": (_I32, [_P, _I64, ctypes.POINTER(_Camera), ctypes.POINTER(_Options), _P, _P,
                                   _I32, _P, _P, _P, _P, _P, _SZ, _P]),
    "epcq_quick_sampling_rays": (_I32, [_P, _I64, ctypes.POINTER(_Camera), ctypes.POINTER(_Options), _P, _P,
                                        _I32, _P, _I64, _P, _P, _P, _P, _P, _SZ, _P]),
    "epcq_sampling_workspace_size": (_SZ, [_I64, _I32, _I32]),
    "epcq_quick_query": (_I32, [_P, _I64, ctypes.POINTER(_Camera), ctypes.POINTER(_Options), _P, _P, _P,
                                _I64, _P, _P]),
    "epcq_knn_sampling": (_I32, [_P, _I64, ctypes.POINTER(_Camera), ctypes.POINTER(_Options), _P, _P, _P,
                                 _P, _I64, _I32, _I32, _P, _P, _P, _P, _P, _P, _I32, _I32, _P, _P, _P, _SZ,
                                 _P]),
    "epcq_max_live_samples": (_I32, [ctypes.POINTER(_Options), _F32]),
    "epcq_idw_features": (_I32, [_P, _P, _I32, _I32, _P, _P, _I64, _I32, _P, _P, _P]),
    "epcq_debug_expf": (_I32, [_P, _P, _I64, _P]),
    "epcq_aggregator_workspace_size": (_SZ, [_I64]),
    "epcq_aggregator_forward": (_I32, [_P, _P, _P, _P, _P, _P, _P, _I64, _I32, _I32, ctypes.POINTER(_AggParams), _P, _P,
                                       _P, _SZ, _P]),
    "epcq_clamp_t": (_I32, [_P, _I64, _I32, _F32, _P, _P]),
    "epcq_render_rays": (_I32, [_P, _P, _F32, _P, _P, _I64, _I32, _I32, _P, _P]),
    "epcq_render_rays_backward": (_I32, [_P, _P, _F32, _P, _P, _P, _I64, _I32, _I32, _P, _P, _P]),
    "epcq_batch_arena_bytes": (_SZ, [ctypes.POINTER(_ViewDesc), _I32, _I32, ctypes.POINTER(_Options), _I32, _I32]),
    "epcq_batch_create": (_I32, [ctypes.POINTER(_Options), _I32, _I32, _I32, _P, _F32, ctypes.POINTER(_Tuning), _P,
                                 _SZ, ctypes.POINTER(ctypes.c_void_p)]),
    "epcq_batch_set_views": (_I32, [_P, ctypes.POINTER(_ViewDesc), _I32, _P]),
    "epcq_batch_view_outputs": (_I32, [_P, _I32, ctypes.POINTER(_BatchViewOut)]),
    "epcq_batch_header_bytes": (_SZ, [_I32]),
    "epcq_batch_view_bytes": (_SZ, [ctypes.POINTER(_ViewDesc), ctypes.POINTER(_Options), _I32, _I32]),
    "epcq_batch_rays": (_I64, [_P]),
    "epcq_batch_run": (_I32, [_P, _P]),
    "epcq_batch_capture": (_I32, [_P, _P]),
    "epcq_batch_launch": (_I32, [_P, _P]),
    "epcq_batch_destroy": (None, [_P]),
    "epcq_batch_profile": (_I32, [_P, _P, ctypes.POINTER(ctypes.c_float)]),
}


def load_library(path=_LIB_PATH):
    """Loads libepcq_hip.so and declares every ABI symbol; raises if absent."""
    if not os.path.exists(path):
        raise ImportError(
            f"epcq.csrc: HIP library not found at {path}; build it with "
            "`python __graft_entry__.py` or `python hashpoint-code_amd/build_lib.py`")
    lib = ctypes.CDLL(path)
    for name, (res, args) in ABI_SYMBOLS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = load_library()
LIBRARY_PATH = _LIB_PATH


# ---------------------------------------------------------------- spec classes
class NeuralPointsSpec:
    """NeuralPointsSpec (epcq/csrc/include/data_spec.hpp:8-13)."""

    def __init__(self):
        self.xyz_data = None

    def check(self):
        _check_input(self.xyz_data, "xyz_data")


class CameraSpec:
    """CameraSpec (epcq/csrc/include/data_spec.hpp:15-31)."""

    def __init__(self):
        self.c2w = None
        self.w2c = None
        self.fx = 0.0
        self.fy = 0.0
        self.cx = 0.0
        self.cy = 0.0
        self.width = 0
        self.height = 0

    def check(self):
        _check_input(self.c2w, "c2w")
        _check_input(self.w2c, "w2c")
        if not self.c2w.is_floating_point():
            raise RuntimeError("Expected c2w.is_floating_point() to be true, but got false.")
        if self.c2w.dim() != 2:
            raise RuntimeError("Expected c2w.ndimension() == 2 to be true, but got false.")
        if self.c2w.size(1) != 4:
            raise RuntimeError("Expected c2w.size(1) == 4 to be true, but got false.")


class RenderOptions:
    """RenderOptions (epcq/csrc/include/data_spec.hpp:49-69)."""

    _FIELDS = [name for name, _ in _Options._fields_]

    def __init__(self):
        for name, ctype in _Options._fields_:
            setattr(self, name, 0 if ctype is ctypes.c_int32 else 0.0)


class NeuralPointsGrads:
    """NeuralPointsGrads (epcq/csrc/include/data_spec.hpp:72-81)."""

    def __init__(self):
        self.grad_sigma_out = None
        self.grad_sh_out = None
        self.mask_out = None

    def check(self):
        _check_input(self.grad_sigma_out, "grad_sigma_out")
        _check_input(self.grad_sh_out, "grad_sh_out")
        _check_input(self.mask_out, "mask_out")


# ---------------------------------------------------------------- helpers
def _check_input(x, name):
    """CHECK_INPUT (epcq/csrc/include/util.hpp:4-11)."""
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not x.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


_DTYPE_NAME = {torch.int32: "Int", torch.float32: "Float", torch.int64: "Long", torch.bool: "Bool",
               torch.float64: "Double", torch.bfloat16: "BFloat16", torch.float16: "Half",
               torch.uint8: "Byte", torch.int16: "Short", torch.int8: "Char"}


def _check_dtype(x, dtype, name):
    if x.dtype != dtype:
        raise RuntimeError(
            f"{name}: expected scalar type {_DTYPE_NAME.get(dtype, dtype)} but found "
            f"{_DTYPE_NAME.get(x.dtype, x.dtype)}")


def _check_same_device(device, *tensors):
    for name, t in tensors:
        if t is not None and t.device != device:
            raise RuntimeError(f"{name} must be on {device}, found {t.device}")


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _raise_on(rc):
    if rc != 0:
        msg = _lib.epcq_last_error()
        raise RuntimeError(msg.decode() if msg else f"epcq error {rc}")


def _img_size(img_size):
    h, w = img_size
    return int(h), int(w)


def _camera_struct(cam):
    return _Camera(_ptr(cam.c2w), _ptr(cam.w2c), float(cam.fx), float(cam.fy), float(cam.cx),
                   float(cam.cy), int(cam.width), int(cam.height))


def _options_struct(opt):
    vals = []
    for name, ctype in _Options._fields_:
        v = getattr(opt, name)
        vals.append(int(v) if ctype is ctypes.c_int32 else float(v))
    return _Options(*vals)


def _hashtable_view(hashtable, cam_h, cam_w):
    """The [P,2] / [H,W,2] hashtable check of epcq_render_topk.cu:351-357."""
    if hashtable.dim() == 2:
        hashtable = hashtable.view(cam_h, cam_w, 2)
    if hashtable.dim() != 3 or hashtable.size(0) != cam_h or hashtable.size(1) != cam_w \
            or hashtable.size(2) != 2:
        dims = list(hashtable.shape) + [0, 0, 0]
        print(f"hashtable size: {dims[0]}, {dims[1]}, {dims[2]}")
        raise RuntimeError("hashtable size error")
    return hashtable


def _check_spec_inputs(neural_points, cam, sorted_point_idx_list, hashtable):
    neural_points.check()
    cam.check()
    _check_input(sorted_point_idx_list, "sorted_point_idx_list")
    _check_input(hashtable, "hashtable")
    _check_dtype(neural_points.xyz_data, torch.float32, "xyz_data")
    if neural_points.xyz_data.dim() != 2 or neural_points.xyz_data.size(1) != 3:
        raise RuntimeError("xyz_data must have shape [N, 3]")
    _check_dtype(cam.c2w, torch.float32, "c2w")
    _check_dtype(cam.w2c, torch.float32, "w2c")
    if cam.c2w.size(0) < 3 or cam.w2c.dim() != 2 or cam.w2c.size(0) < 3 or cam.w2c.size(1) != 4:
        raise RuntimeError("c2w and w2c must have at least 3 rows of 4 values")
    _check_dtype(sorted_point_idx_list, torch.int32, "sorted_point_idx_list")
    _check_dtype(hashtable, torch.int32, "hashtable")
    device = neural_points.xyz_data.device
    _check_same_device(device, ("c2w", cam.c2w), ("w2c", cam.w2c),
                       ("sorted_point_idx_list", sorted_point_idx_list), ("hashtable", hashtable))
    return device


# ---------------------------------------------------------------- ops
def scatter_hashtable(point_idx_list, point2img_idx_list, depth_list, img_size, sort):
    """scatter_hashtable (epcq/csrc/rasterization.cu:130-171).

    Returns (sorted_point_idx_list int32 [N], sorted_depth_list float32 [N],
    hash_table int32 [H*W, 2])."""
    _check_input(point_idx_list, "point_idx_list")
    _check_input(point2img_idx_list, "point2img_idx_list")
    _check_input(depth_list, "depth_list")
    _check_dtype(point_idx_list, torch.int32, "point_idx_list")
    _check_dtype(point2img_idx_list, torch.int32, "point2img_idx_list")
    _check_dtype(depth_list, torch.float32, "depth_list")
    n = point_idx_list.numel()
    if point2img_idx_list.numel() != n or depth_list.numel() != n:
        raise RuntimeError("point_idx_list, point2img_idx_list and depth_list must have the same length")
    H, W = _img_size(img_size)
    dev = point_idx_list.device
    _check_same_device(dev, ("point2img_idx_list", point2img_idx_list), ("depth_list", depth_list))
    out_idx = torch.empty(n, dtype=torch.int32, device=dev)
    out_depth = torch.empty(n, dtype=torch.float32, device=dev)
    out_hash = torch.empty((H * W, 2), dtype=torch.int32, device=dev)
    sort = int(bool(sort))
    ws_bytes = _lib.epcq_scatter_hashtable_workspace_size(n, sort)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev) if ws_bytes else None
    with torch.cuda.device(dev):
        _raise_on(_lib.epcq_scatter_hashtable(_ptr(point_idx_list), _ptr(point2img_idx_list), _ptr(depth_list),
                                              n, H, W, sort, _ptr(out_idx), _ptr(out_depth),
                                              _ptr(out_hash), _ptr(ws) if ws is not None else None, ws_bytes,
                                              _stream(dev)))
    return out_idx, out_depth, out_hash


def scatter(point_idx_list, point2img_idx_list, depth_list, img_size, bins):
    """scatter (epcq/csrc/rasterization.cu:88-128).

    Returns (buffer_depth float32 [H*W, bins], buffer_point_index int32
    [H*W, bins]); per pixel the `bins` smallest depths ascending, -1 padded."""
    _check_input(point_idx_list, "point_idx_list")
    _check_input(point2img_idx_list, "point2img_idx_list")
    _check_input(depth_list, "depth_list")
    _check_dtype(point_idx_list, torch.int32, "point_idx_list")
    _check_dtype(point2img_idx_list, torch.int32, "point2img_idx_list")
    _check_dtype(depth_list, torch.float32, "depth_list")
    n = point_idx_list.numel()
    if point2img_idx_list.numel() != n or depth_list.numel() != n:
        raise RuntimeError("point_idx_list, point2img_idx_list and depth_list must have the same length")
    H, W = _img_size(img_size)
    bins = int(bins)
    dev = point_idx_list.device
    _check_same_device(dev, ("point2img_idx_list", point2img_idx_list), ("depth_list", depth_list))
    out_depth = torch.empty((H * W, bins), dtype=torch.float32, device=dev)
    out_idx = torch.empty((H * W, bins), dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        _raise_on(_lib.epcq_scatter(_ptr(point_idx_list), _ptr(point2img_idx_list), _ptr(depth_list), n, H, W,
                                    bins, _ptr(out_depth), _ptr(out_idx), _stream(dev)))
    return out_depth, out_idx


def build_hashtable(xyz, w2c, K, img_size, sort=False, workspace=None):
    """Fused NeuralPoints.pts2plane (epcq/epcq.py:279-295).

    xyz float32 [N,3]; w2c float32 [>=3, 4]; K float32 [3,3].  Returns
    (sorted_point_idx_list int32 [N], sorted_depth_list float32 [N],
    hash_table int32 [H*W, 2]) with pixel lists ordered by (pixel, index), or
    (pixel, depth, index) when `sort`.  `workspace` is an optional uint8 CUDA
    tensor of at least build_hashtable_workspace_size(N, H, W) bytes."""
    _check_input(xyz, "xyz_data")
    _check_input(w2c, "w2c")
    _check_input(K, "K")
    _check_dtype(xyz, torch.float32, "xyz_data")
    _check_dtype(w2c, torch.float32, "w2c")
    _check_dtype(K, torch.float32, "K")
    if xyz.dim() != 2 or xyz.size(1) != 3:
        raise RuntimeError("xyz_data must have shape [N, 3]")
    if w2c.dim() != 2 or w2c.size(0) < 3 or w2c.size(1) != 4:
        raise RuntimeError("w2c must have at least 3 rows of 4 values")
    if K.dim() != 2 or K.size(0) != 3 or K.size(1) != 3:
        raise RuntimeError("K must have shape [3, 3]")
    H, W = _img_size(img_size)
    n = xyz.size(0)
    dev = xyz.device
    _check_same_device(dev, ("w2c", w2c), ("K", K))
    need = build_hashtable_workspace_size(n, H, W)
    if workspace is None:
        workspace = torch.empty(need, dtype=torch.uint8, device=dev)
    elif workspace.numel() * workspace.element_size() < need:
        raise RuntimeError("build_hashtable: workspace too small")
    out_idx = torch.empty(n, dtype=torch.int32, device=dev)
    out_depth = torch.empty(n, dtype=torch.float32, device=dev)
    out_hash = torch.empty((H * W, 2), dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        _raise_on(_lib.epcq_build_hashtable(_ptr(xyz), n, _ptr(w2c), _ptr(K), H, W, int(bool(sort)),
                                            _ptr(out_idx), _ptr(out_depth), _ptr(out_hash), _ptr(workspace),
                                            workspace.numel() * workspace.element_size(), _stream(dev)))
    return out_idx, out_depth, out_hash


def build_hashtable_workspace_size(n, height, width):
    return int(_lib.epcq_build_hashtable_workspace_size(int(n), int(height), int(width)))


def _workspace(n, H, W, device, cull):
    if not cull:
        return None
    return torch.empty(int(_lib.epcq_sampling_workspace_size(n, H, W)), dtype=torch.uint8, device=device)


def _nbytes(t):
    return 0 if t is None else t.numel() * t.element_size()


def quick_sampling(neural_points, cam, options, sorted_point_idx_list, hashtable, mode="compat", cull=True,
                   rays=None):
    """quick_sampling (epcq/csrc/epcq_render_topk.cu:341-417).

    Returns (output_sp_xyz float32 [H,W,S,3], output_pc_idx int32 [H,W,S,Kc],
    output_sp_t float32 [H,W,S], mask_out bool [H,W,S]).  `mode` "compat"
    keeps the reference's neighbour buffer; "knn" returns the Kc nearest by
    (distance, index).  With `rays` (int32 pixel ids v * W + u, [R]) only
    those rays are sampled and the outputs are [R,S,3], [R,S,Kc], [R,S],
    [R,S]: row r equals the full image's row at rays[r] (an id outside the
    image gives an empty row)."""
    device = _check_spec_inputs(neural_points, cam, sorted_point_idx_list, hashtable)
    H, W = int(cam.height), int(cam.width)
    hashtable = _hashtable_view(hashtable, H, W)
    S = int(options.num_sample_points_per_ray)
    Kc = int(options.num_point_cloud_per_sp)
    xyz = neural_points.xyz_data
    if rays is not None:
        _check_input(rays, "rays")
        _check_dtype(rays, torch.int32, "rays")
        _check_same_device(device, ("rays", rays))
        lead = (rays.numel(),)
    else:
        lead = (H, W)
    sp_xyz = torch.empty(lead + (S, 3), dtype=torch.float32, device=device)
    sp_t = torch.empty(lead + (S,), dtype=torch.float32, device=device)
    pc_idx = torch.empty(lead + (S, Kc), dtype=torch.int32, device=device)
    mask = torch.empty(lead + (S,), dtype=torch.bool, device=device)
    ws = _workspace(xyz.size(0), H, W, device, cull)
    camera, opts = _camera_struct(cam), _options_struct(options)
    with torch.cuda.device(device):
        if rays is None:
            _raise_on(_lib.epcq_quick_sampling(_ptr(xyz), xyz.size(0), ctypes.byref(camera), ctypes.byref(opts),
                                               _ptr(sorted_point_idx_list), _ptr(hashtable), _MODES[mode],
                                               _ptr(sp_xyz), _ptr(pc_idx), _ptr(sp_t), _ptr(mask), _ptr(ws),
                                               _nbytes(ws), _stream(device)))
        else:
            _raise_on(_lib.epcq_quick_sampling_rays(_ptr(xyz), xyz.size(0), ctypes.byref(camera),
                                                    ctypes.byref(opts), _ptr(sorted_point_idx_list),
                                                    _ptr(hashtable), _MODES[mode], _ptr(rays), rays.numel(),
                                                    _ptr(sp_xyz), _ptr(pc_idx), _ptr(sp_t), _ptr(mask), _ptr(ws),
                                                    _nbytes(ws), _stream(device)))
    return sp_xyz, pc_idx, sp_t, mask


def quick_query_for_nearby_point(neural_points, cam, options, sorted_point_idx_list, hashtable, indices_of_ray):
    """quick_query_for_nearby_point (epcq/csrc/epcq_render_topk.cu:287-339).

    Returns int32 [R, num_point_per_ray]: the hits of each ray's neighbour
    buckets in lane order, -1 padded."""
    device = _check_spec_inputs(neural_points, cam, sorted_point_idx_list, hashtable)
    _check_input(indices_of_ray, "indices_of_ray")
    _check_dtype(indices_of_ray, torch.int32, "indices_of_ray")
    _check_same_device(device, ("indices_of_ray", indices_of_ray))
    H, W = int(cam.height), int(cam.width)
    hashtable = _hashtable_view(hashtable, H, W)
    R = indices_of_ray.numel()
    cap = int(options.num_point_per_ray)
    out = torch.empty((R, cap), dtype=torch.int32, device=device)
    xyz = neural_points.xyz_data
    camera, opts = _camera_struct(cam), _options_struct(options)
    with torch.cuda.device(device):
        _raise_on(_lib.epcq_quick_query(_ptr(xyz), xyz.size(0), ctypes.byref(camera), ctypes.byref(opts),
                                        _ptr(sorted_point_idx_list), _ptr(hashtable), _ptr(indices_of_ray), R,
                                        _ptr(out), _stream(device)))
    return out


HEAVY_DEFAULT = 128
LIST_WAVE_DEFAULT = 65536


def set_tuning(tile_w=32, tile_h=8, tile_cap=1024, flags=0, heavy=HEAVY_DEFAULT, list_wave=LIST_WAVE_DEFAULT):
    """Launch shape of the full-image samplers: tile_w x tile_h rays per
    block, tile_cap points staged in LDS per tile, TUNE_* flags; rays whose
    unculled window buckets hold more than `heavy` points are sampled by a
    whole wave (0: never); ray lists of at most `list_wave` rays are sampled
    one wave per ray (0: never)."""
    _raise_on(_lib.epcq_set_tuning(ctypes.byref(_Tuning(int(tile_w), int(tile_h), int(tile_cap), int(flags),
                                                        int(heavy), int(list_wave)))))


def get_tuning():
    t = _Tuning()
    _raise_on(_lib.epcq_get_tuning(ctypes.byref(t)))
    return {"tile_w": t.tile_w, "tile_h": t.tile_h, "tile_cap": t.tile_cap, "flags": t.flags, "heavy": t.heavy,
            "list_wave": t.list_wave}


def squared_threshold(r):
    """Smallest float s with sqrtf(s) >= r."""
    return float(_lib.epcq_squared_threshold(float(r)))


def max_live_samples(options, r_max):
    """Bound on the samples any ray can emit under these options when every
    point radius is below r_max."""
    return int(_lib.epcq_max_live_samples(ctypes.byref(_options_struct(options)), float(r_max)))


_RADIUS_MAX = {}


def radius_max(radius):
    """Largest per-point radius, read from the device once per radius tensor
    (keyed by storage, size and version, so an in-place change is seen)."""
    key = (radius.data_ptr(), radius.numel(), radius._version, str(radius.device))
    r = _RADIUS_MAX.get(key)
    if r is None:
        r = float(radius.max().item()) if radius.numel() else 0.0
        if len(_RADIUS_MAX) >= 64:
            _RADIUS_MAX.clear()
        _RADIUS_MAX[key] = r
    return r


def knn_sampling(neural_points, cam, options, sorted_point_idx_list, hashtable, rays=None, radius=None,
                 mode="knn", s_max=None, feat=None, cull=True, r_max=None):
    """Compact K-nearest sampling.

    Same samples as quick_sampling, stored per ray.  Returns a dict with
    n_samples int32 [R], sample_t float32 [R,s_max], nb_idx int32
    [R,s_max,Kc], and for mode "knn" nb_t / nb_dist float32 [R,s_max,Kc]
    (ray depth and point-to-ray distance of each neighbour); with `feat`
    (float32 or bfloat16 [N,F]) also w float32 [R,s_max,Kc] and agg float32
    [R,s_max,F] (inverse-distance weights and weighted feature sum).
    `rays` int32 pixel ids (default: all pixels); `radius` float32 [N]
    per-point radius (default: options.radius_threshold).  s_max defaults to
    max_live_samples(options, r_max); with per-point radii r_max (their
    maximum) is read once per radius tensor unless given.  The call does not
    synchronise with the device when s_max or r_max is given."""
    device = _check_spec_inputs(neural_points, cam, sorted_point_idx_list, hashtable)
    H, W = int(cam.height), int(cam.width)
    hashtable = _hashtable_view(hashtable, H, W)
    xyz = neural_points.xyz_data
    if rays is not None:
        _check_input(rays, "rays")
        _check_dtype(rays, torch.int32, "rays")
        _check_same_device(device, ("rays", rays))
        R = rays.numel()
    else:
        R = H * W
    if radius is not None:
        _check_input(radius, "radius")
        _check_dtype(radius, torch.float32, "radius")
        _check_same_device(device, ("radius", radius))
        if radius.numel() != xyz.size(0):
            raise RuntimeError("radius must have one entry per point")
    if s_max is None:
        if r_max is None:
            r_max = radius_max(radius) if radius is not None else float(options.radius_threshold)
        s_max = max(1, max_live_samples(options, r_max))
    Kc = int(options.num_point_cloud_per_sp)
    knn = _MODES[mode] == MODE_KNN
    out = {
        "n_samples": torch.empty(R, dtype=torch.int32, device=device),
        "sample_t": torch.empty((R, s_max), dtype=torch.float32, device=device),
        "nb_idx": torch.empty((R, s_max, Kc), dtype=torch.int32, device=device),
    }
    if knn:
        out["nb_t"] = torch.empty((R, s_max, Kc), dtype=torch.float32, device=device)
        out["nb_dist"] = torch.empty((R, s_max, Kc), dtype=torch.float32, device=device)
    F = 0
    bf16 = 0
    if feat is not None:
        _check_input(feat, "feat")
        _check_same_device(device, ("feat", feat))
        if feat.dtype not in (torch.float32, torch.bfloat16):
            raise RuntimeError("feat must be float32 or bfloat16")
        if feat.dim() != 2 or feat.size(0) != xyz.size(0):
            raise RuntimeError("feat must have shape [N, F]")
        F = feat.size(1)
        bf16 = int(feat.dtype == torch.bfloat16)
        out["w"] = torch.empty((R, s_max, Kc), dtype=torch.float32, device=device)
        out["agg"] = torch.empty((R, s_max, F), dtype=torch.float32, device=device)
    ws = _workspace(xyz.size(0), H, W, device, cull)
    camera, opts = _camera_struct(cam), _options_struct(options)
    with torch.cuda.device(device):
        _raise_on(_lib.epcq_knn_sampling(
            _ptr(xyz), xyz.size(0), ctypes.byref(camera), ctypes.byref(opts), _ptr(sorted_point_idx_list),
            _ptr(hashtable), _ptr(radius), _ptr(rays), R, _MODES[mode], int(s_max), _ptr(out["n_samples"]),
            _ptr(out["sample_t"]), _ptr(out["nb_idx"]), _ptr(out.get("nb_t")), _ptr(out.get("nb_dist")),
            _ptr(feat), bf16, F, _ptr(out.get("w")), _ptr(out.get("agg")), _ptr(ws), _nbytes(ws),
            _stream(device)))
    return out


def idw_features(xyz, feat, sp_xyz, nb_idx):
    """IDW weights and weighted features of dense samples
    (epcq/epcq.py:434-437,449-452).  sp_xyz [M,3], nb_idx int32 [M,K] ->
    (w float32 [M,K], agg float32 [M,F])."""
    for name, t in (("xyz_data", xyz), ("feat", feat), ("sp_xyz", sp_xyz), ("nb_idx", nb_idx)):
        _check_input(t, name)
    _check_dtype(xyz, torch.float32, "xyz_data")
    _check_dtype(sp_xyz, torch.float32, "sp_xyz")
    _check_dtype(nb_idx, torch.int32, "nb_idx")
    if feat.dtype not in (torch.float32, torch.bfloat16):
        raise RuntimeError("feat must be float32 or bfloat16")
    dev = xyz.device
    _check_same_device(dev, ("feat", feat), ("sp_xyz", sp_xyz), ("nb_idx", nb_idx))
    M = sp_xyz.numel() // 3
    K = nb_idx.size(-1) if nb_idx.dim() else 0
    F = feat.size(1)
    w = torch.empty((M, K), dtype=torch.float32, device=dev)
    agg = torch.empty((M, F), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        _raise_on(_lib.epcq_idw_features(_ptr(xyz), _ptr(feat), int(feat.dtype == torch.bfloat16), F,
                                         _ptr(sp_xyz), _ptr(nb_idx), M, K, _ptr(w), _ptr(agg), _stream(dev)))
    return w, agg


def _check_rows(name, x, R, S, last=None):
    _check_input(x, name)
    _check_dtype(x, torch.float32, name)
    shape = (R, S) if last is None else (R, S, last)
    if tuple(x.shape) != shape:
        raise RuntimeError(f"{name} must have shape {list(shape)}, found {list(x.shape)}")


def clamp_t(t_val, vsize=0.004):
    """clamp_t (epcq/epcq.py:508-515): per-sample intervals [R,S] of the
    sample depths t_val [R,S] ([R,0] when S = 1, as the reference's slicing
    gives)."""
    _check_input(t_val, "t_val")
    _check_dtype(t_val, torch.float32, "t_val")
    if t_val.dim() != 2:
        raise RuntimeError("t_val must have shape [R, S]")
    R, S = t_val.shape
    out = torch.empty((R, 0 if S == 1 else S), dtype=torch.float32, device=t_val.device)
    with torch.cuda.device(t_val.device):
        _raise_on(_lib.epcq_clamp_t(_ptr(t_val), R, S, float(vsize), _ptr(out), _stream(t_val.device)))
    return out


def render_rays(sigmas, rgbs, white_bg=False, deltas=None, t_val=None, vsize=0.004):
    """render (epcq/epcq.py:517-530) of sigmas [R,S] and rgbs [R,S,3] with
    intervals `deltas` [R,S], or clamp_t(t_val, vsize) applied on the fly
    when `t_val` is given instead.  Returns ray colours [R,3] (no autograd;
    see epcq.epcq.render for the differentiable form)."""
    if (deltas is None) == (t_val is None):
        raise RuntimeError("render_rays: give exactly one of deltas and t_val")
    _check_input(sigmas, "sigmas")
    if sigmas.dim() != 2:
        raise RuntimeError("sigmas must have shape [R, S]")
    R, S = sigmas.shape
    _check_rows("sigmas", sigmas, R, S)
    _check_rows("rgbs", rgbs, R, S, 3)
    src = deltas if deltas is not None else t_val
    _check_rows("deltas" if deltas is not None else "t_val", src, R, S)
    dev = sigmas.device
    _check_same_device(dev, ("rgbs", rgbs), ("deltas/t_val", src))
    out = torch.empty((R, 3), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        _raise_on(_lib.epcq_render_rays(_ptr(t_val), _ptr(deltas), float(vsize), _ptr(sigmas), _ptr(rgbs), R, S,
                                        int(bool(white_bg)), _ptr(out), _stream(dev)))
    return out


def render_rays_backward(sigmas, rgbs, grad_out, white_bg=False, deltas=None, t_val=None, vsize=0.004):
    """Gradients (d sigmas [R,S], d rgbs [R,S,3]) of render_rays for the
    output gradient grad_out [R,3]."""
    if (deltas is None) == (t_val is None):
        raise RuntimeError("render_rays_backward: give exactly one of deltas and t_val")
    R, S = sigmas.shape
    _check_rows("sigmas", sigmas, R, S)
    _check_rows("rgbs", rgbs, R, S, 3)
    src = deltas if deltas is not None else t_val
    _check_rows("deltas" if deltas is not None else "t_val", src, R, S)
    _check_input(grad_out, "grad_out")
    _check_dtype(grad_out, torch.float32, "grad_out")
    if tuple(grad_out.shape) != (R, 3):
        raise RuntimeError("grad_out must have shape [R, 3]")
    dev = sigmas.device
    _check_same_device(dev, ("rgbs", rgbs), ("deltas/t_val", src), ("grad_out", grad_out))
    gs = torch.empty((R, S), dtype=torch.float32, device=dev)
    gc = torch.empty((R, S, 3), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        _raise_on(_lib.epcq_render_rays_backward(_ptr(t_val), _ptr(deltas), float(vsize), _ptr(sigmas), _ptr(rgbs),
                                                 _ptr(grad_out), R, S, int(bool(white_bg)), _ptr(gs), _ptr(gc),
                                                 _stream(dev)))
    return gs, gc


# --------------------------------------------------------------- aggregator
# (in_features padded to a multiple of 16, stored as bf16) of each bf16 layer
_AGG_PAD = {"w1": 288, "w2": 256, "w3": 272, "w4": 256, "c1": 288, "c2": 128, "c3": 128}


def pack_aggregator(layers):
    """Device parameters of the Aggregator kernels from its Linear layers,
    given as {name: (weight [out, in], bias [out])} with the names of
    epcq_aggregator_params (w1..w4 block1/block3, wa alpha_branch, c1..c4
    color_branch).  Returns (struct, tensors): the tensors own the memory the
    struct points to."""
    keep = {}
    for name, (w, b) in layers.items():
        w = w.detach()
        b = b.detach()
        if name in _AGG_PAD:
            pw = torch.zeros((w.shape[0], _AGG_PAD[name]), dtype=torch.bfloat16, device=w.device)
            pw[:, :w.shape[1]] = w.to(torch.bfloat16)
        else:
            pw = w.to(torch.float32).reshape(-1).clone()
        # fresh allocations: the kernels read weights and biases 16 bytes at a time
        keep[name] = pw.contiguous()
        keep["b" + name[1:] if name[0] == "w" else "cb" + name[1:]] = b.to(torch.float32).reshape(-1).clone()
    st = _AggParams(**{n: ctypes.c_void_p(keep[n].data_ptr()) for n in _AGG_FIELDS})
    return st, keep


def aggregator_forward(packed, pc_xyz, pc_rgb, pc_dir, pc_feat, sp_xyz, valid, rays_d):
    """Aggregator.forward (epcq/epcq.py:439-507) on the compacted samples:
    pc_xyz / pc_rgb / pc_dir [M,K,3], pc_feat [M,K,32], sp_xyz [M,3],
    valid [M,K] bool, rays_d [M,3] -> (alpha [M], rgb [M,3]) float32."""
    st, _ = packed
    for t, n in ((pc_xyz, "pc_xyz"), (pc_rgb, "pc_rgb"), (pc_dir, "pc_dir"), (pc_feat, "pc_feat"),
                 (sp_xyz, "sp_xyz"), (rays_d, "rays_d")):
        _check_input(t, n)
        _check_dtype(t, torch.float32, n)
    _check_input(valid, "valid")
    _check_dtype(valid, torch.bool, "valid")
    M, K = valid.shape
    F = pc_feat.shape[-1]
    for t, n, shape in ((pc_xyz, "pc_xyz", (M, K, 3)), (pc_rgb, "pc_rgb", (M, K, 3)), (pc_dir, "pc_dir", (M, K, 3)),
                        (pc_feat, "pc_feat", (M, K, F)), (sp_xyz, "sp_xyz", (M, 3)), (rays_d, "rays_d", (M, 3))):
        if tuple(t.shape) != shape:
            raise RuntimeError(f"{n} must have shape {list(shape)}, got {list(t.shape)}")
    dev = valid.device
    _check_same_device(dev, ("pc_xyz", pc_xyz), ("pc_rgb", pc_rgb), ("pc_dir", pc_dir), ("pc_feat", pc_feat),
                       ("sp_xyz", sp_xyz), ("rays_d", rays_d))
    alpha = torch.empty(M, dtype=torch.float32, device=dev)
    rgb = torch.empty((M, 3), dtype=torch.float32, device=dev)
    ws_bytes = _lib.epcq_aggregator_workspace_size(M)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev) if ws_bytes else None
    with torch.cuda.device(dev):
        _raise_on(_lib.epcq_aggregator_forward(_ptr(pc_xyz), _ptr(pc_rgb), _ptr(pc_dir), _ptr(pc_feat), _ptr(sp_xyz),
                                               _ptr(valid), _ptr(rays_d), M, K, F, ctypes.byref(st), _ptr(alpha),
                                               _ptr(rgb), _ptr(ws), ws_bytes, _stream(dev)))
    return alpha, rgb


def debug_expf(x):
    """exp() as evaluated by the device kernels (test hook)."""
    _check_input(x, "x")
    _check_dtype(x, torch.float32, "x")
    out = torch.empty_like(x)
    with torch.cuda.device(x.device):
        _raise_on(_lib.epcq_debug_expf(_ptr(x), _ptr(out), x.numel(), _stream(x.device)))
    return out
