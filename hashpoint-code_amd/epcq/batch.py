"""Batched multi-view sampling: many views of uneven resolution (and, where
needed, several resident point clouds) in one persistent pass per stage,
replayable as a HIP graph.

    sampler = MultiViewSampler(RenderOptions(...), mode="compat")
    sampler.set_views([(xyz, cam), (xyz, cam2), ...])   # cam: any object with
                                                          # c2w, w2c, K, fx, fy, cx, cy, width, height
    sampler.capture()          # once
    sampler.replay()           # every batch (or run() without a graph)
    out = sampler.view_outputs(i)   # sorted_idx, hash, n_samples, sample_t, nb_idx[, nb_t, nb_dist]

Results per view equal epcq.csrc.knn_sampling on that view (same samples,
neighbours and compact layout).  The arena, a single uint8 CUDA tensor, holds
every intermediate and output; the views' outputs are tensor views into it,
valid until the next set_views().
"""
import ctypes

import numpy as np
import torch

from . import csrc as _C
from .csrc import HEAVY_DEFAULT, LIST_WAVE_DEFAULT, _BatchViewOut, _Tuning, _ViewDesc, _lib, _options_struct, _raise_on


def _host_floats(x, count):
    if hasattr(x, "detach"):
        x = x.detach().float().cpu().numpy()
    a = np.asarray(x, dtype=np.float32).reshape(-1)
    return [float(v) for v in a[:count]]


def view_desc(xyz, cam):
    """Descriptor of view `cam` over the device cloud `xyz` [n, 3] float32."""
    d = _ViewDesc()
    d.xyz = xyz.data_ptr()
    d.n = xyz.shape[0]
    c2w = np.asarray(_host_floats(cam.c2w, 16), np.float32)
    w2c = np.asarray(_host_floats(cam.w2c, 16), np.float32)
    d.c2w[:] = [float(v) for v in c2w[:12]]
    d.w2c[:] = [float(v) for v in w2c[:12]]
    d.K[:] = _host_floats(cam.K, 9)
    d.fx, d.fy, d.cx, d.cy = float(cam.fx), float(cam.fy), float(cam.cx), float(cam.cy)
    d.width, d.height = int(cam.width), int(cam.height)
    return d


class MultiViewSampler:
    """Owns an arena and a batch object of the HIP library."""

    def __init__(self, options, mode="compat", s_max=None, max_views=256, radius=None, tuning=None,
                 device="cuda:0", r_max=None):
        self.device = torch.device(device)
        self.options = options
        self.mode = mode
        self.opts = _options_struct(options)
        self.radius = radius
        if r_max is None:
            r_max = _C.radius_max(radius) if radius is not None else float(options.radius_threshold)
        r_max = float(r_max)
        self.s_max = int(s_max) if s_max is not None else max(1, _C.max_live_samples(options, r_max))
        self.r_max = r_max
        self.max_views = int(max_views)
        self.tuning = tuning
        self.arena = None
        self.handle = ctypes.c_void_p(0)
        self.stream = torch.cuda.Stream(self.device)
        self.views = []
        self.captured = False

    def _create(self, nbytes=None, arena=None):
        if self.handle:
            _lib.epcq_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p(0)
        if arena is None:
            arena = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        self.arena = arena
        tu = None
        if self.tuning is not None:
            tu = ctypes.byref(_Tuning(self.tuning["tile_w"], self.tuning["tile_h"], self.tuning["tile_cap"],
                                      self.tuning.get("flags", 0), self.tuning.get("heavy", HEAVY_DEFAULT),
                                      self.tuning.get("list_wave", LIST_WAVE_DEFAULT)))
        h = ctypes.c_void_p(0)
        rad = ctypes.c_void_p(self.radius.data_ptr()) if self.radius is not None else ctypes.c_void_p(0)
        _raise_on(_lib.epcq_batch_create(ctypes.byref(self.opts), _C._MODES[self.mode], self.s_max, self.max_views,
                                         rad, self.r_max, tu, ctypes.c_void_p(self.arena.data_ptr()),
                                         self.arena.numel(), ctypes.byref(h)))
        self.handle = h
        self.captured = False

    def required_bytes(self, pairs):
        descs = (_ViewDesc * len(pairs))(*[view_desc(x, c) for x, c in pairs])
        return int(_lib.epcq_batch_arena_bytes(descs, len(pairs), self.max_views, ctypes.byref(self.opts),
                                               _C._MODES[self.mode], self.s_max))

    def header_bytes(self):
        """Arena bytes of the descriptor header (for max_views views)."""
        return int(_lib.epcq_batch_header_bytes(self.max_views))

    def view_bytes(self, xyz, cam):
        """Upper bound of the arena bytes one view adds to a batch."""
        d = view_desc(xyz, cam)
        return int(_lib.epcq_batch_view_bytes(ctypes.byref(d), ctypes.byref(self.opts), _C._MODES[self.mode],
                                              self.s_max))

    def reserve(self, nbytes):
        """Allocates an arena of at least nbytes (drops a captured graph)."""
        if self.arena is None or self.arena.numel() < nbytes:
            self._create(nbytes)

    def adopt_arena(self, arena):
        """Uses the caller's contiguous uint8 CUDA tensor as the arena.  Several
        samplers may share one arena if only one of them is used at a time
        (each set_views() rewrites it)."""
        if not (arena.is_cuda and arena.dtype == torch.uint8 and arena.is_contiguous()):
            raise RuntimeError("MultiViewSampler: the arena must be a contiguous uint8 CUDA tensor")
        self._create(arena=arena)

    def set_views(self, pairs):
        """pairs: list of (xyz [n,3] float32 CUDA tensor, camera)."""
        if len(pairs) > self.max_views:
            raise RuntimeError(f"MultiViewSampler: {len(pairs)} views exceed max_views={self.max_views}")
        for xyz, _ in pairs:
            if not (xyz.is_cuda and xyz.dtype == torch.float32 and xyz.is_contiguous() and xyz.dim() == 2
                    and xyz.shape[1] == 3):
                raise RuntimeError("MultiViewSampler: xyz must be a contiguous float32 CUDA tensor [n, 3]")
        need = self.required_bytes(pairs)
        self.reserve(need)
        descs = (_ViewDesc * len(pairs))(*[view_desc(x, c) for x, c in pairs])
        self.views = pairs
        _raise_on(_lib.epcq_batch_set_views(self.handle, descs, len(pairs), ctypes.c_void_p(self.stream.cuda_stream)))

    @property
    def rays(self):
        return int(_lib.epcq_batch_rays(self.handle))

    def run(self):
        """Enqueues the whole pass on the sampler's stream (no graph)."""
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        _raise_on(_lib.epcq_batch_run(self.handle, ctypes.c_void_p(self.stream.cuda_stream)))
        torch.cuda.current_stream(self.device).wait_stream(self.stream)

    def capture(self):
        """Captures the pass into a HIP graph (replayed by replay())."""
        torch.cuda.synchronize(self.device)
        _raise_on(_lib.epcq_batch_capture(self.handle, ctypes.c_void_p(self.stream.cuda_stream)))
        self.captured = True

    def replay(self):
        if not self.captured:
            self.capture()
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        _raise_on(_lib.epcq_batch_launch(self.handle, ctypes.c_void_p(self.stream.cuda_stream)))
        torch.cuda.current_stream(self.device).wait_stream(self.stream)

    def profile(self):
        """Runs the pass eagerly, timing its stages on the sampler's stream:
        {"bin": ms (projection and binning), "place": ms (per-pixel placement,
        hash table and bucket cones), "build": bin + place, "sample": ms}."""
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        ms = (ctypes.c_float * 3)()
        _raise_on(_lib.epcq_batch_profile(self.handle, ctypes.c_void_p(self.stream.cuda_stream), ms))
        return {"bin": ms[0], "place": ms[1], "build": ms[0] + ms[1], "sample": ms[2]}

    def _tensor(self, ptr, count, dtype, shape):
        if not ptr:
            return None
        off = ptr - self.arena.data_ptr()
        esize = torch.empty((), dtype=dtype).element_size()
        return self.arena[off:off + count * esize].view(dtype).view(shape)

    def view_outputs(self, i):
        o = _BatchViewOut()
        _raise_on(_lib.epcq_batch_view_outputs(self.handle, int(i), ctypes.byref(o)))
        xyz, cam = self.views[i]
        n = xyz.shape[0]
        R = int(o.rays)
        K = int(self.options.num_point_cloud_per_sp)
        S = self.s_max
        out = {
            "sorted_idx": self._tensor(o.sorted_idx, n, torch.int32, (n,)),
            "hash": self._tensor(o.hash, 2 * R, torch.int32, (R, 2)),
            "n_samples": self._tensor(o.n_samples, R, torch.int32, (R,)),
            "sample_t": self._tensor(o.sample_t, R * S, torch.float32, (R, S)),
            "nb_idx": self._tensor(o.nb_idx, R * S * K, torch.int32, (R, S, K)),
        }
        if o.nb_t:
            out["nb_t"] = self._tensor(o.nb_t, R * S * K, torch.float32, (R, S, K))
            out["nb_dist"] = self._tensor(o.nb_dist, R * S * K, torch.float32, (R, S, K))
        return out

    def batch_outputs(self):
        """The compact outputs of all views as whole-batch tensors, and each
        caller view's (ray_offset, rays) rows in them."""
        rows = []
        base = None
        for i in range(len(self.views)):
            o = _BatchViewOut()
            _raise_on(_lib.epcq_batch_view_outputs(self.handle, int(i), ctypes.byref(o)))
            rows.append((int(o.ray_offset), int(o.rays)))
            if base is None:
                base = o
        T = self.rays
        K = int(self.options.num_point_cloud_per_sp)
        S = self.s_max
        off = base.ray_offset
        out = {
            "n_samples": self._tensor(base.n_samples - off * 4, T, torch.int32, (T,)),
            "sample_t": self._tensor(base.sample_t - off * S * 4, T * S, torch.float32, (T, S)),
            "nb_idx": self._tensor(base.nb_idx - off * S * K * 4, T * S * K, torch.int32, (T, S, K)),
        }
        if base.nb_t:
            out["nb_t"] = self._tensor(base.nb_t - off * S * K * 4, T * S * K, torch.float32, (T, S, K))
            out["nb_dist"] = self._tensor(base.nb_dist - off * S * K * 4, T * S * K, torch.float32, (T, S, K))
        return out, rows

    def close(self):
        if self.handle:
            _lib.epcq_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
