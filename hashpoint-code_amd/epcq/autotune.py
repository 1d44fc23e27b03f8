"""Autotuner for the samplers' launch shape.

The full-image samplers take a tile shape (tile_w x tile_h rays per block),
a staging capacity (points of a tile's halo region copied to LDS), a flag
(XCD-contiguous tile order) and the outlier threshold of the wave-wide form
(`heavy`); the best choice
depends on the point density per pixel and the image sizes.  tune() times a
caller-supplied workload under candidate settings with HIP events and keeps
the fastest, searching one knob at a time (shape, then capacity, then flags, then heavy).
Results are cached per (device, key) in memory and, when EPCQ_TUNE_CACHE
names a file, in that JSON file.
"""
import json
import os

import torch

from . import csrc as _C

SHAPES = [(16, 16), (8, 8), (16, 8), (32, 8), (32, 4), (64, 2)]
# staged points per tile; on the 64-view 1M-point batch at 16x16 tiles: 1024
# 14.35 ms per step, 1280 14.28, 1536 14.22, 1792 16.7 (fewer blocks per CU)
CAPS = [1024, 512, 1280, 1536, 2048, 3000]
FLAGS = [0, _C.TUNE_XCD_ORDER]
HEAVY = [_C.HEAVY_DEFAULT, 64, 256, 0]
DEFAULT = {"tile_w": 32, "tile_h": 8, "tile_cap": 1024, "flags": 0, "heavy": _C.HEAVY_DEFAULT}

_cache = {}


def _cache_file():
    return os.environ.get("EPCQ_TUNE_CACHE")


def _load():
    path = _cache_file()
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                _cache.update(json.load(f))
        except (OSError, ValueError):
            pass


def _store():
    path = _cache_file()
    if path:
        try:
            with open(path, "w") as f:
                json.dump(_cache, f, indent=1, sort_keys=True)
        except OSError:
            pass


def time_ms(fn, reps=3, warmup=1, stream=None):
    """Average milliseconds of fn() on `stream` (default: current stream)."""
    stream = stream or torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(key, run, reps=3, stream=None, shapes=SHAPES, caps=CAPS, flags=FLAGS, heavy=HEAVY, log=None):
    """Best launch shape for the workload run(tuning) (which applies the
    tuning dict and enqueues the work once).  Returns (tuning, {str: ms})."""
    if not _cache:
        _load()
    full_key = f"{torch.cuda.get_device_name()}|{key}"
    if full_key in _cache:
        return dict(_cache[full_key]["best"]), _cache[full_key]["times"]
    times = {}

    def measure(t):
        name = f'{t["tile_w"]}x{t["tile_h"]}/cap{t["tile_cap"]}/f{t["flags"]}/h{t["heavy"]}'
        if name not in times:
            times[name] = time_ms(lambda: run(t), reps=reps, stream=stream)
            if log:
                log(f"autotune {key}: {name} {times[name]:.4f} ms")
        return times[name]

    best = dict(DEFAULT)
    for tw, th in shapes:
        t = dict(best, tile_w=tw, tile_h=th)
        if measure(t) < measure(best):
            best = t
    for cap in caps:
        t = dict(best, tile_cap=cap)
        if measure(t) < measure(best):
            best = t
    for fl in flags:
        t = dict(best, flags=fl)
        if measure(t) < measure(best):
            best = t
    for hv in heavy:
        t = dict(best, heavy=hv)
        if measure(t) < measure(best):
            best = t
    _cache[full_key] = {"best": best, "times": times}
    _store()
    return dict(best), times


def apply(tuning):
    """Makes `tuning` the process-wide launch shape of the per-view samplers."""
    _C.set_tuning(tuning["tile_w"], tuning["tile_h"], tuning["tile_cap"], tuning.get("flags", 0),
                  tuning.get("heavy", _C.HEAVY_DEFAULT), tuning.get("list_wave", _C.LIST_WAVE_DEFAULT))
