"""Host-side pieces of the HIP library that run without a GPU: the library
loads, exports every symbol its header declares, and its host helpers."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "epcq_abi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(epcq_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import epcq.csrc as C
    names = declared_functions()
    assert len(names) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", C.LIBRARY_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in names if n not in exported]
    assert not missing, f"not exported: {missing}"
    # the ctypes table covers exactly the declared functions
    assert sorted(C.ABI_SYMBOLS) == names
    for n in names:
        assert getattr(C._lib, n) is not None
    assert C._lib.epcq_abi_version() == 3


def test_squared_threshold_is_exact():
    import epcq.csrc as C
    for r in [0.005, 0.01, 0.016, 1e-3, 0.0123456, 1.0, 3.7, 1e-20, 1e20]:
        s = np.float32(C.squared_threshold(r))
        r32 = np.float32(r)
        assert np.sqrt(s) >= r32
        below = np.nextafter(s, np.float32(0))
        assert np.sqrt(below) < r32
        # all floats around the threshold agree with the sqrt test
        xs = np.array([s], np.float32)
        for _ in range(64):
            xs = np.append(xs, [np.nextafter(xs[-1], np.float32(np.inf))])
        ys = np.array([s], np.float32)
        for _ in range(64):
            ys = np.append(ys, [np.nextafter(ys[-1], np.float32(0))])
        for e2 in np.concatenate([xs, ys]):
            assert (e2 < s) == (np.sqrt(e2) < r32)
    assert C.squared_threshold(0.0) == 0.0 and C.squared_threshold(-1.0) == 0.0
    assert C.squared_threshold(float("nan")) == 0.0  # no distance is below NaN
    assert C.squared_threshold(float("inf")) == float("inf")
    assert C.squared_threshold(1e30) == float("inf")  # every finite distance is below it
    # radii whose square underflows: only the distance 0 is below them
    for r in [1e-23, 1e-40]:
        s = np.float32(C.squared_threshold(r))
        assert s == np.nextafter(np.float32(0), np.float32(1)) and np.sqrt(s) >= np.float32(r)


def test_max_live_samples_bound():
    import epcq.csrc as C
    from epcq.epcq import RenderOptions
    assert C.max_live_samples(RenderOptions(radius_threshold=0.005), 0.005) == 2
    assert C.max_live_samples(RenderOptions(radius_threshold=0.01), 0.01) == 2
    assert C.max_live_samples(RenderOptions(), 0.016) == 2
    assert C.max_live_samples(RenderOptions(num_sample_points_per_ray=1), 0.01) == 1
    # small alpha: many samples possible, capped by S
    opt = RenderOptions(sdf2weight_gaussian_gamma=0.2, sdf2weight_gaussian_alpha=0.01)
    assert C.max_live_samples(opt, 0.016) == 16
    # brute force the walk in float32 with the smallest possible alpha
    for g, ap, r in [(0.9, 0.2, 0.005), (0.5, 0.05, 0.02), (0.3, 0.1, 0.05)]:
        opt = RenderOptions(sdf2weight_gaussian_gamma=g, sdf2weight_gaussian_alpha=ap)
        amin = np.float32(g) * np.float32(np.exp(-np.float32(r) / np.float32(ap)))
        T, n = np.float32(1.0), 0
        while T >= np.float32(1e-3) and n < 16:
            T = np.float32(T - amin)
            n += 1
        assert C.max_live_samples(opt, r) >= n


def test_tuning_roundtrip_and_validation():
    import epcq.csrc as C
    old = C.get_tuning()
    try:
        C.set_tuning(32, 4, 1024, C.TUNE_XCD_ORDER, heavy=7)
        assert C.get_tuning() == {"tile_w": 32, "tile_h": 4, "tile_cap": 1024, "flags": C.TUNE_XCD_ORDER,
                                  "heavy": 7, "list_wave": C.LIST_WAVE_DEFAULT}
        C.set_tuning(32, 4, 1024, list_wave=0)
        assert C.get_tuning()["list_wave"] == 0
        with pytest.raises(RuntimeError, match="unknown flags"):
            C.set_tuning(16, 16, 1024, flags=2)
        with pytest.raises(RuntimeError, match="heavy"):
            C.set_tuning(16, 16, 1024, heavy=-1)
        with pytest.raises(RuntimeError, match="list_wave"):
            C.set_tuning(16, 16, 1024, list_wave=-1)
        with pytest.raises(RuntimeError, match="tile_w"):
            C.set_tuning(12, 8, 1024)
        with pytest.raises(RuntimeError, match="tile_w"):
            C.set_tuning(16, 3, 1024)
    finally:
        C.set_tuning(**old)
